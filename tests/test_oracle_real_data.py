"""The oracle on the reference's real-data fixture (tests/golden/
real_sources.json, 30 ZTF sources), CPU only: the acceptance ranges of the
reference's own real-data tests (tests/test_real_data.rs:79-150, 203-250)
hold for the oracle's process_band and fit_thermal, so the parity tests that
compare the device against it on these sources compare against a pipeline
that meets the reference's bar."""
import json
from pathlib import Path

import numpy as np
import pytest

from test_gpu_real_data import np_csr

FIXTURE = Path(__file__).resolve().parent / "golden" / "real_sources.json"


@pytest.fixture(scope="module")
def mag_bands(pkg):
    out = []
    for s in json.loads(FIXTURE.read_text()):
        ts, ms, es, bs = [], [], [], []
        for name, b in s["bands"].items():
            ts += b["times"]
            ms += b["values"]
            es += b["errors"]
            bs += [name[3:]] * len(b["times"])
        out.append(pkg.common.build_mag_bands(ts, ms, es, bs))
    return out


def test_fixture_shape(mag_bands):
    assert len(mag_bands) == 30
    sizes = [len(b.times) for bands in mag_bands for b in bands.values()]
    assert min(sizes) == 3 and max(sizes) == 1005 and len(sizes) == 83
    assert all(min(min(b.times) for b in bands.values()) == 0.0 for bands in mag_bands)


def test_oracle_nonparametric_acceptance(orc, mag_bands):
    csr, lo, hi = np_csr(mag_bands)
    o = orc.batch_nonparametric(csr, lo, hi)
    names = list(orc.NP_FEATURES)
    f = o["features"]
    col = lambda k: f[:, names.index(k)]
    n_obs = np.diff(csr["offsets"])
    ok = (o["status"] > 0) & (n_obs >= 10)
    assert ok.sum() >= 60
    pm = col("peak_mag")[ok]
    pm = pm[np.isfinite(pm)]
    assert pm.size > 0 and np.all((pm > 10.0) & (pm < 30.0))
    for k, lo_ in (("chi2", 0.0), ("rise_time", 0.0), ("decay_efold", 0.0)):
        v = col(k)[ok]
        v = v[np.isfinite(v)]
        assert np.all(v >= lo_), k
    vn = col("von_neumann_ratio")[ok]
    vn = vn[np.isfinite(vn)]
    assert np.all((vn > 0.0) & (vn < 100.0))


def test_oracle_thermal_acceptance(pkg, orc, mag_bands):
    pack, meta = pkg.thermal.pack_sources(mag_bands)
    o = orc.batch_thermal(pack, seed=42)
    fit = o["status"] == 2
    assert fit.sum() >= 20
    lt = o["out"][fit, 0]
    assert np.all((10 ** lt > 500.0) & (10 ** lt < 1_000_001.0))
    assert np.all(o["out"][fit, 5] >= 0.0)
