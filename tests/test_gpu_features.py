"""The feature extractor end to end on the device (needs an MI355X), on six
of the reference's real ZTF sources (tests/golden/real_sources.json):

  * extract_features_batch runs the nonparametric pipeline, thermal on its
    GPs, the parametric fit over all models with SVI and the 2D GP with
    thermal, and returns per source exactly the reference's feature keys
    (tests/golden/feature_keys.json);
  * the nonparametric and thermal features equal the fields of those
    fitters' own results for the reference band (the extractor adds no
    arithmetic of its own there), and every value is finite or None;
  * a single source through extract_features equals its batch entry, and
    the per-source JSON record of the real results reads back with the
    reference's top-level schema.
"""
import json
import math
from pathlib import Path

import pytest

from test_gpu_real_data import real  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "feature_keys.json").read_text())
PICK = [0, 3, 7, 12, 18, 25]


def test_feature_extractor_on_real_sources(pkg, real, tmp_path):
    from lightcurve_fitting_amd import features, io, nonparametric, thermal

    src, mag_all, flux_all = real
    mags = [mag_all[k] for k in PICK]
    fluxes = [flux_all[k] for k in PICK]
    gpu = pkg.GpuContext(0)
    feats = features.extract_features_batch(gpu, mags, fluxes, "r")
    assert len(feats) == len(PICK)
    npo = nonparametric.fit_nonparametric_batch(gpu, mags)
    th = thermal.fit_thermal_batch(gpu, mags, [g for _, g in npo])
    for k, f in enumerate(feats):
        assert sorted(f) == GOLD["keys"], k
        assert all(v is None or math.isfinite(v) for v in f.values()), k
        bands = npo[k][0]
        ref = next((b for b in bands if b.band == "r"), bands[0] if bands else None)
        if ref is not None:
            for name in features._NP_FIELDS:
                v = getattr(ref, name)
                want = v if v is not None and math.isfinite(v) else None
                assert f[f"np_{name}"] == want, (k, name)
        t = th[k]
        assert f["thermal_log_temp_peak"] == (t.log_temp_peak if t is not None else None), k
    one = features.extract_features(gpu, mags[1], fluxes[1], "r")
    assert one == feats[1]
    # the JSON record of the first source's real results
    par = pkg.parametric.fit_parametric_batch(gpu, fluxes[:1], True, pkg.parametric.UncertaintyMethod.Svi)
    g2 = pkg.gp2d.fit_gp_2d_with_thermal_batch(gpu, mags[:1], 50)
    r2, t2 = g2[0] if g2[0] is not None else (None, None)
    p = io.write_source_json(tmp_path, src[PICK[0]].get("name", "s0"), npo[0][0], par[0], th[0], r2, t2, feats[0])
    rec = json.loads(p.read_text())
    assert list(rec) == ["source_id", "nonparametric", "parametric", "thermal", "gp2d", "gp2d_thermal", "features"]
    assert sorted(rec["features"]) == GOLD["keys"]
    assert [b["band"] for b in rec["nonparametric"]] == [b.band for b in npo[0][0]]
    gpu.close()
