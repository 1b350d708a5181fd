"""Host logic of the nonparametric interface (CPU only): the pooled band of
fit_combined_band (src/nonparametric.rs:1164-1203) and the feature-row to
NonparametricBandResult mapping (src/nonparametric.rs:74-160, 1462-1522)."""
import math

import numpy as np


def test_combined_band(pkg):
    npm = pkg.nonparametric
    bands = {"g": npm.BandData([3.0, 1.0], [18.0, 18.4], [0.1, 0.1]),
             "r": npm.BandData([2.0, 0.5, 4.0], [17.0, float("nan"), 17.6], [0.1, 0.2, 0.1]),
             "i": npm.BandData([], [], [])}
    assert npm._combined_band(bands) is None  # 4 finite points
    bands["i"] = npm.BandData([2.0], [16.0], [0.3])
    t, v, e, gm = npm._combined_band(bands)
    # per-band means include the NaN (it propagates), so the r band's pooled values are NaN
    assert list(t) == [1.0, 2.0, 2.0, 3.0, 4.0]
    assert v[0] == 18.4 - 18.2 and v[3] == 18.0 - 18.2 and math.isnan(v[1]) and v[2] == 0.0
    assert gm == (18.0 + 18.4 + 17.0 + 17.6 + 16.0) / 5


def test_result_mapping(pkg):
    npm = pkg.nonparametric
    row = np.full(npm.NP_NFEAT if hasattr(npm, "NP_NFEAT") else 50, np.nan)
    names = npm.FEATURE_NAMES
    row[names.index("t0")] = 12.5
    row[names.index("n_obs")] = 40
    row[names.index("n_upper_limits")] = 3
    row[names.index("n_local_maxima")] = 2
    r = npm._result("g", row)
    assert r.band == "g" and r.t0 == 12.5 and r.n_obs == 40 and r.n_upper_limits == 3
    assert r.n_local_maxima == 2.0 and r.peak_mag is None and r.fwhm is None
    assert len(names) == 50 and names[0] == "rise_time" and names[-1] == "cusum_range"


def test_state_gp(pkg):
    npm = pkg.nonparametric
    cap = 25
    st = np.zeros(2 * cap + 4)
    assert npm._state_gp(st) is None
    st[:3] = [0.5, -0.2, 0.1]
    st[cap:cap + 3] = [0.0, 10.0, 20.0]
    st[2 * cap:2 * cap + 4] = [3.0, 0.5 / 100.0, 18.0, 3]
    gp = npm._state_gp(st)
    assert gp.kernel_lengthscale() == 10.0 and gp.kernel_amp() == 3.0
    q = np.array([5.0])
    want = 18.0 + sum(a * 3.0 * math.exp(-(5.0 - x) ** 2 / 200.0) for a, x in zip([0.5, -0.2, 0.1], [0, 10, 20]))
    assert gp.predict(q)[0] == want or abs(gp.predict(q)[0] - want) < 1e-12
