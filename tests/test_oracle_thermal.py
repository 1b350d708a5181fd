"""Pins the oracle's colour-temperature fit (CPU only).

  * bb_color_mag: zero colour at the reference wavelength, equal to
    -2.5 log10 of a ratio of Planck functions in frequency evaluated
    independently (numpy expm1), continuous across the Wien branch;
  * recovery: colours generated from a known log T0 / cooling rate are fitted
    back (the swarm's optimum costs no more than the truth);
  * the reference's own tests (tests/test_thermal.rs:5-120) on its fixture
    generator (tests/synthetic.rs, restated): a result, log T in [3, 6],
    finite non-negative chi2, reference band g, at least one other band,
    at least 10 colours, None for no bands, no colours for a single band;
  * fit_ref_band_gp's selection against numpy (Cholesky / solves).
"""
import math

import numpy as np
import pytest

HC_K = 1.4388e8


def planck_color(lref, lam, temp):
    """-2.5 log10(B_nu(lref) / B_nu(lam)), B_nu ~ nu^3 / expm1(h nu / k T)."""
    b = lambda l: (1.0 / l) ** 3 / np.expm1(HC_K / (l * temp))
    return -2.5 * np.log10(b(lref) / b(lam))


def test_bb_color_mag(orc):
    for temp in (3000.0, 6000.0, 1e4, 3e4):
        assert orc.bb_color_mag(4770.0, 4770.0, temp) == pytest.approx(0.0, abs=1e-12)
        for lam in (6231.0, 7625.0, 3671.0):
            assert orc.bb_color_mag(4770.0, lam, temp) == pytest.approx(planck_color(4770.0, lam, temp), rel=1e-9)
    # Wien branch (exponent > 500) joins the Planck form
    t_w = HC_K / (4770.0 * 501.0)
    assert orc.bb_color_mag(4770.0, 6231.0, t_w) == pytest.approx(planck_color(4770.0, 6231.0, t_w), rel=1e-9)


def pack_one(pkg, bands):
    return pkg.thermal.pack_sources([bands])


def test_recovers_known_temperature(orc, pkg):
    rng = np.random.default_rng(3)
    lt0, rate = 4.1, -0.004
    t = np.sort(rng.uniform(0, 80, 40))
    ref = pkg.nonparametric.BandData(np.linspace(0, 80, 30), 18 + 0.01 * np.linspace(0, 80, 30), np.full(30, 0.02))
    bands = {"g": ref}
    for name, lam in (("r", 6231.0), ("i", 7625.0)):
        temp = 10 ** (lt0 + rate * t)
        col = planck_color(4770.0, lam, temp)
        # colour = ref - band: band = ref_gp(t) - colour, the ref GP ~ its linear trend
        bands[name] = pkg.nonparametric.BandData(t, 18 + 0.01 * t - col + 0.005 * rng.normal(size=t.size),
                                                 np.full(t.size, 0.01))
    pack, meta = pack_one(pkg, bands)
    r = orc.batch_thermal(pack, seed=7)
    assert r["status"][0] == 2 and meta[0][0] == "g"
    o = r["out"][0]
    assert o[7] == 80
    assert abs(o[0] - lt0) < 0.1 and abs(o[2] - rate) < 0.003
    assert 2.0 <= o[1] <= 7.0 and o[3] >= 0 and o[4] >= 0


@pytest.fixture(scope="module")
def ref_sources(pkg):
    out = {}
    for seed in (700, 800, 900, 1000, 1100, 1200):
        out[seed] = pkg.common.build_mag_bands(*pkg.synthetic.reference_test_source(30, seed))
    return out


def test_reference_thermal_tests(orc, pkg, ref_sources):
    srcs = [ref_sources[s] for s in (700, 800, 900, 1000, 1100, 1200)]
    pack, meta = pkg.thermal.pack_sources(srcs)
    r = orc.batch_thermal(pack)
    res = pkg.thermal.results_from_arrays(r, pack, meta)
    assert all(x is not None for x in res)                                   # thermal_returns_result
    assert res[1].log_temp_peak is None or 3.0 <= res[1].log_temp_peak <= 6.0  # thermal_log_temp_in_range
    assert res[2].chi2 is None or (math.isfinite(res[2].chi2) and res[2].chi2 >= 0)
    assert res[3].ref_band in ("g", "r")
    assert res[4].n_bands_used >= 1
    assert res[5].n_color_obs >= 10
    pack0, meta0 = pkg.thermal.pack_sources([{}])
    assert meta0 == [None]                                                   # thermal_empty_bands
    single = pkg.common.build_mag_bands([0.0, 1.0, 2.0, 3.0, 4.0], [20.0, 19.5, 19.0, 19.5, 20.0], [0.1] * 5,
                                        ["r"] * 5)
    p1, m1 = pkg.thermal.pack_sources([single])
    r1 = orc.batch_thermal(p1)
    res1 = pkg.thermal.results_from_arrays(r1, p1, m1)[0]
    assert res1 is not None and res1.n_color_obs == 0                        # thermal_single_band


def test_ref_gp_selection_against_numpy(orc, pkg, ref_sources):
    bands = ref_sources[900]
    pack, meta = pkg.thermal.pack_sources([bands])
    r = orc.batch_thermal(pack)
    st = r["gp_state"][0]
    cap = (st.size - 4) // 2
    m = int(st[2 * cap + 3])
    g = bands["g"]
    t, y, e = (np.asarray(x) for x in (g.times, g.values, g.errors))
    assert m == 30 and np.array_equal(st[cap:cap + m], t)  # 30 points: no subsampling
    dur = pack["duration"][0]
    best = (np.inf, None)
    for amp in (0.1, 0.3):
        for f in (6.0, 12.0, 24.0):
            ls = max(dur / f, 0.1)
            k = amp * np.exp(-np.subtract.outer(t, t) ** 2 * (0.5 / ls ** 2))
            L = np.linalg.cholesky(k + np.diag(np.maximum(np.maximum(e ** 2, 1e-6), 1e-10)))
            al = np.linalg.solve(L.T, np.linalg.solve(L, y - y.mean()))
            rms = np.sqrt(np.mean((y - (k @ al + y.mean())) ** 2))
            if abs((k @ al + y.mean()).min() - y.min()) > 6:
                continue
            if rms < best[0]:
                best = (rms, (amp, 0.5 / ls ** 2, al))
    amp, it, al = best[1]
    assert st[2 * cap] == amp and st[2 * cap + 1] == pytest.approx(it, rel=1e-15)
    assert np.allclose(st[:m], al, rtol=1e-7, atol=1e-9)
