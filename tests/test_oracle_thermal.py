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


def _bb_color_mag(lref, lam, temp):
    """bb_color_mag (src/thermal.rs:75-92) in plain Python."""
    x_ref, x = 1.4388e8 / (lref * temp), 1.4388e8 / (lam * temp)
    if x > 500.0 or x_ref > 500.0:
        return -2.5 * (3.0 * math.log10(lam / lref) + (x - x_ref) * 0.434294481903251827651128918916605082)
    q = lam / lref
    ratio = q * q * q * (math.exp(x) - 1.0) / (math.exp(x_ref) - 1.0)
    if ratio <= 0.0 or not math.isfinite(ratio):
        return math.nan
    return -2.5 * math.log10(ratio)


def _replay_swarm(orc, obs, lref, src, run, seed):
    """One run of argmin 0.11's ParticleSwarm as fit_thermal configures it
    (src/thermal.rs:336-339; 10 particles x 30 iterations, box [3, 6] x
    [-0.05, 0.01]), written from the algorithm: inertia 1 / (2 ln 2),
    cognitive = social = 1/2 + ln 2; positions uniform in the box and
    velocities uniform in [-range, range]; each iteration first moves every
    particle (momentum + cognitive pull x U + social pull x U per dimension,
    position clamped to the box), then evaluates all of them and updates the
    personal and global bests on strict improvement in particle order; the
    initial global best is the first particle of least cost.  Uniforms from
    the oracle's Philox stream (counter (0, 64 run + particle,
    iteration or 0xFFFFFFFF, 0x54480000), the same for every source)."""
    u24 = lambda x: float(int(x) >> 8) / 16777216.0
    key = [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF]
    lo, hi = (3.0, -0.05), (6.0, 0.01)
    w, cc = 1.0 / (2.0 * math.log(2.0)), 0.5 + math.log(2.0)

    def cost(lt0, rate):
        chi2 = 0.0
        for t, c, e, lam in obs:
            mc = _bb_color_mag(lref, lam, math.pow(10.0, lt0 + rate * t))
            if not math.isfinite(mc):
                return 1e10
            chi2 += (c - mc) ** 2 / (e * e + 1e-10)
        return chi2 / max(len(obs), 1)

    x, v, pb, pc = [], [], [], []
    for p in range(10):
        r = orc.philox([0, 64 * run + p, 0xFFFFFFFF, 0x54480000], key)
        xi = [lo[d] + u24(r[d]) * (hi[d] - lo[d]) for d in range(2)]
        vi = [-(hi[d] - lo[d]) + u24(r[2 + d]) * (2.0 * (hi[d] - lo[d])) for d in range(2)]
        x.append(xi), v.append(vi), pb.append(list(xi)), pc.append(cost(*xi))
    g = min(range(10), key=lambda p: (pc[p], p))
    gp, gc = list(x[g]), pc[g]
    for it in range(30):
        for p in range(10):
            r = orc.philox([0, 64 * run + p, it, 0x54480000], key)
            for d in range(2):
                v[p][d] = v[p][d] * w + (pb[p][d] - x[p][d]) * u24(r[d]) * cc + (gp[d] - x[p][d]) * u24(r[2 + d]) * cc
                x[p][d] = min(max(x[p][d] + v[p][d], lo[d]), hi[d])
        c = [cost(*x[p]) for p in range(10)]
        for p in range(10):
            if c[p] < pc[p]:
                pc[p], pb[p] = c[p], list(x[p])
                if c[p] < gc:
                    gc, gp = c[p], list(x[p])
    return gp, gc


def test_python_swarm_reproduces_oracle_thermal(orc):
    """The oracle's two swarm runs replayed in plain Python on colours from a
    constant reference GP (state row with one zero weight: mean = ym): best
    positions, best cost and the run-to-run spread agree exactly."""
    rng = np.random.default_rng(31)
    ns, cap = 3, 32
    lref, lams = 4770.0, (6231.0, 7625.0)
    col_t, col_m, col_e, col_l, col_off = [], [], [], [], [0]
    st = np.zeros((ns, 2 * cap + 4))
    for s in range(ns):
        st[s, 2 * cap:2 * cap + 4] = [1.0, 1.0, 18.0, 1.0]  # amp, 1/(2l^2), mean, m = 1 (alpha = 0)
        n = 8 + 4 * s
        t = np.sort(rng.uniform(0.0, 40.0, n))
        lam = np.array([lams[i % 2] for i in range(n)])
        temp = 10 ** (4.0 - 0.01 * t)
        colour = np.array([_bb_color_mag(lref, l_, T) for l_, T in zip(lam, temp)])
        col_t.append(t), col_m.append(18.0 - colour + rng.normal(0, 0.03, n)), col_e.append(np.full(n, 0.05))
        col_l.append(lam), col_off.append(col_off[-1] + n)
    pack = {"ref_off": np.zeros(ns + 1, np.int64), "ref_t": np.zeros(1), "ref_m": np.zeros(1), "ref_e": np.zeros(1),
            "col_off": np.array(col_off), "col_t": np.concatenate(col_t), "col_m": np.concatenate(col_m),
            "col_e": np.concatenate(col_e), "col_lambda": np.concatenate(col_l), "lambda_ref": np.full(ns, lref),
            "duration": np.full(ns, 40.0), "gp_state": st}
    seed = 0x5EED
    r = orc.batch_thermal(pack, seed=seed)
    for s in range(ns):
        assert r["status"][s] == 2
        a, b = col_off[s], col_off[s + 1]
        e = np.sqrt(pack["col_e"][a:b] ** 2 + 0.02 ** 2)
        obs = list(zip(pack["col_t"][a:b], 18.0 - pack["col_m"][a:b], e, pack["col_lambda"][a:b]))
        runs = [_replay_swarm(orc, obs, lref, s, run, seed) for run in (0, 1)]
        bi = 1 if runs[1][1] < runs[0][1] else 0
        out = r["out"][s]
        assert out[0] == runs[bi][0][0] and out[2] == runs[bi][0][1] and out[6] == runs[bi][1]
        m0 = (runs[0][0][0] + runs[1][0][0]) / 2.0
        sd0 = math.sqrt(((runs[0][0][0] - m0) ** 2 + (runs[1][0][0] - m0) ** 2) / 2.0)
        assert out[3] == pytest.approx(sd0, rel=1e-12, abs=1e-15)
