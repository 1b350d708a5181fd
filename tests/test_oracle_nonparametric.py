"""Pins the oracle's nonparametric band pipeline (CPU only).

  * light-curve metrics: the reference's known answers (tests/test_common.rs:
    229-234 rise NaN at peak 0, :260-290 rate signs and NaN below two points)
    and closed forms on piecewise-linear magnitude curves (crossing times,
    dm15, FWHM on grid points, rise time scale, least-squares slopes);
  * variability features: closed forms (alternating sequences for von
    Neumann / Stetson K / excess variance / CUSUM, three-point pre-peak
    statistics), time-order invariance of the von Neumann ratio;
  * grid features: exact power law recovered with zero chi2, two-point
    distributions (skewness 0, kurtosis -2), counted dips and inflections;
  * FITC: predictive mean / std / approx NLML against a dense numpy
    reconstruction of the FITC equations (Quinonero-Candela & Rasmussen 2005),
    and against the exact dense GP when the inducing points are the inputs;
  * the band pipeline: detection rule err <= 1.0857/3, the minimum length
    scale and train-RMS selection against numpy (sorted gaps, numpy.linalg
    Cholesky / solves), grid mean / std against numpy, chi2 and peak rules.
"""
import math

import numpy as np
import pytest

FAC = 1.0857362047581294
HALF = 0.7525749891599529


def test_reference_known_answers(orc):
    t = np.arange(10.0)
    m = 20.0 - 0.3 * t
    out = orc.np_curve_metrics(t, np.r_[18.0, m[1:]], 0)
    assert math.isnan(out[4])  # rise time scale at peak index 0
    out = orc.np_curve_metrics(t, m, 9)  # brightening
    assert out[7] < 0
    out = orc.np_curve_metrics(t, m[::-1], 0)  # fading
    assert out[8] > 0
    one = orc.np_curve_metrics(t[:1], m[:1], 0)
    assert math.isnan(one[7]) and math.isnan(one[8])


def test_linear_rise_and_decay_closed_forms(orc):
    t = np.arange(0.0, 101.0)
    tp, r, s = 40.0, 0.25, 0.1
    m = np.where(t <= tp, 18.0 + r * (tp - t), 18.0 + s * (t - tp))
    o = orc.np_curve_metrics(t, m, 40)
    assert o[0] == pytest.approx(FAC / s, rel=1e-12)          # decay e-fold
    assert o[1] == pytest.approx(15 * s, rel=1e-12)           # dm15
    assert o[2] == pytest.approx(HALF / s, rel=1e-12)         # decay half-max
    assert o[3] == pytest.approx(8.0 + 3.0)                   # FWHM on grid points: ceil(.75/s) + ceil(.75/r)
    assert o[5] == pytest.approx(HALF / r, rel=1e-12)         # rise half-max
    assert o[6] == pytest.approx(FAC / r, rel=1e-12)          # rise e-fold
    assert o[9] == pytest.approx(-r, rel=1e-10)               # near-peak rise rate (30 d window)
    assert o[10] == pytest.approx(s, rel=1e-10)               # near-peak decay rate
    line = 20.0 - 0.05 * t
    o2 = orc.np_curve_metrics(t, line, 100)
    assert o2[7] == pytest.approx(-0.05, rel=1e-10) and o2[8] == pytest.approx(-0.05, rel=1e-10)


def test_rise_timescale_closed_form(orc):
    t = np.arange(0.0, 31.0)
    m = np.where(t < 10, 20.0, np.where(t <= 20, 20.0 - 0.2 * (t - 10), 18.0 + 0.05 * (t - 20)))
    o = orc.np_curve_metrics(t, m, 20)
    # target 20 + (18 - 20)(1 - 1/e) = 18.7358 is crossed at t = 16.32: nearest grid point t = 16
    assert o[4] == pytest.approx(4.0)


def test_variability_closed_forms(orc):
    n = 12
    alt = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    t = np.arange(n, dtype=float)
    e = np.full(n, 0.1)
    out = orc.np_raw_features(t, alt, e, 100.0, 0.0)
    assert out[0] == pytest.approx(4 * (n - 1) / n, rel=1e-12)        # von Neumann
    assert out[3] == pytest.approx(1.0, rel=1e-12)                    # Stetson K
    perm = np.random.default_rng(0).permutation(n)
    assert orc.np_raw_features(t[perm], alt[perm], e, 100.0, 0.0)[0] == pytest.approx(out[0], rel=1e-12)
    mags = 20.0 + 0.3 * alt
    ev = orc.np_raw_features(t, mags, e, 100.0, 0.0)[4]
    assert ev == pytest.approx((0.09 - 0.01) / 400.0, rel=1e-9)      # excess variance
    sd = math.sqrt(n * 0.09 / (n - 1))
    assert orc.np_raw_features(t, mags, e, 100.0, 0.0)[5] == pytest.approx(0.3 / (n * sd), rel=1e-9)  # CUSUM
    tt = np.array([0.0, 1.0, 2.0, 5.0, 6.0])
    mm = np.array([20.0, 21.0, 22.0, 18.0, 18.5])
    ee = np.array([0.1, 0.3, 0.2, 0.05, 0.05])
    out = orc.np_raw_features(tt, mm, ee, 4.0, 18.0)
    assert out[1] == pytest.approx(1.0, rel=1e-12)                    # pre-peak rms (sample std)
    assert out[2] == pytest.approx((21.0 - 18.0) / 0.2, rel=1e-12)    # median pre-peak error 0.2
    assert np.all(np.isnan(orc.np_raw_features(tt[:4], mm[:4], ee[:4], 4.0, 18.0)[[0, 3, 4, 5]]))


def test_grid_feature_closed_forms(orc):
    tg = np.linspace(0.0, 100.0, 101)
    pk = 10
    dt = tg - tg[pk]
    p = np.where(dt >= 0, 18.0 + 2.5 * np.log10(np.maximum(dt, 0) + 1.0), 18.0 + 0.1 * (-dt))
    sd = np.full(tg.size, 0.2)
    o = orc.np_grid_features(tg, p, sd, pk)
    assert o[10] == pytest.approx(2.5, rel=1e-10) and abs(o[11]) < 1e-20   # power law
    assert o[12] == pytest.approx(1.0)                                      # strictly fading after peak
    assert o[5] == pytest.approx(0.04, rel=1e-12) and o[6] == pytest.approx(0.04, rel=1e-12)
    two = np.where(np.arange(100) % 2 == 0, 1.0, -1.0)
    o2 = orc.np_grid_features(np.arange(100.0), two, np.ones(100), 1)
    assert abs(o2[7]) < 1e-12 and o2[8] == pytest.approx(-2.0, rel=1e-12)
    assert o2[0] == pytest.approx(1.0) and o2[1] == pytest.approx(2.0)
    assert o2[2] == pytest.approx(2.0) and o2[3] == pytest.approx(-2.0)
    dips = np.full(50, 20.0)
    dips[[10, 30]] = 19.5
    dips[0] = 19.0
    assert orc.np_grid_features(np.arange(50.0), dips, np.ones(50), 0)[13] == 3
    xs = np.linspace(0.0, 4 * np.pi, 401) + np.pi / 400  # no grid point on a zero of sin
    o3 = orc.np_grid_features(xs, np.sin(xs) + 20.0, np.ones(xs.size), 0)
    assert o3[9] == 3  # curvature changes sign at pi, 2 pi, 3 pi
    xz = np.linspace(0.0, 4 * np.pi, 401)  # second differences through an exact zero are not counted
    assert orc.np_grid_features(xz, np.sin(xz) + 20.0, np.ones(xz.size), 0)[9] < 3


def fitc_numpy(t, y, nv, amp, ls, m, q):
    """Dense numpy FITC: Lambda = diag(max(a - Q_jj, 0) + nv), B = Kmm + Kmn
    Lambda^-1 Knm, mean = K*m B^-1 Kmn Lambda^-1 (y - ybar) + ybar, var = a -
    K*m Kmm^-1 Km* + K*m B^-1 Km*."""
    it = 0.5 / ls ** 2
    k = lambda a, b: amp * np.exp(-np.subtract.outer(a, b) ** 2 * it)
    z = t.min() + np.arange(m) * (t.max() - t.min()) / max(m - 1, 1)
    kmm = k(z, z) + 1e-6 * np.eye(m)
    kmn = k(z, t)
    qjj = np.einsum("ij,ij->j", kmn, np.linalg.solve(kmm, kmn))
    lam = np.maximum(amp - qjj, 0) + nv
    b = kmm + (kmn / lam) @ kmn.T
    yb = y.mean()
    w = kmn @ ((y - yb) / lam)
    alpha = np.linalg.solve(b, w)
    ks = k(q, z)
    mean = ks @ alpha + yb
    var = amp - np.einsum("ij,ij->i", ks, np.linalg.solve(kmm, ks.T).T) + np.einsum(
        "ij,ij->i", ks, np.linalg.solve(b, ks.T).T)
    pt = k(t, z) @ alpha + yb
    return mean, np.sqrt(np.maximum(var, 1e-10)), np.mean((y - pt) ** 2 / nv + np.log(nv))


@pytest.mark.parametrize("n,m,amp,ls", [(60, 30, 1.0, 12.0), (150, 30, 3.0, 6.0), (45, 10, 0.3, 25.0)])
def test_fitc_against_numpy(orc, n, m, amp, ls):
    rng = np.random.default_rng(n)
    t = np.sort(rng.uniform(0, 150, n))
    y = 19 - 1.5 * np.exp(-((t - 60) / 20) ** 2) + 0.05 * rng.normal(size=n)
    nv = rng.uniform(1e-3, 1e-2, n)
    q = np.linspace(-5, 155, 37)
    mean, std, nl = orc.sparse_gp(t, y, nv, amp, ls, m, q)
    wm, ws, wn = fitc_numpy(t, y, nv, amp, ls, m, q)
    assert np.allclose(mean, wm, rtol=1e-7, atol=1e-7)
    assert np.allclose(std, ws, rtol=1e-5, atol=1e-7)
    assert nl == pytest.approx(wn, rel=1e-7)


def test_fitc_with_inducing_points_at_inputs_is_exact_gp(orc):
    t = np.linspace(0.0, 100.0, 25)
    y = 18 + np.sin(t / 15.0)
    nv = np.full(t.size, 1e-2)
    q = np.linspace(0, 100, 41)
    mean, std, _ = orc.sparse_gp(t, y, nv, 1.0, 20.0, 25, q)
    dm, ds, _, _, _ = orc.dense_gp(t, y, nv, 1.0, 20.0, q)
    assert np.allclose(mean, dm, atol=1e-4) and np.allclose(std, ds, atol=1e-3)


def numpy_band(t, m, e, t_lo, t_hi, n_grid, amps, lsf):
    """Independent numpy restatement of the dense branch of process_band:
    selection, grid mean / std, chi2, peak."""
    flux = 10 ** (-0.4 * (m - 23.9))
    det = (flux > 0) & (e / FAC * flux > 0) & (flux / (e / FAC * flux) >= 3)
    td, md, ed = t[det], m[det], e[det]
    gaps = np.diff(np.sort(td))
    gaps = np.sort(gaps[np.isfinite(gaps) & (gaps > 0)])
    min_ls = max(2 * (gaps[gaps.size // 2] if gaps.size else 1.0), 0.1)
    nv = np.maximum(ed ** 2, 1e-6)
    dur = t_hi - t_lo
    best = (np.inf, None)
    for a in amps:
        for f in lsf:
            ls = max(dur / f, 0.1)
            if ls < min_ls:
                continue
            k = a * np.exp(-np.subtract.outer(td, td) ** 2 * (0.5 / ls ** 2))
            try:
                L = np.linalg.cholesky(k + np.diag(np.maximum(nv, 1e-10)))
            except np.linalg.LinAlgError:
                continue
            yb = md.mean()
            alpha = np.linalg.solve(L.T, np.linalg.solve(L, md - yb))
            rms = np.sqrt(np.mean((k @ alpha + yb - md) ** 2))
            if np.isfinite(rms) and rms < best[0]:
                best = (rms, (a, ls, L, alpha, yb))
    a, ls, L, alpha, yb = best[1]
    tg = t_lo + np.arange(n_grid) * dur / (n_grid - 1)
    ks = a * np.exp(-np.subtract.outer(tg, td) ** 2 * (0.5 / ls ** 2))
    pred = ks @ alpha + yb
    v = np.linalg.solve(L, ks.T)
    std = np.sqrt(np.maximum(a - np.sum(v * v, axis=0), 1e-10))
    ko = a * np.exp(-np.subtract.outer(td, td) ** 2 * (0.5 / ls ** 2))
    res = md - (ko @ alpha + yb)
    chi2 = np.mean(res ** 2 / (ed ** 2 + 1e-10))
    return dict(n_upper=int((~det).sum()), amp=a, ls=ls, pred=pred, std_raw=std, chi2=chi2,
                rms=np.sqrt(np.mean(res ** 2)), td=td)


def test_band_pipeline_dense_against_numpy(orc, pkg):
    c = pkg.synthetic.bazin_mag_batch(12, seed=5, n_min=10, n_max=90)
    r = orc.batch_nonparametric(c, c["t_lo"], c["t_hi"], n_grid=60, dense_max=100)
    off = c["offsets"]
    checked = 0
    for b in range(off.size - 1):
        a0, a1 = off[b], off[b + 1]
        t, m, e = c["times"][a0:a1], c["mags"][a0:a1], c["mag_errs"][a0:a1]
        f = r["features"][b]
        assert f[10] == a1 - a0
        nd = int(np.sum(e <= FAC / 3.0))
        if a1 - a0 < 5:
            assert r["status"][b] == 0
            continue
        assert f[44] == (a1 - a0) - nd
        if nd < 5:
            assert r["status"][b] == 0
            continue
        w = numpy_band(t, m, e, c["t_lo"][b], c["t_hi"][b], 60, orc.NP_AMPS, orc.NP_LS_FACTORS)
        assert r["status"][b] == 1
        assert f[45] == w["amp"] and f[46] == pytest.approx(w["ls"], rel=1e-15)
        assert np.allclose(r["pred"][b], w["pred"], rtol=1e-8, atol=1e-8)
        assert f[8] == pytest.approx(w["chi2"], rel=1e-6)
        pk = int(np.argmin(r["pred"][b]))
        tg = c["t_lo"][b] + np.arange(60) * (c["t_hi"][b] - c["t_lo"][b]) / 59
        assert f[6] == tg[pk] and f[7] == max(r["pred"][b][pk], m[e <= FAC / 3.0].min() - 2.0)
        # the std rescale factor is 0.6 clamp(rms / mean std at the detections, 0.05, 5)
        ratio = r["std"][b] / w["std_raw"]
        assert np.allclose(ratio, ratio[0], rtol=1e-6) and 0.03 - 1e-12 <= ratio[0] <= 3.0 + 1e-12
        checked += 1
    assert checked >= 20


def test_band_pipeline_fitc_branch(orc, pkg):
    c = pkg.synthetic.bazin_mag_batch(3, seed=9, n_min=120, n_max=200)
    r = orc.batch_nonparametric(c, c["t_lo"], c["t_hi"], n_grid=50, dense_max=100)
    off = c["offsets"]
    for b in range(off.size - 1):
        e = c["mag_errs"][off[b]:off[b + 1]]
        nd = int(np.sum(e <= FAC / 3.0))
        assert r["status"][b] == (2 if nd > 100 else (1 if nd >= 5 else 0))
        if r["status"][b] == 2:
            st = r["state"][b]
            cap = (st.size - 4) // 2
            assert st[2 * cap + 3] == 25  # thermal dense GP on the 25-point subsample
            assert np.all(np.isfinite(r["pred"][b])) and np.all(r["std"][b] > 0)
