"""Pins the oracle's dense GP (CPU only).

  * fit / predictive mean / std / NLML against numpy.linalg (an independent
    Cholesky and solve) on the same kernel matrix (Rasmussen & Williams
    eqs. 2.25-2.26, 2.30);
  * the analytic derivative of the mean against central differences;
  * the reference's in-module acceptance: a DenseGP recovers sin(t/2)
    (src/sparse_gp.rs:334-360, tolerance 0.15);
  * batch selection: the chosen candidate is the first strict minimum of the
    score over the amp-major grid, skipped candidates (ls < 0.1), bands with
    fewer than 3 points, both subsample rules.
"""
import math

import numpy as np
import pytest


def np_gp(t, y, nv, amp, ls, q):
    K = amp * np.exp(-(t[:, None] - t[None, :]) ** 2 / (2 * ls * ls)) + np.diag(np.maximum(nv, 1e-10))
    L = np.linalg.cholesky(K)
    ym = y.mean()
    alpha = np.linalg.solve(K, y - ym)
    ks = amp * np.exp(-(q[:, None] - t[None, :]) ** 2 / (2 * ls * ls))
    mean = ks @ alpha + ym
    v = np.linalg.solve(L, ks.T)
    var = np.maximum(amp - (v * v).sum(axis=0), 1e-10)
    nlml = 0.5 * (y - ym) @ alpha + np.log(np.diag(L)).sum() + 0.5 * t.size * math.log(2 * math.pi)
    return mean, np.sqrt(var), nlml


@pytest.mark.parametrize("seed", range(4))
def test_dense_gp_against_numpy(orc, seed):
    rng = np.random.default_rng(seed)
    n = 10 + 7 * seed
    t = np.sort(rng.uniform(0, 60, n))
    y = 18 + np.sin(t / 7) + 0.05 * rng.normal(size=n)
    nv = rng.uniform(1e-4, 4e-3, n)
    q = np.linspace(-5, 65, 33)
    amp, ls = [0.3, 1.0, 3.0, 10.0][seed], [4.0, 8.0, 15.0, 2.5][seed]
    mean, std, dm, nlml, tr = orc.dense_gp(t, y, nv, amp, ls, q)
    m2, s2, nl2 = np_gp(t, y, nv, amp, ls, q)
    assert np.allclose(mean, m2, rtol=1e-9, atol=1e-9)
    assert np.allclose(std, s2, rtol=1e-6, atol=1e-9)
    assert nlml == pytest.approx(nl2, rel=1e-9, abs=1e-9)
    h = 1e-5
    fd = (orc.dense_gp(t, y, nv, amp, ls, q + h)[0] - orc.dense_gp(t, y, nv, amp, ls, q - h)[0]) / (2 * h)
    assert np.allclose(dm, fd, rtol=1e-5, atol=1e-7)
    mt = orc.dense_gp(t, y, nv, amp, ls, t)[0]
    assert tr == pytest.approx(math.sqrt(np.mean((mt - y) ** 2)), rel=1e-12)


def test_reference_sin_recovery(orc):
    t = np.linspace(0, 20, 40)
    y = np.sin(t / 2)
    mean = orc.dense_gp(t, y, np.full(40, 0.01), 1.0, 3.0, np.linspace(1, 19, 30))[0]
    assert np.all(np.abs(mean - np.sin(np.linspace(1, 19, 30) / 2)) < 0.15)


def make_bands(rng, sizes):
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    t, y, nv = [], [], []
    for n in sizes:
        tt = np.sort(rng.uniform(0, 100, n))
        t.append(tt)
        y.append(19 + 1.5 * np.exp(-((tt - 40) / 15) ** 2) * -1 + 0.05 * rng.normal(size=n))
        nv.append(rng.uniform(1e-4, 1e-2, n))
    cat = lambda a: np.concatenate(a) if a else np.zeros(0)
    return {"offsets": off, "times": cat(t), "mags": cat(y), "noise_var": cat(nv)}


@pytest.mark.parametrize("criterion", [0, 1])
@pytest.mark.parametrize("rule", [0, 1])
def test_batch_selection_semantics(orc, criterion, rule):
    rng = np.random.default_rng(3)
    csr = make_bands(rng, [2, 5, 30, 80])
    q = np.linspace(0, 100, 50)
    amps = np.array([0.3, 1.0, 3.0, 10.0])
    lss = np.array([0.05, 5.0, 10.0, 20.0])
    cap = 25
    out = orc.batch_gp_fit(csr, q, amps, lss, cap, rule, criterion)
    off = csr["offsets"]
    assert np.all(np.isnan(out["pred"][0])) and out["state"][0, 2 * cap + 3] == 0
    for b in range(1, 4):
        a, e = off[b], off[b + 1]
        n = e - a
        m = min(n, cap)
        if n <= m:
            idx = np.arange(n)
        elif rule == 0:
            idx = np.minimum(((np.arange(m) * (n - 1) / (m - 1)) + 0.5).astype(int), n - 1)
        else:
            idx = np.minimum(((np.arange(m) + 0.5) * n / m).astype(int), n - 1)
        tt, yy, vv = csr["times"][a:e][idx], csr["mags"][a:e][idx], csr["noise_var"][a:e][idx]
        scores = []
        for A in amps:
            for ls in lss:
                if ls < 0.1:
                    scores.append(1e99)
                    continue
                r = orc.dense_gp(tt, yy, vv, A, ls, q)
                scores.append(1e99 if r is None else (r[3] if criterion == 0 else r[4]))
        c = int(np.argmin(scores))  # first minimum
        A, ls = amps[c // lss.size], lss[c % lss.size]
        st = out["state"][b]
        assert st[2 * cap] == A and st[2 * cap + 1] == pytest.approx(0.5 / ls ** 2) and st[2 * cap + 3] == m
        assert np.array_equal(st[cap:cap + m], tt)
        want = orc.dense_gp(tt, yy, vv, A, ls, q)
        assert np.allclose(out["pred"][b], want[0], rtol=1e-12) and np.allclose(out["std"][b], want[1], rtol=1e-12)
        po = orc.dense_gp(tt, yy, vv, A, ls, csr["times"][a:e])[0]
        assert np.allclose(out["pred_obs"][a:e], po, rtol=1e-12)
