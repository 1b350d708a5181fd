"""Pins the oracle's fit_gp_2d (src/gp2d.rs:336-439) against an independent
numpy implementation (linalg.solve, not the oracle's Cholesky-Crout): the
subsample floor(i n / 200), the per-source grid from the subsample's time and
log-wavelength ranges (floored at 1 d / 0.01 dex), first strict minimum of the
train RMS, predictive mean and std (variance floored at 1e-10)."""
import numpy as np
import pytest


def make(rng, sizes):
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    t, w, y = [], [], []
    for n in sizes:
        tt = np.sort(rng.uniform(0.0, 80.0, n))
        ww = np.log10(rng.choice([4770.0, 6231.0, 7625.0], n))
        t.append(tt)
        w.append(ww)
        y.append(18.0 + np.sin(tt / 12.0) + (ww - 3.8) + 0.05 * rng.normal(size=n))
    return {"offsets": off, "times": np.concatenate(t), "waves": np.concatenate(w), "mags": np.concatenate(y),
            "noise_var": np.full(int(off[-1]), 0.05 ** 2)}


def numpy_fit(t, w, y, nv, qt, qw):
    n = t.size
    if n > 200:
        idx = [int(i * (n / 200)) for i in range(200)]
        t, w, y, nv = t[idx], w[idx], y[idx], nv[idx]
    dur = max(t.max() - t.min(), 1.0)
    wr = max(w.max() - w.min(), 0.01)
    best = (np.inf, None)
    for a in (0.3, 1.0, 3.0, 10.0):
        for lt in (dur / 4, dur / 8, dur / 16):
            for lw in (wr * 2, wr, wr / 2):
                k = lambda t1, w1, t2, w2: a * np.exp(-(t1[:, None] - t2[None]) ** 2 / (2 * lt * lt)
                                                      - (w1[:, None] - w2[None]) ** 2 / (2 * lw * lw))
                K = k(t, w, t, w) + np.diag(np.maximum(nv, 1e-10))
                al = np.linalg.solve(K, y - y.mean())
                rms = np.sqrt(np.mean((k(t, w, t, w) @ al + y.mean() - y) ** 2))
                if rms < best[0] - 1e-12:
                    ks = k(qt, qw, t, w)
                    var = a - np.einsum("ij,ji->i", ks, np.linalg.solve(K, ks.T))
                    best = (rms, (a, lt, lw), ks @ al + y.mean(), np.sqrt(np.maximum(var, 1e-10)), t.size)
    return best


def test_fit_gp_2d_against_numpy(orc):
    rng = np.random.default_rng(7)
    sizes = [6, 40, 120, 450]
    c = make(rng, sizes)
    qt = np.tile(np.linspace(0, 80, 9), (len(sizes), 1))
    qw = np.tile(np.log10([4770.0, 6231.0, 7625.0]), (len(sizes), 3))
    o = orc.batch_fit_gp_2d(c, qt, qw)
    off = c["offsets"]
    for s in range(len(sizes)):
        sl = slice(off[s], off[s + 1])
        rms, hyp, mu, sd, m = numpy_fit(c["times"][sl], c["waves"][sl], c["mags"][sl], c["noise_var"][sl],
                                        qt[s], qw[s])
        assert o["n_train"][s] == m
        assert np.allclose(o["hyp"][s], hyp, rtol=1e-14)
        assert o["train_rms"][s] == pytest.approx(rms, rel=1e-8)
        assert np.allclose(o["pred"][s], mu, rtol=1e-9, atol=1e-9)
        assert np.allclose(o["std"][s], sd, rtol=1e-6, atol=1e-9)
