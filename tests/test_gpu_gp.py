"""GPU parity of the batched dense GP (FP64) against the oracle (needs an
MI355X): selected hyperparameters, alpha, predictive mean / derivative / std
on the grid and the mean at every observation, to 1e-8 relative, for
training sets up to the LDS cap (128), both subsample rules and both
selection criteria; plus the reference API (GpBandInput / batch_gp_fit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bands(seed, sizes):
    rng = np.random.default_rng(seed)
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    t, y, nv = [], [], []
    for n in sizes:
        tt = np.sort(rng.uniform(0, 150, n))
        t.append(tt)
        y.append(20 - 2.0 * np.exp(-((tt - rng.uniform(30, 90)) / rng.uniform(8, 30)) ** 2)
                 + 0.05 * rng.normal(size=n))
        nv.append(rng.uniform(1e-4, 1e-2, n))
    cat = lambda a: np.concatenate(a) if a else np.zeros(0)
    return {"offsets": off, "times": cat(t), "mags": cat(y), "noise_var": cat(nv)}


AMPS = np.array([0.3, 1.0, 3.0, 10.0])
LSS = 150.0 / np.array([3.0, 6.0, 12.0, 24.0, 48.0])


def close(a, b, rel=1e-8):
    fin = np.isfinite(b)
    return np.array_equal(fin, np.isfinite(a)) and np.allclose(a[fin], b[fin], rtol=rel, atol=rel * np.abs(b[fin]).max(initial=1.0))


@pytest.mark.parametrize("cap,rule,crit", [(25, 0, 0), (64, 1, 1), (128, 0, 0), (128, 1, 1)])
def test_gp_matches_oracle(ctx, orc, cap, rule, crit):
    sizes = [0, 2, 3, 5, 17, 25, 26, 63, 64, 65, 100, 127, 128, 129, 200]
    csr = bands(cap + rule, sizes)
    q = np.linspace(0, 150, 200)
    g = ctx.gp_fit(csr, q, AMPS, LSS, cap=cap, subsample_rule=rule, criterion=crit)
    o = orc.batch_gp_fit(csr, q, AMPS, LSS, cap, rule, crit)
    assert np.array_equal(g["state"][:, 2 * cap:2 * cap + 2], o["state"][:, 2 * cap:2 * cap + 2])
    assert np.array_equal(g["state"][:, 2 * cap + 3], o["state"][:, 2 * cap + 3])
    assert close(g["state"][:, :cap], o["state"][:, :cap], 1e-7)
    assert close(g["pred"], o["pred"]) and close(g["dpred"], o["dpred"], 1e-7)
    assert close(g["std"], o["std"], 1e-6)
    assert close(g["pred_obs"], o["pred_obs"])


def test_gp_per_band_grid(ctx, orc):
    csr = bands(7, [40, 90, 12])
    q = np.stack([np.linspace(0, 150, 64), np.linspace(10, 120, 64), np.linspace(-5, 160, 64)])
    g = ctx.gp_fit(csr, q, AMPS, LSS, cap=128, query_per_band=True)
    for b in range(3):
        sub = {"offsets": np.array([0, csr["offsets"][b + 1] - csr["offsets"][b]]),
               **{k: csr[k][csr["offsets"][b]:csr["offsets"][b + 1]] for k in ("times", "mags", "noise_var")}}
        o = orc.batch_gp_fit(sub, q[b], AMPS, LSS, 128)
        assert close(g["pred"][b:b + 1], o["pred"]) and close(g["std"][b:b + 1], o["std"], 1e-6)


def test_reference_api(ctx, pkg, orc):
    csr = bands(11, [30, 8, 2, 60])
    off = csr["offsets"]
    inputs = [pkg.GpBandInput(csr["times"][off[b]:off[b + 1]], csr["mags"][off[b]:off[b + 1]],
                              csr["noise_var"][off[b]:off[b + 1]]) for b in range(4)]
    q = np.linspace(0, 150, 50)
    out = ctx.batch_gp_fit(inputs, q, AMPS, LSS, 25)
    o = orc.batch_gp_fit(csr, q, AMPS, LSS, 25)
    for b in range(4):
        assert close(np.asarray(out[b].pred_grid)[None], o["pred"][b:b + 1])
        assert close(np.asarray(out[b].pred_at_obs), o["pred_obs"][off[b]:off[b + 1]])
    assert out[2].dense_gp is None and np.all(np.isnan(out[2].pred_grid))
    gp = out[0].dense_gp
    assert np.allclose(gp.predict(q), out[0].pred_grid, rtol=1e-10)
    assert gp.kernel_amp() in AMPS
