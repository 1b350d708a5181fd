"""GPU parity of the batched dense GP (FP64) against the oracle (needs an
MI355X): selected hyperparameters, alpha, predictive mean / derivative / std
on the grid and the mean at every observation, to 1e-8 relative, for
training sets up to the LDS cap (200), both subsample rules and both
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


@pytest.mark.parametrize("cap,rule,crit", [(25, 0, 0), (64, 1, 1), (128, 0, 0), (128, 1, 1), (200, 0, 0), (200, 1, 1)])
def test_gp_matches_oracle(ctx, orc, cap, rule, crit):
    sizes = [0, 2, 3, 5, 17, 25, 26, 63, 64, 65, 100, 127, 128, 129, 199, 200, 201, 333]
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


def sources_2d(seed, sizes):
    rng = np.random.default_rng(seed)
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    cols = {k: [] for k in ("times", "waves", "mags", "noise_var")}
    for n in sizes:
        tt = rng.uniform(0, 120, n)
        ww = rng.choice(np.log10([3551.0, 4770.0, 6231.0, 7625.0, 9134.0]), n)
        cols["times"].append(tt)
        cols["waves"].append(ww)
        cols["mags"].append(19 - 1.5 * np.exp(-((tt - rng.uniform(20, 80)) / 15) ** 2) + 0.6 * (ww - 3.75)
                            + 0.03 * rng.normal(size=n))
        cols["noise_var"].append(rng.uniform(1e-4, 4e-3, n))
    return {"offsets": off, **{k: np.concatenate(v) if v else np.zeros(0) for k, v in cols.items()}}


AMPS2 = np.array([0.1, 0.5, 2.0])
LST2 = np.array([0.05, 5.0, 15.0, 40.0])
LSW2 = np.array([0.0005, 0.05, 0.2, 1.0])


@pytest.mark.parametrize("cap,rule,crit", [(40, 0, 0), (100, 1, 1), (128, 0, 0), (200, 1, 1)])
def test_gp2d_matches_oracle(ctx, orc, cap, rule, crit):
    sizes = [0, 2, 3, 7, 39, 40, 41, 99, 100, 128, 129, 300]
    csr = sources_2d(cap + rule, sizes)
    qt = np.tile(np.linspace(0, 120, 40), 3)
    qw = np.repeat(np.log10([4770.0, 6231.0, 7625.0]), 40)
    g = ctx.gp2d_fit(csr, qt, qw, AMPS2, LST2, LSW2, cap=cap, subsample_rule=rule, criterion=crit)
    o = orc.batch_gp2d_fit(csr, qt, qw, AMPS2, LST2, LSW2, cap, rule, crit)
    h = 3 * cap
    assert np.array_equal(g["state"][:, h:h + 3], o["state"][:, h:h + 3])
    assert np.array_equal(g["state"][:, h + 4], o["state"][:, h + 4])
    assert np.array_equal(g["state"][:, cap:h], o["state"][:, cap:h])
    assert close(g["state"][:, :cap], o["state"][:, :cap], 1e-7)
    assert close(g["pred"], o["pred"]) and close(g["dpred"], o["dpred"], 1e-7)
    assert close(g["std"], o["std"], 1e-6)
    assert close(g["pred_obs"], o["pred_obs"])


def test_gp2d_reference_api(ctx, pkg, orc):
    """batch_gp_2d (src/gpu_cuda.rs:1152-1306): cap min(max_subsample, 40),
    hyperparameters as length scales, train RMS over at most 100 of the
    source's observations at floor(i n / 100)."""
    csr = sources_2d(3, [25, 2, 160, 60])
    off = csr["offsets"]
    srcs = [pkg.Gp2dInput(*(csr[k][off[s]:off[s + 1]] for k in ("times", "waves", "mags", "noise_var")))
            for s in range(4)]
    qt, qw = np.linspace(0, 120, 30), np.full(30, np.log10(6231.0))
    out = ctx.batch_gp_2d(srcs, qt, qw, AMPS2, LST2, LSW2, 64)
    o = orc.batch_gp2d_fit(csr, qt, qw, AMPS2, LST2, LSW2, 40)
    assert not out[1].success and np.isnan(out[1].train_rms) and out[1].n_train == 2
    for s in (0, 2, 3):
        r = out[s]
        assert r.success and r.n_train == off[s + 1] - off[s]
        assert close(np.asarray(r.pred_grid)[None], o["pred"][s:s + 1])
        assert close(np.asarray(r.std_grid)[None], o["std"][s:s + 1], 1e-6)
        st = o["state"][s]
        assert r.amp == st[120] and r.ls_time == pytest.approx(np.sqrt(0.5 / st[121]), rel=1e-14)
        assert r.ls_wave == pytest.approx(np.sqrt(0.5 / st[122]), rel=1e-14)
        n = r.n_train
        ne = min(n, 100)
        idx = (np.arange(ne) * (n / ne if ne < n else 1.0)).astype(int)
        po = o["pred_obs"][off[s]:off[s + 1]]
        want = np.sqrt(np.mean((po[idx] - csr["mags"][off[s]:off[s + 1]][idx]) ** 2))
        assert r.train_rms == pytest.approx(want, rel=1e-8)
