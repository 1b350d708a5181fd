"""GPU parity of the blackbody grid fit and of fit_gp_2d_batch_gpu (needs an
MI355X): lcf_batch_bb_temperature against the oracle's fit_bb_temperature
(grid temperature identical, chi2 / offset to 1e-9), and the 2D GP + thermal
orchestrator against the oracle's batch_gp2d_fit + fit_bb_temperature on the
same bookkeeping (log T identical at >= 99% of the time steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bb_matches_oracle(orc, ctx):
    rng = np.random.default_rng(5)
    sizes = rng.integers(0, 6, size=300)
    off = np.zeros(sizes.size + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    lam = rng.choice([3671.0, 4770.0, 6231.0, 7625.0, 8691.0], size=off[-1])
    mag = 18 + rng.normal(size=off[-1])
    err = rng.uniform(0.01, 0.3, size=off[-1])
    g = ctx.bb_temperature(off, lam, mag, err)
    for s in range(sizes.size):
        o = orc.fit_bb_temperature(lam[off[s]:off[s + 1]], mag[off[s]:off[s + 1]], err[off[s]:off[s + 1]])
        if o is None:
            assert g["ok"][s] == 0
            continue
        assert g["ok"][s] == 1 and g["log_t"][s] == o[0]
        assert g["chi2"][s] == pytest.approx(o[1], rel=1e-9) and g["offset"][s] == pytest.approx(o[2], rel=1e-9)


def test_gp2d_thermal_orchestrator(pkg, orc, ctx):
    srcs = [pkg.common.build_mag_bands(*pkg.synthetic.reference_test_source(n, 70 + k))
            for k, n in enumerate([12, 20, 30, 45])]
    srcs.append({"g": pkg.nonparametric.BandData([0.0, 1.0], [18.0, 18.1], [0.1, 0.1])})  # too few points
    nt = 12
    res = pkg.gp2d.fit_gp_2d_batch_gpu(ctx, srcs, nt)
    assert res[4] is None and all(r is not None for r in res[:4])
    lw = np.log10([4770.0, 6231.0, 7625.0])
    tmin = min(min(b.times) for s in srcs[:4] for b in s.values())
    tmax = max(max(b.times) for s in srcs[:4] for b in s.values())
    dur = max(tmax - tmin, 1.0)
    et = tmin + dur * np.arange(nt) / (nt - 1)
    qt, qw = np.repeat(et, 3), np.tile(lw, nt)
    wr = max(lw[-1] - lw[0], 0.01)
    agree = total = 0
    for s in range(4):
        b = srcs[s]
        t = np.concatenate([b[n].times for n in "gri"])
        w = np.concatenate([np.full(len(b[n].times), np.log10(x)) for n, x in zip("gri", (4770.0, 6231.0, 7625.0))])
        m = np.concatenate([b[n].values for n in "gri"])
        e = np.concatenate([b[n].errors for n in "gri"])
        csr = {"offsets": np.array([0, t.size]), "times": t, "waves": w, "mags": m, "noise_var": e * e}
        o = orc.batch_gp2d_fit(csr, qt, qw, [0.3, 1.0, 3.0, 10.0], [dur / 4, dur / 8, dur / 16, dur / 32],
                               [wr * 2, wr, wr / 2, wr / 4], 40)
        r2, th = res[s]
        assert r2.bands == ["g", "r", "i"] and r2.n_train == t.size
        for ti in range(nt):
            p, sd = o["pred"][0, 3 * ti:3 * ti + 3], o["std"][0, 3 * ti:3 * ti + 3]
            want = orc.fit_bb_temperature([4770.0, 6231.0, 7625.0], p, sd)
            total += 1
            if et[ti] in th.times:
                k = th.times.index(et[ti])
                agree += want is not None and th.log_temps[k] == want[0]
    assert agree >= 0.99 * total
