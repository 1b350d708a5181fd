"""GPU parity of the batched fit_gp_2d (needs an MI355X) against the oracle,
through lcf_batch_fit_gp_2d: selected hyper-parameters identical, train RMS,
predictive mean and std within 1e-9 relative (1e-7 for the std), sources of
3-600 points (the subsample path above 200), too-short sources unfitted; and
the reference API (fit_gp_2d_with_thermal_batch on magnitude-band maps)
against the oracle's fit followed by its blackbody fit."""
import numpy as np
import pytest

from test_oracle_fit_gp_2d import make

pytestmark = pytest.mark.gpu


def test_matches_oracle(orc, ctx):
    rng = np.random.default_rng(11)
    sizes = [3, 8, 64, 150, 200, 201, 333, 600]
    c = make(rng, sizes)
    qt = np.tile(np.linspace(-5, 85, 12), (len(sizes), 1))
    qw = np.tile(np.log10([4770.0, 6231.0, 7625.0]), (len(sizes), 4))
    g = ctx.fit_gp_2d_arrays(c, qt, qw)
    o = orc.batch_fit_gp_2d(c, qt, qw)
    assert np.array_equal(g["n_train"], o["n_train"])
    fit = o["n_train"] > 0
    assert fit.all()
    assert np.array_equal(g["hyp"], o["hyp"])
    assert np.allclose(g["train_rms"], o["train_rms"], rtol=1e-9)
    assert np.allclose(g["pred"], o["pred"], rtol=1e-9, atol=1e-9)
    assert np.allclose(g["std"], o["std"], rtol=1e-7, atol=1e-9)
    short = make(rng, [2, 0])
    gs = ctx.fit_gp_2d_arrays(short, np.zeros((2, 3)), np.zeros((2, 3)))
    assert list(gs["n_train"]) == [0, 0] and np.isnan(gs["hyp"]).all() and np.isnan(gs["pred"]).all()


def test_reference_api(pkg, orc, ctx):
    gp2 = pkg.gp2d
    rng = np.random.default_rng(12)
    sources = []
    for s in range(4):
        bands = {}
        for nm in ("g", "r", "i", "x"):  # "x": no known wavelength, ignored
            n = int(rng.integers(4, 90))
            t = np.sort(rng.uniform(0, 60, n))
            bands[nm] = pkg.BandData(t, 18 + np.sin(t / 9 + s) + 0.02 * rng.normal(size=n), np.full(n, 0.05))
        sources.append(bands)
    sources.append({"g": pkg.BandData([1.0, 2.0, 3.0], [18.0, 18.1, 18.2], [0.1] * 3)})  # one band: None
    out = gp2.fit_gp_2d_with_thermal_batch(ctx, sources, 7)
    assert out[4] is None
    for s in range(4):
        res, th = out[s]
        b = {k: v for k, v in sources[s].items() if k != "x"}
        t = np.concatenate([v.times for v in b.values()])
        w = np.concatenate([np.full(len(v.times), np.log10(gp2.BAND_WAVELENGTH[k])) for k, v in b.items()])
        y = np.concatenate([v.values for v in b.values()])
        nv = np.concatenate([np.asarray(v.errors) ** 2 for v in b.values()])
        lo, hi = gp2._subsample_range(t)  # the training subsample's time span
        et = lo + (hi - lo) * np.arange(7) / 6.0
        lw = np.log10([gp2.BAND_WAVELENGTH[k] for k in b])
        o = orc.batch_fit_gp_2d({"offsets": np.array([0, t.size]), "times": t, "waves": w, "mags": y,
                                 "noise_var": nv}, np.repeat(et, 3)[None], np.tile(lw, 7)[None])
        assert (res.amp, res.ls_time, res.ls_wave) == tuple(o["hyp"][0])
        assert res.train_rms == pytest.approx(o["train_rms"][0], rel=1e-9)
        assert res.bands == ["g", "r", "i"] and res.n_train == t.size
        want = []
        for ti in range(7):
            r = orc.fit_bb_temperature([gp2.BAND_WAVELENGTH[k] for k in b], o["pred"][0, 3 * ti:3 * ti + 3],
                                       o["std"][0, 3 * ti:3 * ti + 3])
            if r is not None:
                want.append((et[ti], r[0], r[1]))
        assert len(th.times) == len(want)
        for k, (tt, lt, c2) in enumerate(want):
            assert th.times[k] == pytest.approx(tt, rel=1e-12)
            assert th.log_temps[k] == pytest.approx(lt, rel=1e-9, abs=1e-9)
            assert th.chi2s[k] == pytest.approx(c2, rel=1e-6, abs=1e-9)
