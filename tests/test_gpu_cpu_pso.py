"""GPU parity of the CPU-path swarm engine (needs an MI355X) against the
oracle through lcf_cpu_pso: per-band swarms of all eight models (plain,
seeded, t0 held fixed), joint multi-band swarms of 1-3 bands, MultiBazin
swarms of K = 1..4 components, and L-BFGS
polish jobs -- swarm best positions and costs identical to 1e-12 relative
(the same Philox stream, the same summation order; only the device
exponential and logarithm round differently), polish end points close and
never above their start -- and the reference API (fit_parametric on
BandData maps, multi-band and per-band routes, both posterior methods)."""
import numpy as np
import pytest

from test_oracle_cpu_pso import job, OBS

pytestmark = pytest.mark.gpu


def bands_csr(pkg, orc, rng, n_bands):
    c = pkg.synthetic.bazin_mag_batch(1, seed=int(rng.integers(1 << 30)), n_bands=n_bands, n_min=15, n_max=60)
    flux = 10 ** (-0.4 * (c["mags"] - 23.9))
    err = c["mag_errs"] * flux / 1.0857362047581294
    off = c["offsets"]
    cols = {k: [] for k in OBS}
    for b in range(n_bands):
        sl = slice(off[b], off[b + 1])
        d = pkg.parametric.band_fit_data(pkg.BandData(c["times"][sl], flux[sl], err[sl]))[0]
        cols["times"].append(d.times)
        cols["flux"].append(d.flux)
        cols["obs_var"].append(d.obs_var)
        cols["is_upper"].append(d.is_upper)
        cols["upper_flux"].append(d.upper_flux)
    return {"offsets": off, **{k: np.concatenate(v) for k, v in cols.items()}}


def test_engine_matches_oracle(pkg, orc, ctx):
    rng = np.random.default_rng(21)
    csr = bands_csr(pkg, orc, rng, 6)
    jobs = []
    for m in range(8):
        lo, hi = orc.bounds(m, cpu=True)
        for b in range(3):
            x0 = lo + 0.5 * (hi - lo)
            jobs.append(job(orc, kind=orc.CPSO_SWARM_BAND, model=m, band=b, dim=lo.size, n_particles=16,
                            max_iters=25, stall_iters=6, use_prior=1, seeded=b % 2, lo=lo, hi=hi, x0=x0,
                            seed=42 + b, stream=b))
        t0i = [2, 3, 2, 1, 1, 1, 1, 3][m]
        lr, hr = np.delete(lo, t0i), np.delete(hi, t0i)
        jobs.append(job(orc, kind=orc.CPSO_SWARM_BAND, model=m, band=3, dim=lr.size, n_particles=12, max_iters=20,
                        stall_iters=5, use_prior=0, fixed_t0=1, t0_value=20.0, lo=lr, hi=hr, seed=42, stream=3))
    for B in (1, 2, 3):
        lo, hi = orc.bounds(1, cpu=True)
        lo = np.r_[lo, np.full(7 * (B - 1), -0.5)]
        hi = np.r_[hi, np.full(7 * (B - 1), 0.5)]
        jobs.append(job(orc, kind=orc.CPSO_SWARM_JOINT, model=1, band=B, n_bands=B, dim=7 * B, n_particles=40,
                        max_iters=30, stall_iters=10, seeded=1, lo=lo, hi=hi, x0=0.5 * (lo + hi), seed=137,
                        stream=B))
    for k in (1, 2, 3, 4):  # fit_multi_bazin's swarms
        t = csr["times"][csr["offsets"][4]:csr["offsets"][5]]
        lo, hi = [], []
        for _ in range(k):
            lo += [-3.0, t.min() - 30.0, -2.0, -2.0]
            hi += [3.0, t.max() + 30.0, 5.0, 6.0]
        lo, hi = np.array(lo + [-0.3, -5.0]), np.array(hi + [0.3, 0.0])
        jobs.append(job(orc, kind=orc.CPSO_SWARM_MULTIBAZIN, model=k, band=4, dim=4 * k + 2, n_particles=30,
                        max_iters=60, stall_iters=12, seeded=int(k > 1), lo=lo, hi=hi, x0=0.5 * (lo + hi),
                        seed=137))
    jobs = np.array(jobs)
    o = orc.cpu_pso(csr, jobs)
    # polish from the swarm results of the closed-form models
    pol = []
    for k, j in enumerate(jobs):
        if j["kind"] == orc.CPSO_SWARM_BAND and j["model"] != 7 and not j["fixed_t0"] and o["cost"][k] < 1e98:
            pol.append(job(orc, kind=orc.CPSO_POLISH, model=j["model"], band=j["band"], dim=j["dim"], max_iters=50,
                           use_prior=1, lo=j["lo"][:j["dim"]], hi=j["hi"][:j["dim"]], x0=o["x"][k, :j["dim"]],
                           start_cost=o["cost"][k]))
    allj = np.concatenate([jobs, np.array(pol)])
    o = orc.cpu_pso(csr, allj)
    data = pkg.GpuBatchData(ctx, csr=csr)
    try:
        g = ctx.cpu_pso(data, allj)
    finally:
        data.close()
    ns = jobs.size
    # swarms: the same trajectory
    assert np.array_equal(g["iters"][:ns], o["iters"][:ns]), np.nonzero(g["iters"][:ns] != o["iters"][:ns])[0]
    assert np.allclose(g["cost"][:ns], o["cost"][:ns], rtol=1e-12, atol=0)
    assert np.allclose(g["x"][:ns], o["x"][:ns], rtol=1e-10, atol=1e-12)
    # polish: an L-BFGS path is decided by line-search tests that last-bit
    # differences of the gradient's exponentials can flip, so the end points
    # agree closely for most jobs and stay valid (never above the start) for all
    gp, op, start = g["cost"][ns:], o["cost"][ns:], allj["start_cost"][ns:]
    assert np.all(gp <= start) and np.all(op <= start)
    close = np.abs(gp - op) <= 1e-6 * np.abs(op)
    assert close.mean() >= 0.75, (gp, op)
    assert np.all(np.abs(gp - op) <= 0.05 * np.abs(op) + 1e-9)


def test_invalid_jobs(pkg, orc, ctx):
    rng = np.random.default_rng(22)
    csr = bands_csr(pkg, orc, rng, 2)
    data = pkg.GpuBatchData(ctx, csr=csr)
    lo, hi = orc.bounds(0, cpu=True)
    try:
        for bad in (dict(dim=7), dict(band=5), dict(model=9), dict(n_particles=0)):
            kw = dict(kind=orc.CPSO_SWARM_BAND, model=0, band=0, dim=6, n_particles=8, max_iters=5, stall_iters=2,
                      lo=lo, hi=hi)
            kw.update(bad)
            with pytest.raises(pkg.GpuError):
                ctx.cpu_pso(data, np.array([job(orc, **kw)]))
        with pytest.raises(pkg.GpuError):  # MetzgerKN has no polish
            ctx.cpu_pso(data, np.array([job(orc, kind=orc.CPSO_POLISH, model=7, band=0, dim=5, max_iters=5,
                                            lo=orc.bounds(7)[0], hi=orc.bounds(7)[1])]))
    finally:
        data.close()


def _sources(pkg, n, n_bands, seed):
    c = pkg.synthetic.bazin_mag_batch(n, seed=seed, n_bands=n_bands, n_min=20, n_max=40, t_span=100.0)
    flux = 10 ** (-0.4 * (c["mags"] - 23.9))
    err = c["mag_errs"] * flux / 1.0857362047581294
    off = c["offsets"]
    out = []
    for s in range(n):
        out.append({nm: pkg.BandData(c["times"][off[n_bands * s + k]:off[n_bands * s + k + 1]],
                                     flux[off[n_bands * s + k]:off[n_bands * s + k + 1]],
                                     err[off[n_bands * s + k]:off[n_bands * s + k + 1]])
                    for k, nm in enumerate("gri"[:n_bands])})
    return out


@pytest.mark.parametrize("method", ["Laplace", "Svi"])
def test_fit_parametric(pkg, ctx, method):
    par = pkg.parametric
    um = par.UncertaintyMethod[method]
    multi = _sources(pkg, 2, 3, 31)
    single_band = _sources(pkg, 2, 1, 32)
    out = par.fit_parametric_batch(ctx, multi + single_band, method=um)
    for s in range(2):  # joint Villar route: one result per band, shared cost
        assert [r.band for r in out[s]] == [nm for nm, _, _ in par.prepare_bands(multi[s])]
        assert all(r.model == pkg.GpuModelName.Villar for r in out[s])
        assert len({r.pso_chi2 for r in out[s]}) == 1 and out[s][0].pso_chi2 is not None
        assert out[s][0].multi_bazin is not None and all(r.multi_bazin is None for r in out[s][1:])
        for r in out[s]:
            assert len(r.svi_mu) == 7 and np.all(np.isfinite(r.svi_log_sigma)) and r.uncertainty_method == um
    for s in range(2, 4):  # per-band route: model selection by BIC
        assert len(out[s]) == 1
        r = out[s][0]
        assert r.per_model_chi2[pkg.GpuModelName.Bazin] is not None
        assert len(r.pso_params) == r.model.n_params() == len(r.svi_mu)
        assert r.mag_chi2 is None or np.isfinite(r.mag_chi2)
    # fit_all_models forces the per-band route and fills every model
    allm = par.fit_parametric_batch(ctx, multi[:1], fit_all_models=True, method=um)
    assert len(allm[0]) == 3
    for r in allm[0]:
        assert set(r.per_model_params) == set(pkg.GpuModelName)
