"""GPU parity of the batched finalize step (needs an MI355X) against the
oracle's profile_t0_refine / profile_refine_all / magnitude chi2, through
lcf_batch_finalize:

  * all eight models (MetzgerKN in its 200-point batch form), ragged bands of
    1-150 observations with upper limits, negative fluxes, an empty band;
  * each step alone (flag bits), all three together, and with the magnitude
    chi2 taken from the posterior mean only (the multi-band route);
  * refined t0 identical (grid values), log sigma and chi2 within 1e-9
    relative: every likelihood is one lane's sum over the observations in
    the reference's order, so only the device exponential / logarithm
    rounding differs;
  * the reference API (finalize_parametric_with_gpu_svi and
    finalize_all_models_with_gpu_svi on BandData maps, and the feature
    extractor's whole GPU parametric path).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def random_jobs(orc, n_per_model, seed, n_max=150):
    rng = np.random.default_rng(seed)
    bands, models, pso, mu, ls, fe_all, mags, peaks = [], [], [], [], [], [], [], []
    for m in range(8):
        lo, hi = orc.bounds(m)
        np_ = lo.size
        k = 0
        while k < n_per_model:
            n = int(rng.integers(1, n_max + 1))
            t = np.sort(rng.uniform(0.0, 120.0, n))
            p = lo + (hi - lo) * rng.uniform(0.2, 0.8, np_)
            if m != 7:
                p[{0: 2, 1: 3, 2: 2}.get(m, 1)] = rng.uniform(-10.0, 60.0)
            else:
                p[3] = rng.uniform(-5.0, 30.0)
            with np.errstate(all="ignore"):
                f = orc.eval_batch(m, p, t).reshape(-1)
            if not np.all(np.isfinite(f)):
                continue
            raw = 100.0 * f + rng.normal(0.0, 3.0, n) + 2.0
            err = np.full(n, 3.0) * rng.uniform(0.5, 1.5, n)
            peak = raw.max()
            if peak <= 0:
                continue
            nf, ne = raw / peak, err / peak
            up = ((err > 0) & (raw / err < 3.0)).astype(np.int32)
            bands.append((t, nf, ne * ne + 1e-10, up, 3.0 * err / peak))
            fe_all.append(ne)
            with np.errstate(all="ignore"):
                mags.append(-2.5 * np.log10(raw) + 23.9)
            peaks.append(peak)
            models.append(m)
            q = np.zeros(7)
            q[:np_] = p
            pso.append(q)
            mm = q.copy()
            mm[:np_] += rng.normal(0.0, 0.05, np_)
            mu.append(mm)
            ll = np.zeros(7)
            ll[:np_] = rng.uniform(-4.0, 0.5, np_)
            ls.append(ll)
            k += 1
    # an empty band
    bands.append((np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0, np.int32), np.zeros(0)))
    fe_all.append(np.zeros(0)); mags.append(np.zeros(0)); peaks.append(1.0); models.append(1)
    pso.append(np.zeros(7)); mu.append(np.arange(7.0)); ls.append(-np.arange(7.0))
    off = np.zeros(len(bands) + 1, dtype=np.int64)
    np.cumsum([b[0].size for b in bands], out=off[1:])
    csr = {"offsets": off}
    for i, k in enumerate(("times", "flux", "obs_var", "is_upper", "upper_flux")):
        csr[k] = np.concatenate([b[i] for b in bands])
    return csr, np.concatenate(fe_all), np.concatenate(mags), np.array(peaks), np.array(models, np.int32), \
        np.array(pso), np.array(mu), np.array(ls)


@pytest.fixture(scope="module")
def jobs(orc):
    return random_jobs(orc, 6, seed=11)


@pytest.mark.parametrize("flags", [1, 2, 4, 7, 15])
def test_matches_oracle(pkg, orc, ctx, jobs, flags):
    csr, fe, mags, peaks, models, pso, mu, ls = jobs
    data = pkg.GpuBatchData(ctx, csr=csr)
    try:
        g = ctx.finalize_arrays(data, models, pso, mu, ls, csr["flux"], fe, mags, peaks, flags=flags)
    finally:
        data.close()
    o = orc.finalize(csr, fe, mags, peaks, models, pso, mu, ls, flags)
    assert np.array_equal(g["mu"], o["mu"]), np.argwhere(g["mu"] != o["mu"])[:5]
    assert np.allclose(g["log_sigma"], o["log_sigma"], rtol=1e-9, atol=1e-12)
    fin = np.isfinite(o["mag_chi2"])
    assert np.array_equal(fin, np.isfinite(g["mag_chi2"]))
    assert np.allclose(g["mag_chi2"][fin], o["mag_chi2"][fin], rtol=1e-9, atol=0)
    if flags & 4:
        assert fin.sum() > 0.8 * fin.size
    if flags & 1:
        t0_idx = np.array([2, 3, 2, 1, 1, 1, 1, 3])[models]
        moved = g["mu"][np.arange(models.size), t0_idx] != mu[np.arange(models.size), t0_idx]
        assert moved.sum() > 0.8 * (models.size - 1)
    # the empty band passes through
    assert np.array_equal(g["mu"][-1], mu[-1]) and np.array_equal(g["log_sigma"][-1], ls[-1])
    assert math.isnan(g["mag_chi2"][-1])


def test_invalid_arguments(pkg, ctx, jobs):
    csr, fe, mags, peaks, models, pso, mu, ls = jobs
    data = pkg.GpuBatchData(ctx, csr=csr)
    try:
        bad = models.copy()
        bad[0] = 9
        with pytest.raises(pkg.GpuError):
            ctx.finalize_arrays(data, bad, pso, mu, ls, csr["flux"], fe, mags, peaks)
        with pytest.raises(pkg.GpuError):
            ctx.finalize_arrays(data, models, pso, mu, ls, csr["flux"], fe, mags, peaks, flags=16)
        with pytest.raises(pkg.GpuError):
            ctx.finalize_arrays(data, models, pso, mu, ls, csr["flux"][:-1], fe, mags, peaks)
    finally:
        data.close()


def _flux_bands(pkg, rng, n_bands=3):
    out = {}
    for k, name in enumerate(["g", "r", "i"][:n_bands]):
        n = int(rng.integers(8, 60))
        t = np.sort(rng.uniform(0.0, 90.0, n))
        f = 200.0 * np.exp(-(t - 30.0 - 3 * k) / 20.0) / (1.0 + np.exp(-(t - 30.0 - 3 * k) / 3.0))
        f = f + rng.normal(0.0, 4.0, n)
        out[name] = pkg.BandData(t, f, np.full(n, 4.0))
    return out


def test_reference_api(pkg, orc, ctx):
    par = pkg.parametric
    rng = np.random.default_rng(5)
    fb = _flux_bands(pkg, rng)
    prepared = par.prepare_bands(fb)
    assert [p[1].times.size for p in prepared] == sorted((b.times.size for b in fb.values()), reverse=True)
    gpu_results, svi_outputs = [], []
    for k, (name, data, mags) in enumerate(prepared):
        m = [pkg.GpuModelName.Bazin, pkg.GpuModelName.Villar, pkg.GpuModelName.Afterglow][k]
        lo, hi = orc.bounds(int(m))
        p = list(lo + 0.5 * (hi - lo))
        gpu_results.append(par.GpuPsoBandResult(model=m, pso_params=p, pso_cost=1.5,
                                                per_model_chi2={m: 3.0}))
        svi_outputs.append((list(np.array(p) + 0.01), list(np.full(len(p), -2.0)), -12.5))
    res = par.finalize_parametric_with_gpu_svi(ctx, fb, gpu_results, svi_outputs)
    assert [r.band for r in res] == [p[0] for p in prepared]
    for k, r in enumerate(res):
        name, data, mags = prepared[k]
        csr = {"offsets": np.array([0, data.times.size]), "times": data.times, "flux": data.flux,
               "obs_var": data.obs_var, "is_upper": data.is_upper, "upper_flux": data.upper_flux}
        np_ = len(gpu_results[k].pso_params)
        pad = lambda v: np.r_[v, np.zeros(7 - len(v))]
        o = orc.finalize(csr, data.flux_err, mags, [data.peak_flux_obs], [int(gpu_results[k].model)],
                         pad(gpu_results[k].pso_params), pad(svi_outputs[k][0]), pad(svi_outputs[k][1]))
        assert np.array_equal(r.svi_mu, o["mu"][0, :np_])
        assert np.allclose(r.svi_log_sigma, o["log_sigma"][0, :np_], rtol=1e-9)
        assert r.mag_chi2 == pytest.approx(o["mag_chi2"][0], rel=1e-9)
        assert r.svi_elbo == -12.5 and r.pso_chi2 == 1.5 and r.n_obs == data.times.size
        assert r.uncertainty_method == par.UncertaintyMethod.Svi
    # all-models form: one entry per (band, model), names given explicitly
    names = [prepared[0][0]] * 2 + ["missing"]
    res2 = par.finalize_all_models_with_gpu_svi(ctx, fb, gpu_results[:2] + gpu_results[:1], svi_outputs[:3], names)
    assert [r.band for r in res2] == [prepared[0][0]] * 2


def test_from_gpu_and_pipeline(pkg, ctx):
    par = pkg.parametric
    rng = np.random.default_rng(9)
    sources = [_flux_bands(pkg, rng, n_bands=1 + s % 3) for s in range(4)]
    out = par.fit_parametric_batch_gpu(ctx, sources)
    assert [len(r) for r in out] == [8 * len(s) for s in sources]
    for s, rs in enumerate(out):
        for r in rs:
            assert len(r.svi_mu) == r.model.n_params() == len(r.svi_log_sigma)
            assert r.svi_elbo is not None and all(np.isfinite(r.svi_mu)) and all(np.isfinite(r.svi_log_sigma))
            assert r.pso_chi2 == pytest.approx(r.per_model_chi2[r.model] / 2.0)
        # MultiBazin on the best model of the first band only
        assert sum(r.multi_bazin is not None for r in rs) == 1
    # finalize_parametric_from_gpu, both posterior methods
    gpu_results = [par.GpuPsoBandResult(model=r.model, pso_params=r.pso_params, pso_cost=r.pso_chi2 / 2.0)
                   for r in out[0] if r.model in (pkg.GpuModelName.Bazin,)]
    for method in (par.UncertaintyMethod.Svi, par.UncertaintyMethod.Laplace):
        res = par.finalize_parametric_from_gpu(ctx, sources[0], gpu_results, method)
        assert len(res) == 1 and res[0].uncertainty_method == method
        assert all(np.isfinite(res[0].svi_log_sigma)) and res[0].mag_chi2 is not None
