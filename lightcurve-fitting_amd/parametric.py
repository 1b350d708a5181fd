"""Parametric band results from device fits: the reference's finalize
functions (src/parametric.rs:3126-3553) and the GPU parametric pipeline of its
feature extractor (src/bin/extract_features.rs:585-739), with every per-band
step on the device:

  * PSO over all models (``lcf_batch_all_models``) and MultiBazin on each
    source's most populated band (``lcf_batch_multi_bazin``);
  * SVI (``lcf_batch_svi_fit_ex``) or Laplace (``lcf_batch_laplace``)
    posteriors seeded at the PSO points;
  * profile_t0_refine, profile_refine_all and the magnitude chi2
    (``lcf_batch_finalize``), which the reference runs per band on the host.

The host keeps only the bookkeeping: band normalisation into BandFitData
(src/parametric.rs:3313-3354), the (band, model) job lists and the result
records.  There is no CPU path: every call needs the HIP library.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import Mapping, Sequence

import numpy as np

from .gpu import (ALL_GPU_MODELS, FINALIZE_DEFAULT, FP64, MAX_PARAMS, GpuBatchData, GpuContext, GpuModelName,
                  MultiBazinResult, svi_prior_for_model)
from .nonparametric import BandData

ZP = 23.9
SNR_THRESHOLD = 3.0

# SviModelName (src/parametric.rs:85-110) carries the same eight names as
# GpuModelName; results use the GpuModelName ids.
SviModelName = GpuModelName

# finalize_parametric_from_gpu's CPU svi_fit settings (src/parametric.rs:3136-3138)
SVI_LR, SVI_N_STEPS, SVI_N_SAMPLES = 0.01, 1000, 4
# the feature extractor's batched SVI (src/bin/extract_features.rs:697)
PIPELINE_SVI_STEPS, PIPELINE_SVI_SAMPLES = 150, 2
_SVI_CPU_STOPPING, _SVI_INFLATE_SIGMA, _SVI_FD_GRADIENT, _SVI_SHARED_STREAM = 1, 2, 4, 32


class UncertaintyMethod(enum.Enum):
    """``UncertaintyMethod`` (src/parametric.rs:14-21)."""

    Svi = "Svi"
    Laplace = "Laplace"


@dataclass
class BandFitData:
    """``BandFitData`` (src/parametric.rs:3014-3025): one band normalised by
    its peak flux."""

    times: np.ndarray
    flux: np.ndarray
    flux_err: np.ndarray
    obs_var: np.ndarray
    is_upper: np.ndarray
    upper_flux: np.ndarray
    noise_frac_median: float
    peak_flux_obs: float


@dataclass
class GpuPsoBandResult:
    """``GpuPsoBandResult`` (src/parametric.rs:3100-3114)."""

    model: GpuModelName
    pso_params: list
    pso_cost: float
    per_model_chi2: dict = field(default_factory=dict)
    per_model_params: dict = field(default_factory=dict)
    multi_bazin: MultiBazinResult | None = None


@dataclass
class ParametricBandResult:
    """``ParametricBandResult`` (src/parametric.rs:3049-3072)."""

    band: str
    model: GpuModelName
    pso_params: list
    pso_chi2: float | None
    svi_mu: list
    svi_log_sigma: list
    svi_elbo: float | None
    n_obs: int
    mag_chi2: float | None
    per_model_chi2: dict
    per_model_params: dict
    uncertainty_method: UncertaintyMethod = UncertaintyMethod.Svi
    multi_bazin: MultiBazinResult | None = None


def finite_or_none(v: float) -> float | None:
    """``finite_or_none`` (src/common.rs:463-469)."""
    return float(v) if math.isfinite(v) else None


def _median(xs: np.ndarray) -> float | None:
    """``median_f64`` (src/parametric.rs:3027-3038): total order, mean of the
    two middle values for even counts."""
    if xs.size == 0:
        return None
    s = np.sort(xs, kind="stable")
    mid = s.size // 2
    return float((s[mid - 1] + s[mid]) / 2.0) if s.size % 2 == 0 else float(s[mid])


def band_fit_data(band: BandData) -> tuple[BandFitData, np.ndarray] | None:
    """Peak normalisation of one flux band (src/parametric.rs:3313-3354);
    returns the BandFitData and the observed magnitudes, or None for an empty
    band or a band whose peak flux is not positive."""
    f = np.asarray(band.values, dtype=np.float64)
    e = np.asarray(band.errors, dtype=np.float64)
    t = np.asarray(band.times, dtype=np.float64)
    if f.size == 0:
        return None
    fin = f[~np.isnan(f)]
    peak = max(float(fin.max()), np.finfo(np.float64).min) if fin.size else np.finfo(np.float64).min
    if peak <= 0.0:
        return None
    with np.errstate(divide="ignore", invalid="ignore"):
        nf = f / peak
        ne = e / peak
        is_up = (e > 0.0) & (f / e < SNR_THRESHOLD)
        uf = SNR_THRESHOLD * e / peak
        pos = nf > 0.0
        noise = _median(ne[pos] / nf[pos])
        mags = -2.5 * np.log10(f) + ZP
    return (BandFitData(times=t.copy(), flux=nf, flux_err=ne, obs_var=ne * ne + 1e-10,
                        is_upper=is_up.astype(np.int32), upper_flux=uf,
                        noise_frac_median=noise if noise is not None else 0.0, peak_flux_obs=peak), mags)


def prepare_bands(flux_bands: Mapping[str, BandData]) -> list[tuple[str, BandFitData, np.ndarray]]:
    """Normalised bands, most populated first (stable for equal counts, in
    the mapping's order)."""
    out = []
    for name, b in flux_bands.items():
        r = band_fit_data(b)
        if r is not None:
            out.append((name, r[0], r[1]))
    out.sort(key=lambda x: -x[1].times.size)
    return out


def _fit_csr(datas: Sequence[BandFitData]) -> dict:
    off = np.zeros(len(datas) + 1, dtype=np.int64)
    np.cumsum([d.times.size for d in datas], out=off[1:])
    cat = lambda k, dt=np.float64: (np.concatenate([np.asarray(getattr(d, k), dtype=dt) for d in datas])
                                    if datas else np.zeros(0, dtype=dt))
    return {"offsets": off, "times": cat("times"), "flux": cat("flux"), "obs_var": cat("obs_var"),
            "is_upper": cat("is_upper", np.int32), "upper_flux": cat("upper_flux"), "flux_err": cat("flux_err")}


def finalize_jobs(ctx: GpuContext, datas: Sequence[BandFitData], mags: Sequence[np.ndarray], model_ids,
                  pso_params, mu, log_sigma, *, flags: int = FINALIZE_DEFAULT) -> dict:
    """``lcf_batch_finalize`` over a list of (band, model) jobs."""
    n = len(datas)
    if n == 0:
        return {"mu": np.zeros((0, MAX_PARAMS)), "log_sigma": np.zeros((0, MAX_PARAMS)), "mag_chi2": np.zeros(0)}
    csr = _fit_csr(datas)
    data = GpuBatchData(ctx, csr=csr)
    try:
        return ctx.finalize_arrays(data, model_ids, pso_params, mu, log_sigma, csr["flux"], csr["flux_err"],
                                   np.concatenate([np.asarray(m, dtype=np.float64) for m in mags]),
                                   [d.peak_flux_obs for d in datas], flags=flags)
    finally:
        data.close()


def _pad7(v: Sequence[float]) -> np.ndarray:
    out = np.zeros(MAX_PARAMS)
    k = min(len(v), MAX_PARAMS)
    out[:k] = np.asarray(v, dtype=np.float64)[:k]
    return out


def _results(ctx, jobs, method: UncertaintyMethod) -> list[ParametricBandResult]:
    """Finalize the (name, data, mags, gpu result, (mu, log sigma, elbo),
    multi_bazin) jobs on the device and build the result records."""
    if not jobs:
        return []
    models = [int(j[3].model) for j in jobs]
    fin = finalize_jobs(ctx, [j[1] for j in jobs], [j[2] for j in jobs], models,
                        np.stack([_pad7(j[3].pso_params) for j in jobs]),
                        np.stack([_pad7(j[4][0]) for j in jobs]), np.stack([_pad7(j[4][1]) for j in jobs]))
    out = []
    for k, (name, data, _mags, gpu, svi, mb) in enumerate(jobs):
        np_ = GpuModelName(models[k]).n_params()
        out.append(ParametricBandResult(
            band=name, model=gpu.model, pso_params=list(gpu.pso_params), pso_chi2=finite_or_none(gpu.pso_cost),
            svi_mu=[float(x) for x in fin["mu"][k, :np_]], svi_log_sigma=[float(x) for x in fin["log_sigma"][k, :np_]],
            svi_elbo=finite_or_none(svi[2]), n_obs=int(data.times.size),
            mag_chi2=finite_or_none(fin["mag_chi2"][k]), per_model_chi2=dict(gpu.per_model_chi2),
            per_model_params=dict(gpu.per_model_params), uncertainty_method=method, multi_bazin=mb))
    return out


def finalize_parametric_with_gpu_svi(ctx: GpuContext, flux_bands: Mapping[str, BandData],
                                     gpu_results: Sequence[GpuPsoBandResult],
                                     svi_outputs: Sequence[tuple]) -> list[ParametricBandResult]:
    """``finalize_parametric_with_gpu_svi`` (src/parametric.rs:3298-3421):
    gpu_results[i] and svi_outputs[i] = (mu, log_sigma, elbo) belong to the
    i-th most populated band."""
    if not flux_bands or not gpu_results or not svi_outputs:
        return []
    jobs = []
    for bi, (name, data, mags) in enumerate(prepare_bands(flux_bands)):
        if bi >= len(gpu_results) or bi >= len(svi_outputs):
            break
        gpu = gpu_results[bi]
        if len(gpu.pso_params) == 0:
            continue
        jobs.append((name, data, mags, gpu, svi_outputs[bi], gpu.multi_bazin if bi == 0 else None))
    return _results(ctx, jobs, UncertaintyMethod.Svi)


def finalize_all_models_with_gpu_svi(ctx: GpuContext, flux_bands: Mapping[str, BandData],
                                     gpu_results: Sequence[GpuPsoBandResult], svi_outputs: Sequence[tuple],
                                     band_names: Sequence[str]) -> list[ParametricBandResult]:
    """``finalize_all_models_with_gpu_svi`` (src/parametric.rs:3429-3553): one
    result per (band, model) entry, band_names[i] naming entry i's band."""
    if not flux_bands or not gpu_results or not svi_outputs:
        return []
    prepared = {name: (data, mags) for name, data, mags in prepare_bands(flux_bands)}
    jobs = []
    for i, gpu in enumerate(gpu_results):
        if i >= len(svi_outputs) or i >= len(band_names):
            break
        name = band_names[i]
        if name not in prepared:
            continue
        if len(gpu.pso_params) == 0 or len(svi_outputs[i][0]) == 0:
            continue
        data, mags = prepared[name]
        jobs.append((name, data, mags, gpu, svi_outputs[i], gpu.multi_bazin))
    return _results(ctx, jobs, UncertaintyMethod.Svi)


def _svi_on_device(ctx: GpuContext, datas, models, pso, n_steps: int, n_samples: int, lr: float,
                   cpu_semantics: bool):
    """Device SVI for (band, model) jobs; returns [(mu, log_sigma, elbo)].
    With ``cpu_semantics`` the CPU fit's stopping rule and ln 4 inflation
    apply, every job draws the same noise stream (svi_fit seeds a fresh
    SmallRng(42) per band), and MetzgerKN (no analytic gradient) is
    differentiated numerically as the CPU fit does."""
    mids = np.asarray(models, dtype=np.int32)
    flags = (_SVI_CPU_STOPPING | _SVI_INFLATE_SIGMA | _SVI_SHARED_STREAM) if cpu_semantics else 0
    if cpu_semantics and ctx.precision != FP64:
        # the CPU fit is FP64 throughout: run its SVI leg in FP64 whatever
        # the context's swarm precision
        prev = ctx.precision
        ctx.set_precision(FP64)
        try:
            return _svi_on_device(ctx, datas, models, pso, n_steps, n_samples, lr, cpu_semantics)
        finally:
            ctx.set_precision(prev)
    res = {}
    for fd in (False, True):
        sel = np.nonzero((mids == int(GpuModelName.MetzgerKN)) == fd)[0]
        if sel.size == 0:
            continue
        sub = GpuBatchData(ctx, csr=_fit_csr([datas[k] for k in sel]))
        try:
            pc, pw = _priors([models[k] for k in sel], [pso[k] for k in sel])
            r = ctx.svi_fit(sub, mids[sel], np.stack([_pad7(pso[k]) for k in sel]), pc, pw, n_steps, n_samples, lr,
                            flags=flags | (_SVI_FD_GRADIENT if fd and cpu_semantics else 0))
        finally:
            sub.close()
        for j, k in enumerate(sel):
            np_ = GpuModelName(models[k]).n_params()
            res[int(k)] = (list(r["mu"][j, :np_]), list(r["log_sigma"][j, :np_]), float(r["elbo"][j]))
    return [res[k] for k in range(len(models))]


def finalize_parametric_from_gpu(ctx: GpuContext, flux_bands: Mapping[str, BandData],
                                 gpu_results: Sequence[GpuPsoBandResult],
                                 method: UncertaintyMethod = UncertaintyMethod.Svi) -> list[ParametricBandResult]:
    """``finalize_parametric_from_gpu`` (src/parametric.rs:3126-3253): the
    posterior of each band's PSO point, by SVI (the CPU fit's 1000 steps x 4
    samples, lr 0.01, stopping rule and ln 4 inflation, on the device's
    counter-based normal stream) or by the Laplace approximation, then the
    profile refinement and magnitude chi2."""
    if not flux_bands or not gpu_results:
        return []
    sel = []
    for bi, (name, data, mags) in enumerate(prepare_bands(flux_bands)):
        if bi >= len(gpu_results):
            break
        if len(gpu_results[bi].pso_params) == 0:
            continue
        sel.append((bi, name, data, mags))
    if not sel:
        return []
    models = [int(gpu_results[bi].model) for bi, *_ in sel]
    pso = [list(gpu_results[bi].pso_params) for bi, *_ in sel]
    datas = [d for _, _, d, _ in sel]
    if method == UncertaintyMethod.Svi:
        post = _svi_on_device(ctx, datas, models, pso, SVI_N_STEPS, SVI_N_SAMPLES, SVI_LR, True)
    else:
        data = GpuBatchData(ctx, csr=_fit_csr(datas))
        try:
            r = ctx.laplace_fit(data, models, np.stack([_pad7(p) for p in pso]))
        finally:
            data.close()
        post = [(list(_pad7(pso[k])[:len(pso[k])]), list(r["log_sigma"][k, :len(pso[k])]), float(r["elbo"][k]))
                for k in range(len(sel))]
    jobs = [(name, data, mags, gpu_results[bi], post[k], gpu_results[bi].multi_bazin if bi == 0 else None)
            for k, (bi, name, data, mags) in enumerate(sel)]
    return _results(ctx, jobs, method)


def fit_parametric_batch_gpu(ctx: GpuContext, sources: Sequence[Mapping[str, BandData]], *,
                             n_particles: int = 30, max_iters: int = 60, stall_iters: int = 12,
                             svi_steps: int = PIPELINE_SVI_STEPS, svi_samples: int = PIPELINE_SVI_SAMPLES,
                             svi_lr: float = 0.01) -> list[list[ParametricBandResult]]:
    """The feature extractor's GPU parametric path
    (src/bin/extract_features.rs:585-739) for a list of sources (flux bands
    per source): PSO over all eight models for every band, MultiBazin on each
    source's first band, SVI for every (band, model) pair, then
    finalize_all_models_with_gpu_svi per source, all on the device."""
    from .gpu import BatchSource

    band_map, batch_sources = [], []
    for vi, fb in enumerate(sources):
        for name, data, _m in prepare_bands(fb):
            band_map.append((vi, name))
            batch_sources.append(BatchSource(data.times, data.flux, data.obs_var, data.is_upper.astype(bool),
                                             data.upper_flux))
    per_source: list[list[ParametricBandResult]] = [[] for _ in sources]
    if not batch_sources:
        return per_source
    gdata = GpuBatchData(ctx, batch_sources)
    try:
        all_models = ctx.batch_all_models(gdata, n_particles, max_iters, stall_iters)
        first_bands, seen = [], set()
        for bi, (vi, _) in enumerate(band_map):
            if vi not in seen:
                seen.add(vi)
                first_bands.append(bi)
        mb_data = GpuBatchData(ctx, [batch_sources[bi] for bi in first_bands])
        try:
            mb = ctx.batch_pso_multi_bazin(mb_data, None, n_particles, max_iters, stall_iters, 42)
        finally:
            mb_data.close()
    finally:
        gdata.close()
    nb = len(batch_sources)
    chi2 = [dict() for _ in range(nb)]
    params = [dict() for _ in range(nb)]
    best = [(GpuModelName.Bazin, math.inf) for _ in range(nb)]
    for model, results in all_models:
        for bi, r in enumerate(results):
            chi2[bi][model] = r.cost * 2.0
            params[bi][model] = list(r.params)
            if r.cost < best[bi][1]:
                best[bi] = (model, r.cost)
    mb_of = {bi: mb[k] for k, bi in enumerate(first_bands)}
    entries = [[] for _ in sources]  # (GpuPsoBandResult, band name, svi job index)
    svi_models, svi_pso, svi_bands = [], [], []
    for bi, (vi, name) in enumerate(band_map):
        for model, results in all_models:
            if bi >= len(results):
                continue
            r = results[bi]
            g = GpuPsoBandResult(model=model, pso_params=list(r.params), pso_cost=r.cost,
                                 per_model_chi2=dict(chi2[bi]), per_model_params={},
                                 multi_bazin=mb_of.get(bi) if model == best[bi][0] else None)
            k = None
            if len(r.params):
                k = len(svi_models)
                svi_models.append(int(model))
                svi_pso.append(list(r.params))
                svi_bands.append(bi)
            entries[vi].append((g, name, k))
    fit_data = {}
    for vi, fb in enumerate(sources):
        for name, data, mags in prepare_bands(fb):
            fit_data[(vi, name)] = (data, mags)
    # SVI runs on the extractor's batch sources (its own normalisation)
    svi_data = GpuBatchData(ctx, [batch_sources[bi] for bi in svi_bands]) if svi_models else None
    try:
        svi = (ctx.svi_fit(svi_data, svi_models, np.stack([_pad7(p) for p in svi_pso]),
                           *_priors(svi_models, svi_pso), svi_steps, svi_samples, svi_lr)
               if svi_data is not None else None)
    finally:
        if svi_data is not None:
            svi_data.close()
    jobs, owners = [], []
    for vi in range(len(sources)):
        for g, name, k in entries[vi]:
            if k is None or (vi, name) not in fit_data:
                continue
            np_ = g.model.n_params()
            data, mags = fit_data[(vi, name)]
            post = (list(svi["mu"][k, :np_]), list(svi["log_sigma"][k, :np_]), float(svi["elbo"][k]))
            jobs.append((name, data, mags, g, post, g.multi_bazin))
            owners.append(vi)
    for vi, r in zip(owners, _results(ctx, jobs, UncertaintyMethod.Svi)):
        per_source[vi].append(r)
    return per_source


def _priors(models, pso):
    n = len(models)
    pc, pw = np.zeros((n, MAX_PARAMS)), np.zeros((n, MAX_PARAMS))
    for k in range(n):
        c, w = svi_prior_for_model(models[k], pso[k])
        pc[k, :len(c)], pw[k, :len(w)] = c, w
    return pc, pw


__all__ = ["ALL_GPU_MODELS", "BandFitData", "GpuPsoBandResult", "ParametricBandResult", "SviModelName",
           "UncertaintyMethod", "band_fit_data", "finalize_all_models_with_gpu_svi", "finalize_jobs",
           "finalize_parametric_from_gpu", "finalize_parametric_with_gpu_svi", "fit_parametric_batch_gpu",
           "prepare_bands"]


# ---------------------------------------------------------------------------
# CPU path of the reference (fit_parametric, src/parametric.rs:3563-4131) on
# the device: the swarm fits, L-BFGS polish, model selection, t0 profile
# search, posteriors, finalize and MultiBazin as batched device calls; the
# host builds the jobs and assembles the records.
# ---------------------------------------------------------------------------
from .gpu import (CPSO_POLISH, CPSO_SWARM_BAND, CPSO_SWARM_JOINT, CPSO_SWARM_MULTIBAZIN, FINALIZE_SVI_CHI2,
                  PSO_JOB_DTYPE, _PSO_BOUNDS)

_T0_IDX = (2, 3, 2, 1, 1, 1, 1, 3)
_LN10 = 2.30258509299404568402
BAZIN_GOOD_ENOUGH = 2.0  # src/parametric.rs:1862
# pso_model_select's order after Bazin (src/parametric.rs:2002-2011)
_SELECT_REST = (GpuModelName.Arnett, GpuModelName.Tde, GpuModelName.Afterglow, GpuModelName.Villar,
                GpuModelName.Magnetar, GpuModelName.ShockCooling, GpuModelName.MetzgerKN)
_SEEDS = (42, 137, 271)
# relative_band_priors (src/parametric.rs:3771-3785): mean, sigma per Villar parameter
_REL_MEAN = (-0.015 * _LN10, -0.001, -0.095 * _LN10, -1.181, -0.195 * _LN10, -0.323 * _LN10, -0.025 * _LN10)
_REL_SIG = (0.08 * _LN10, 0.008, 0.25 * _LN10, 1.5, 0.35 * _LN10, 0.25 * _LN10, 0.30 * _LN10)


def population_priors(model: int) -> tuple[list, list]:
    """population_priors (src/parametric.rs:1588-1621): centres, widths."""
    if model == 1:
        return ([-0.009 * _LN10, 0.009, 1.327 * _LN10, -12.0, 0.360 * _LN10, 1.454 * _LN10, -1.033 * _LN10],
                [0.10 * _LN10, 0.010, 0.40 * _LN10, 15.0, 0.35 * _LN10, 0.40 * _LN10, 0.50 * _LN10])
    if model == 0:
        return [0.0, 0.0, 0.0, 1.5, 3.0, -3.0], [2.0, 0.2, 30.0, 2.5, 2.5, 2.0]
    n = GpuModelName(model).n_params()
    return [0.0] * n, [5.0] * n


def pso_bounds_cpu(model: int) -> tuple[list, list]:
    """pso_bounds (src/parametric.rs:1541-1579): the GPU box with the Bazin
    baseline at +-0.3."""
    lo, hi = (list(x) for x in _PSO_BOUNDS[int(model)])
    if model == 0:
        lo[1], hi[1] = -0.3, 0.3
    return lo, hi


def _peak_and_first(data: BandFitData) -> tuple[float, float]:
    """Time of the largest detected flux (first occurrence; times[0] when
    none) and the earliest time."""
    t = data.times
    pt, pf = float(t[0]), -math.inf
    for i in range(t.size):
        if not data.is_upper[i] and data.flux[i] > pf:
            pf, pt = float(data.flux[i]), float(t[i])
    return pt, float(np.min(t))


def _clamp(v, lo, hi):
    return lo if v < lo else (hi if v > hi else v)


def _job(kind, model, band, dim, lo, hi, *, n_bands=1, particles=30, iters=60, stall=12, prior=1, fixed_t0=0,
         t0=0.0, seeded=0, x0=None, seed=42, stream=0, start_cost=0.0):
    """One lcf_pso_job record.  Every job of the CPU path keeps stream 0: the
    reference seeds a fresh SmallRng(seed) for each pso_minimize call, so
    its swarms draw the same numbers for every band, and a band's fit does
    not depend on which batch it is in."""
    j = np.zeros((), dtype=PSO_JOB_DTYPE)
    j["kind"], j["model"], j["band"], j["n_bands"], j["dim"] = kind, model, band, n_bands, dim
    j["n_particles"], j["max_iters"], j["stall_iters"] = particles, iters, stall
    j["use_prior"], j["fixed_t0"], j["seeded"], j["stream"], j["seed"] = prior, fixed_t0, seeded, stream, seed
    j["t0_value"], j["start_cost"] = t0, start_cost
    j["lo"][:dim], j["hi"][:dim] = lo[:dim], hi[:dim]
    if x0 is not None:
        j["x0"][:dim] = x0[:dim]
    return j


def _single_model_setup(data: BandFitData, model: int):
    """pso_fit_single_model's box with the data-informed t0 bounds and its
    two seed particles (src/parametric.rs:1864-1960)."""
    lo, hi = pso_bounds_cpu(model)
    t0i = _T0_IDX[model]
    peak_t, t_first = _peak_and_first(data)
    t0_lo, t0_hi = max(t_first - 30.0, lo[t0i]), min(peak_t + 10.0, hi[t0i])
    if t0_lo < t0_hi:
        lo[t0i], hi[t0i] = t0_lo, t0_hi
    pc, _ = population_priors(model)
    base = [_clamp(pc[d], lo[d], hi[d]) for d in range(len(lo))]
    at_peak, early = list(base), list(base)
    at_peak[t0i] = _clamp(peak_t - 5.0, lo[t0i], hi[t0i])
    early[t0i] = _clamp(t_first - 15.0, lo[t0i], hi[t0i])
    return lo, hi, (None, at_peak, early)


def fit_single_models(ctx: GpuContext, batch: GpuBatchData, datas: Sequence[BandFitData], pairs) -> dict:
    """pso_fit_single_model (src/parametric.rs:1864-1999) for (band, model)
    pairs: three restarts (seeds 42 / 137 / 271, the second seeded at the
    peak, the third before the first observation), best in restart order,
    then lbfgs_refine (not for MetzgerKN).  Returns {(band, model): (params,
    cost)}; params empty when no restart found a finite cost."""
    jobs, meta = [], []
    for b, m in pairs:
        lo, hi, seeds = _single_model_setup(datas[b], m)
        np_ = len(lo)
        for r, sd in enumerate(_SEEDS):
            x0 = seeds[r]
            jobs.append(_job(CPSO_SWARM_BAND, m, b, np_, lo, hi, seeded=int(x0 is not None), x0=x0, seed=sd))
        meta.append((b, m, lo, hi))
    if not jobs:
        return {}
    r = ctx.cpu_pso(batch, np.array(jobs))
    best = {}
    pol, pol_meta = [], []
    for k, (b, m, lo, hi) in enumerate(meta):
        np_ = len(lo)
        bc, bp = math.inf, []
        for rr in range(3):
            c = r["cost"][3 * k + rr]
            if c < bc:
                bc, bp = float(c), list(r["x"][3 * k + rr, :np_])
        best[(b, m)] = (bp, bc)
        if bp and m != int(GpuModelName.MetzgerKN):
            pol.append(_job(CPSO_POLISH, m, b, np_, lo, hi, iters=50, x0=bp, start_cost=bc))
            pol_meta.append((b, m, np_))
    if pol:
        rp = ctx.cpu_pso(batch, np.array(pol))
        for k, (b, m, np_) in enumerate(pol_meta):
            best[(b, m)] = (list(rp["x"][k, :np_]), float(rp["cost"][k]))
    return best


def pso_model_select_batch(ctx: GpuContext, batch: GpuBatchData, datas: Sequence[BandFitData], bands,
                           fit_all_models: bool = False) -> dict:
    """pso_model_select (src/parametric.rs:2001-2066) per band: Bazin first,
    the other seven models unless Bazin's cost is below 2 (or fit_all_models),
    selection by BIC = 2 cost n + k ln n.  Returns {band: (model, params,
    cost, per_model_chi2, per_model_params)}."""
    bz = fit_single_models(ctx, batch, datas, [(b, 0) for b in bands])
    need = [b for b in bands if fit_all_models or not (bz[(b, 0)][1] < BAZIN_GOOD_ENOUGH)]
    rest = fit_single_models(ctx, batch, datas, [(b, int(m)) for b in need for m in _SELECT_REST])
    out = {}
    for b in bands:
        n = max(datas[b].times.size, 1)
        ln_n = math.log(n)
        bic = lambda cost, m: 2.0 * cost * n + GpuModelName(m).n_params() * ln_n
        p0, c0 = bz[(b, 0)]
        chi2 = {GpuModelName.Bazin: finite_or_none(c0)}
        params = {GpuModelName.Bazin: p0} if fit_all_models else {}
        best = (GpuModelName.Bazin, p0, c0, bic(c0, 0))
        if b in need:
            for m in _SELECT_REST:
                pm, cm = rest[(b, int(m))]
                chi2[m] = finite_or_none(cm)
                if fit_all_models:
                    params[m] = pm
                bm = bic(cm, int(m))
                if bm < best[3]:
                    best = (m, pm, cm, bm)
        else:
            for m in _SELECT_REST:
                chi2.setdefault(m, None)
        out[b] = (best[0], best[1], best[2], chi2, params)
    return out


def _linspace(a: float, b: float, n: int) -> list:
    """linspace (src/parametric.rs:1759-1765)."""
    if n <= 1:
        return [a]
    step = (b - a) / (n - 1)
    return [a + i * step for i in range(n)]


def profile_t0_search_batch(ctx: GpuContext, batch: GpuBatchData, datas: Sequence[BandFitData], fits: dict) -> dict:
    """profile_t0_search (src/parametric.rs:1767-1849) for {band: (model,
    params, cost)}: fixed-t0 swarms (30 x 60, stall 12, seed 42, no prior)
    at 5 t0 within +-5 d of the pilot (not before t_first - 10), then 3
    within +-0.5 d of the best; returns {band: (params, cost)}."""
    def run(points):
        jobs, meta = [], []
        for b, t0 in points:
            m, p, _ = fits[b]
            lo, hi = pso_bounds_cpu(m)
            t0i = _T0_IDX[m]
            del lo[t0i], hi[t0i]
            jobs.append(_job(CPSO_SWARM_BAND, m, b, len(lo), lo, hi, prior=0, fixed_t0=1, t0=t0, seed=42))
            meta.append((b, t0, t0i, len(lo)))
        r = ctx.cpu_pso(batch, np.array(jobs)) if jobs else None
        res = []
        for k, (b, t0, t0i, d) in enumerate(meta):
            red = list(r["x"][k, :d])
            res.append((b, t0, red[:t0i] + [t0] + red[t0i:], float(r["cost"][k])))
        return res
    state, rng = {}, {}
    coarse = []
    for b, (m, p, c) in fits.items():
        t0i = _T0_IDX[m]
        pilot = p[t0i]
        t_first = float(np.min(datas[b].times))
        lo, hi = max(pilot - 5.0, t_first - 10.0), pilot + 5.0
        rng[b] = (lo, hi)
        state[b] = [list(p), c, pilot]
        coarse += [(b, t) for t in _linspace(lo, hi, 5)]
    for b, t0, full, cost in run(coarse):
        if cost < state[b][1] and full:
            state[b] = [full, cost, t0]
    fine = []
    for b, (p, c, bt) in state.items():
        lo, hi = rng[b]
        fine += [(b, t) for t in _linspace(max(bt - 0.5, lo), min(bt + 0.5, hi), 3)]
    for b, t0, full, cost in run(fine):
        if cost < state[b][1] and full:
            state[b][0], state[b][1] = full, cost
    return {b: (s[0], s[1]) for b, s in state.items()}


def _posteriors(ctx, datas, models, params, method):
    if method == UncertaintyMethod.Svi:
        return _svi_on_device(ctx, datas, models, params, SVI_N_STEPS, SVI_N_SAMPLES, SVI_LR, True)
    data = GpuBatchData(ctx, csr=_fit_csr(datas))
    try:
        r = ctx.laplace_fit(data, models, np.stack([_pad7(p) for p in params]))
    finally:
        data.close()
    return [(list(params[k]), list(r["log_sigma"][k, :len(params[k])]), float(r["elbo"][k])) for k in range(len(datas))]


# fit_multi_bazin (src/parametric.rs:1008-1157)
MB_MAX_K = 4


def _mb_bounds(k: int, t_min: float, t_max: float) -> tuple[list, list]:
    """multi_bazin_bounds (src/parametric.rs:966-981)."""
    lo, hi = [], []
    for _ in range(k):
        lo += [-3.0, t_min - 30.0, -2.0, -2.0]
        hi += [3.0, t_max + 30.0, 5.0, 6.0]
    return lo + [-0.3, -5.0], hi + [0.3, 0.0]


def _mb_eval(p, k: int, t: float) -> float:
    """multi_bazin_eval (src/parametric.rs:931-940), the reference's
    expression in binary64 (math.exp: the C library's exponential)."""
    s = 0.0
    for c in range(k):
        a, tr, tf = math.exp(p[4 * c]), math.exp(p[4 * c + 2]), math.exp(p[4 * c + 3])
        dt = t - p[4 * c + 1]
        s += a * math.exp(-dt / tf) * (1.0 / (1.0 + math.exp(-dt / tr)))
    return s + p[4 * k]


def _mb_seed(data: BandFitData, prev, k: int, lo, hi) -> list:
    """fit_multi_bazin's K-component start (:1045-1082): the K - 1 components,
    a component at the largest detection residual (first index on ties;
    observation 0 and ln 1e-10 when there is no detection), B and sigma_x of
    the K - 1 fit, clamped to the box."""
    nc = 4 * (k - 1)
    peak_idx, peak_val = 0, -math.inf
    for i in range(data.times.size):
        if data.is_upper[i]:
            continue
        r = float(data.flux[i]) - _mb_eval(prev, k - 1, float(data.times[i]))
        if r > peak_val:
            peak_val, peak_idx = r, i
    init = list(prev[:nc]) + [math.log(max(peak_val, 1e-10)), float(data.times[peak_idx]), 1.0, 1.0,
                              prev[nc], prev[nc + 1]]
    return [_clamp(init[i], lo[i], hi[i]) for i in range(len(init))]


def fit_multi_bazin_batch(ctx: GpuContext, datas: Sequence[BandFitData | None]) -> list:
    """fit_multi_bazin (src/parametric.rs:1008-1157) for a list of bands
    (None entries skipped): for K = 1..4 the swarms run as device jobs
    (30 particles x 60 iterations, stall 12; seed 42, then 137 from the
    residual-seeded start for K >= 2, then 271; the best in that order),
    BIC = 2 cost n + (4 K + 2) ln n, stop once a K's BIC exceeds K - 1's by
    more than 2.  Bands of fewer than 5 points give None, as does a fit
    without parameters."""
    from .gpu import BatchSource

    idx = [k for k, d in enumerate(datas) if d is not None and d.times.size >= 5]
    out = [None] * len(datas)
    if not idx:
        return out
    sub = [datas[k] for k in idx]
    batch = GpuBatchData(ctx, csr=_fit_csr(sub))
    st = [{"best_k": 1, "params": [], "cost": math.inf, "bic": math.inf, "pk_cost": [], "pk_bic": [],
           "prev": None, "done": False} for _ in sub]
    try:
        for k in range(1, MB_MAX_K + 1):
            live = [b for b in range(len(sub)) if not st[b]["done"]]
            if not live:
                break
            jobs, meta = [], []
            for b in live:
                d = sub[b]
                lo, hi = _mb_bounds(k, float(np.min(d.times)), float(np.max(d.times)))
                init = _mb_seed(d, st[b]["prev"], k, lo, hi) if k > 1 and st[b]["prev"] is not None else None
                kw = dict(particles=30, iters=60, stall=12, prior=0)
                r = [_job(CPSO_SWARM_MULTIBAZIN, k, b, 4 * k + 2, lo, hi, seed=42, **kw)]
                if init is not None:
                    r.append(_job(CPSO_SWARM_MULTIBAZIN, k, b, 4 * k + 2, lo, hi, seed=137, seeded=1, x0=init, **kw))
                r.append(_job(CPSO_SWARM_MULTIBAZIN, k, b, 4 * k + 2, lo, hi, seed=271, **kw))
                meta.append((b, len(jobs), len(r)))
                jobs += r
            res = ctx.cpu_pso(batch, np.array(jobs))
            for b, j0, nr in meta:
                np_ = 4 * k + 2
                p1, c1 = list(res["x"][j0, :np_]), float(res["cost"][j0])
                for j in range(j0 + 1, j0 + nr):
                    if res["cost"][j] < c1:
                        p1, c1 = list(res["x"][j, :np_]), float(res["cost"][j])
                s = st[b]
                n = float(sub[b].times.size)
                bic = 2.0 * c1 * n + np_ * math.log(n)
                s["pk_cost"].append(c1)
                s["pk_bic"].append(bic)
                if bic < s["bic"]:
                    s["bic"], s["cost"], s["best_k"], s["params"] = bic, c1, k, p1
                s["prev"] = p1
                if k > 1 and bic > s["pk_bic"][k - 2] + 2.0:
                    s["pk_cost"] += [math.nan] * (MB_MAX_K - k)
                    s["pk_bic"] += [math.nan] * (MB_MAX_K - k)
                    s["done"] = True
    finally:
        batch.close()
    for b, k in enumerate(idx):
        s = st[b]
        if s["params"]:
            out[k] = MultiBazinResult(s["best_k"], s["params"], s["cost"], s["bic"], s["pk_cost"], s["pk_bic"])
    return out


def _multi_bazin_first(ctx, datas_first):
    """MultiBazin on each source's first band: the CPU path's
    fit_multi_bazin."""
    return fit_multi_bazin_batch(ctx, datas_first)


def fit_parametric_model_batch(ctx: GpuContext, sources: Sequence[Mapping[str, BandData]],
                               fit_all_models: bool = False, method: UncertaintyMethod = UncertaintyMethod.Svi,
                               forced_model: GpuModelName | None = None) -> list[list[ParametricBandResult]]:
    """fit_parametric_model (src/parametric.rs:3563-3739) for a list of
    sources: per band model selection (or the forced model), the t0 profile
    search on each source's first band for SVI, the posterior, profile
    refinement, magnitude chi2 and MultiBazin on the first band."""
    prepared = [prepare_bands(fb) for fb in sources]
    flat = [(s, i) for s, pb in enumerate(prepared) for i in range(len(pb))]
    out: list[list[ParametricBandResult]] = [[] for _ in sources]
    if not flat:
        return out
    datas = [prepared[s][i][1] for s, i in flat]
    batch = GpuBatchData(ctx, csr=_fit_csr(datas))
    try:
        bands = list(range(len(flat)))
        if forced_model is not None:
            fm = int(forced_model)
            fits = fit_single_models(ctx, batch, datas, [(b, fm) for b in bands])
            sel = {b: (GpuModelName(fm), fits[(b, fm)][0], fits[(b, fm)][1],
                       {GpuModelName(fm): finite_or_none(fits[(b, fm)][1])},
                       {GpuModelName(fm): fits[(b, fm)][0]} if fit_all_models else {}) for b in bands}
        else:
            sel = pso_model_select_batch(ctx, batch, datas, bands, fit_all_models)
        keep = [b for b in bands if len(sel[b][1]) > 0]
        pso = {b: (list(sel[b][1]), sel[b][2]) for b in keep}
        if method == UncertaintyMethod.Svi:
            first = {b: (int(sel[b][0]), pso[b][0], pso[b][1]) for b in keep if flat[b][1] == 0}
            pso.update(profile_t0_search_batch(ctx, batch, datas, first))
    finally:
        batch.close()
    if not keep:
        return out
    models = [int(sel[b][0]) for b in keep]
    post = _posteriors(ctx, [datas[b] for b in keep], models, [pso[b][0] for b in keep], method)
    mbz = _multi_bazin_first(ctx, [datas[b] if flat[b][1] == 0 else None for b in keep])
    jobs = []
    for k, b in enumerate(keep):
        s, i = flat[b]
        name, data, mags = prepared[s][i]
        g = GpuPsoBandResult(model=sel[b][0], pso_params=pso[b][0], pso_cost=pso[b][1], per_model_chi2=sel[b][3],
                             per_model_params=sel[b][4])
        jobs.append((name, data, mags, g, post[k], mbz[k]))
    for b, r in zip(keep, _results(ctx, jobs, method)):
        out[flat[b][0]].append(r)
    return out


def fit_parametric_multiband_batch(ctx: GpuContext, sources: Sequence[Mapping[str, BandData]],
                                   method: UncertaintyMethod = UncertaintyMethod.Svi) -> list:
    """fit_parametric_multiband (src/parametric.rs:3813-4131) for a list of
    sources: the joint Villar swarm (reference band + offsets, 200 particles x
    1500 iterations, stall 60, three restarts, the second seeded at the prior
    means) over all bands, then per band the posterior, profile refinement
    and magnitude chi2 (from the posterior mean), MultiBazin on the first
    band.  A source yields [] when it has no usable band or more than 9."""
    prepared = [prepare_bands(fb) for fb in sources]
    out: list = [[] for _ in sources]
    srcs = [s for s, pb in enumerate(prepared) if 0 < len(pb) <= 9]
    if not srcs:
        return out
    datas, first = [], {}
    for s in srcs:
        first[s] = len(datas)
        datas += [e[1] for e in prepared[s]]
    batch = GpuBatchData(ctx, csr=_fit_csr(datas))
    jobs, meta = [], []
    pc, _ = population_priors(1)
    for s in srcs:
        B = len(prepared[s])
        lo, hi = pso_bounds_cpu(1)
        for _ in range(1, B):
            lo += [-3.0 * sg for sg in _REL_SIG]
            hi += [3.0 * sg for sg in _REL_SIG]
        peak_t, t_first = _peak_and_first(prepared[s][0][1])
        t0_lo, t0_hi = max(t_first - 30.0, lo[3]), min(peak_t + 10.0, hi[3])
        if t0_lo < t0_hi:  # the reference narrows unconditionally and would panic on an empty box
            lo[3], hi[3] = t0_lo, t0_hi
        x0 = [_clamp(pc[j], lo[j], hi[j]) for j in range(7)]
        x0[3] = _clamp(peak_t - 5.0, lo[3], hi[3])
        for b in range(1, B):
            x0 += [_clamp(_REL_MEAN[j], lo[7 * b + j], hi[7 * b + j]) for j in range(7)]
        for r, sd in enumerate(_SEEDS):
            jobs.append(_job(CPSO_SWARM_JOINT, 1, first[s], 7 * B, lo, hi, n_bands=B, particles=200, iters=1500,
                             stall=60, seeded=int(r == 1), x0=x0, seed=sd))
        meta.append((s, B))
    try:
        r = ctx.cpu_pso(batch, np.array(jobs))
    finally:
        batch.close()
    band_jobs = []
    for k, (s, B) in enumerate(meta):
        bc, bp = math.inf, None
        for rr in range(3):
            c = r["cost"][3 * k + rr]
            if c < bc:
                bc, bp = float(c), r["x"][3 * k + rr, :7 * B]
        if bp is None:
            continue
        ref = list(bp[:7])
        for b in range(B):
            params = ref if b == 0 else [ref[j] + bp[7 * b + j] for j in range(7)]
            band_jobs.append((s, b, params, bc))
    if not band_jobs:
        return out
    bdatas = [prepared[s][b][1] for s, b, _, _ in band_jobs]
    post = _posteriors(ctx, bdatas, [1] * len(band_jobs), [p for _, _, p, _ in band_jobs], method)
    mbz = _multi_bazin_first(ctx, [prepared[s][b][1] if b == 0 else None for s, b, _, _ in band_jobs])
    fin = finalize_jobs(ctx, bdatas, [prepared[s][b][2] for s, b, _, _ in band_jobs], [1] * len(band_jobs),
                        np.stack([_pad7(p) for _, _, p, _ in band_jobs]),
                        np.stack([_pad7(q[0]) for q in post]), np.stack([_pad7(q[1]) for q in post]),
                        flags=FINALIZE_DEFAULT | FINALIZE_SVI_CHI2)
    for k, (s, b, params, bc) in enumerate(band_jobs):
        name, data, _ = prepared[s][b]
        out[s].append(ParametricBandResult(
            band=name, model=GpuModelName.Villar, pso_params=list(params), pso_chi2=finite_or_none(bc),
            svi_mu=[float(x) for x in fin["mu"][k, :7]], svi_log_sigma=[float(x) for x in fin["log_sigma"][k, :7]],
            svi_elbo=finite_or_none(post[k][2]), n_obs=int(data.times.size),
            mag_chi2=finite_or_none(fin["mag_chi2"][k]), per_model_chi2={GpuModelName.Villar: finite_or_none(bc)},
            per_model_params={}, uncertainty_method=method, multi_bazin=mbz[k]))
    return out


def fit_parametric_batch(ctx: GpuContext, sources: Sequence[Mapping[str, BandData]], fit_all_models: bool = False,
                         method: UncertaintyMethod = UncertaintyMethod.Svi) -> list[list[ParametricBandResult]]:
    """fit_parametric (src/parametric.rs:3742-3761) for a list of sources:
    the joint multi-band Villar fit for sources with at least two non-empty
    bands (unless fit_all_models), the per-band path otherwise or when the
    joint fit yields nothing."""
    multi = [s for s, fb in enumerate(sources)
             if not fit_all_models and sum(1 for b in fb.values() if len(b.values) > 0) >= 2]
    out: list = [[] for _ in sources]
    if multi:
        mr = fit_parametric_multiband_batch(ctx, [sources[s] for s in multi], method)
        for s, r in zip(multi, mr):
            out[s] = r
    rest = [s for s in range(len(sources)) if not out[s]]
    if rest:
        pr = fit_parametric_model_batch(ctx, [sources[s] for s in rest], fit_all_models, method)
        for s, r in zip(rest, pr):
            out[s] = r
    return out
