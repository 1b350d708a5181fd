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

from .gpu import (ALL_GPU_MODELS, FINALIZE_DEFAULT, MAX_PARAMS, GpuBatchData, GpuContext, GpuModelName,
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
_SVI_CPU_STOPPING, _SVI_INFLATE_SIGMA, _SVI_FD_GRADIENT = 1, 2, 4


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
    apply, and MetzgerKN (no analytic gradient) is differentiated
    numerically as the CPU fit does."""
    mids = np.asarray(models, dtype=np.int32)
    flags = (_SVI_CPU_STOPPING | _SVI_INFLATE_SIGMA) if cpu_semantics else 0
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
