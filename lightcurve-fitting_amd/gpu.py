"""Host-side mirror of the reference's GPU batch interface.

Same names, argument meaning and error behaviour as the reference's
``src/gpu_types.rs`` (types) and ``src/gpu_cuda.rs`` (``GpuContext``,
``GpuBatchData``); every call goes through the C ABI of liblcf_hip.so.
Errors surface as ``GpuError`` carrying the library's message, the analogue
of the reference's ``Err(String)``.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib

MAX_PARAMS = 7
NP_NFEAT = 50  # NonparametricBandResult fields without the band name
FP32, FP64 = 0, 1
SCHED_PERSISTENT, SCHED_PER_CLASS, SCHED_TUNED, SCHED_PER_MODEL = 0, 1, 2, 3
TUNE_BUCKETS, TUNE_FIELDS = 7, 6
NOALLOC_MAX_BANDS = 16384  # LCF_NOALLOC_MAX_BANDS
PSO_STALL_FROM_FIRST_EVAL = 1
FINALIZE_T0, FINALIZE_ALL, FINALIZE_MAG_CHI2 = 1, 2, 4
FINALIZE_DEFAULT = FINALIZE_T0 | FINALIZE_ALL | FINALIZE_MAG_CHI2
FINALIZE_SVI_CHI2 = 8
CPSO_SWARM_BAND, CPSO_SWARM_JOINT, CPSO_POLISH, CPSO_SWARM_MULTIBAZIN = 0, 1, 2, 3
# struct lcf_pso_job (include/lcf_hip.h)
PSO_JOB_DTYPE = np.dtype([("kind", "<i4"), ("model", "<i4"), ("band", "<i4"), ("n_bands", "<i4"), ("dim", "<i4"),
                          ("n_particles", "<i4"), ("max_iters", "<i4"), ("stall_iters", "<i4"),
                          ("use_prior", "<i4"), ("fixed_t0", "<i4"), ("seeded", "<i4"), ("stream", "<u4"),
                          ("seed", "<u8"), ("t0_value", "<f8"), ("start_cost", "<f8"), ("lo", "<f8", 64),
                          ("hi", "<f8", 64), ("x0", "<f8", 64)], align=True)


class GpuError(RuntimeError):
    """Failure reported by the native library (reference: ``Err(String)``)."""


class GpuModelName(enum.IntEnum):
    """Model ids of ``GpuModelName::model_id`` (src/gpu_types.rs:51-62)."""

    Bazin = 0
    Villar = 1
    Tde = 2
    Arnett = 3
    Magnetar = 4
    ShockCooling = 5
    Afterglow = 6
    MetzgerKN = 7

    def n_params(self) -> int:
        """``GpuModelName::n_params`` (src/gpu_types.rs:23-34)."""
        return (6, 7, 7, 5, 5, 5, 6, 5)[int(self)]

    def model_id(self) -> int:
        return int(self)

    def pso_bounds(self) -> tuple[list[float], list[float]]:
        """``GpuModelName::pso_bounds`` (src/gpu_types.rs:64-103)."""
        return [list(x) for x in _PSO_BOUNDS[int(self)]]


_PSO_BOUNDS = (
    ((-3.0, -1.0, -100.0, -2.0, -2.0, -5.0), (3.0, 1.0, 100.0, 5.0, 6.0, 0.0)),
    ((-1.5, -0.02, 1.0, -60.0, -1.0, 1.0, -6.0), (1.5, 0.05, 5.0, 40.0, 3.0, 5.5, 0.0)),
    ((-3.0, -1.0, -100.0, -2.0, -1.0, 0.5, -5.0), (3.0, 1.0, 100.0, 5.0, 6.0, 4.0, 0.0)),
    ((-3.0, -100.0, 0.5, -3.0, -5.0), (3.0, 100.0, 4.5, 3.0, 0.0)),
    ((-3.0, -100.0, 0.0, 0.5, -5.0), (3.0, 100.0, 6.0, 4.5, 0.0)),
    ((-3.0, -100.0, 0.1, -1.0, -5.0), (3.0, 100.0, 3.0, 4.0, 0.0)),
    ((-3.0, -100.0, -2.0, -2.0, 0.5, -5.0), (3.0, 100.0, 6.0, 3.0, 5.0, 0.0)),
    ((-3.0, -2.0, -1.0, -2.0, -5.0), (-0.5, -0.5, 2.0, 1.0, 0.0)),
)

# ALL_GPU_MODELS (src/gpu_types.rs:107-116)
ALL_GPU_MODELS = (
    GpuModelName.Bazin, GpuModelName.Arnett, GpuModelName.Tde, GpuModelName.Afterglow,
    GpuModelName.Villar, GpuModelName.Magnetar, GpuModelName.ShockCooling, GpuModelName.MetzgerKN,
)


@dataclass
class BatchSource:
    """Pre-normalised observations of one source/band (src/gpu_types.rs:122-130)."""

    times: Sequence[float]
    flux: Sequence[float]
    obs_var: Sequence[float]
    is_upper: Sequence[bool]
    upper_flux: Sequence[float]


@dataclass
class BatchPsoResult:
    """Best position and cost of one swarm (src/gpu_types.rs:133-137)."""

    params: list = field(default_factory=list)
    cost: float = float("inf")


@dataclass
class MultiBazinResult:
    """Greedy MultiBazin fit of one source (src/parametric.rs:1269-1283)."""

    best_k: int
    params: list
    cost: float
    bic: float
    per_k_cost: list
    per_k_bic: list


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class GpuContext:
    """``GpuContext`` (src/gpu_cuda.rs:297-1423) on one HIP device."""

    def __init__(self, device: int = 0, precision: int = FP32, schedule: int = SCHED_PERSISTENT):
        self._lib = _lib.load()
        h = C.c_void_p()
        rc = self._lib.lcf_ctx_create(int(device), C.byref(h))
        if rc != 0 or not h.value:
            raise GpuError(f"HIP device {device} unavailable (status {rc})")
        self._h = h
        self.device = device
        self.set_precision(precision)
        self.set_schedule(schedule)

    @classmethod
    def new(cls, device: int = 0) -> "GpuContext":
        return cls(device)

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self._lib.lcf_last_error(self._h)
            raise GpuError((msg or b"").decode() or f"status {rc}")

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lcf_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: int) -> None:
        self._check(self._lib.lcf_ctx_set_precision(self._h, int(precision)))
        self.precision = int(precision)

    def set_schedule(self, schedule: int) -> None:
        self._check(self._lib.lcf_ctx_set_schedule(self._h, int(schedule)))
        self.schedule = int(schedule)

    # -- autotuned launch shapes (lcf_ctx_autotune_pso) ---------------------------
    def autotune_pso(self, data: "GpuBatchData", models: Sequence[int], n_particles: int, max_iters: int,
                     stall_iters: int, seed: int = 42) -> list[dict]:
        """Measure the PSO swarms' launch shape (LDS tile, waves per CU) per
        band-size bucket of ``data`` for ``models`` at this context's
        precision; returns the resulting table (``tuning()``).  Fits use it
        under ``SCHED_TUNED``."""
        ms = np.ascontiguousarray(np.asarray([int(m) for m in models], dtype=np.int32))
        self._check(self._lib.lcf_ctx_autotune_pso(self._h, data._h, _lib.ptr(ms, C.c_int32), ms.size,
                                                   int(n_particles), int(max_iters), int(stall_iters),
                                                   C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF)))
        return self.tuning()

    def tuning(self) -> list[dict]:
        """The context's tuning table, one dict per (precision, model,
        particles per lane, bucket): its tile and waves per CU."""
        n = self._lib.lcf_ctx_get_tuning(self._h, None, 0)
        if n < 0:
            self._check(n)
        buf = np.zeros((max(n, 1), TUNE_FIELDS), dtype=np.int32)
        self._lib.lcf_ctx_get_tuning(self._h, _lib.ptr(buf, C.c_int32), n)
        keys = ("precision", "model", "ppl", "bucket", "tile", "waves_per_cu")
        return [dict(zip(keys, map(int, row))) for row in buf[:n]]

    def set_tuning(self, table: Sequence[dict]) -> None:
        """Replace the tuning table (e.g. one saved as JSON from ``tuning()``)."""
        keys = ("precision", "model", "ppl", "bucket", "tile", "waves_per_cu")
        buf = np.ascontiguousarray(np.asarray([[int(e[k]) for k in keys] for e in table], dtype=np.int32)
                                   .reshape(-1, TUNE_FIELDS))
        self._check(self._lib.lcf_ctx_set_tuning(self._h, _lib.ptr(buf, C.c_int32), len(table)))

    @property
    def num_cus(self) -> int:
        return int(self._lib.lcf_ctx_num_cus(self._h))

    def synchronize(self) -> None:
        """Wait for all work on the context's device (``lcf_ctx_synchronize``)."""
        self._check(self._lib.lcf_ctx_synchronize(self._h))

    @property
    def last_kernel_ms(self) -> float:
        return float(self._lib.lcf_last_kernel_ms(self._h))

    def last_model_ms(self, slot: int) -> float:
        """Device time of model slot `slot` in the last batched fit."""
        return float(self._lib.lcf_last_model_ms(self._h, int(slot)))

    # -- GpuContext::eval_batch (src/gpu_cuda.rs:317-360) --------------------
    def eval_batch(self, model: GpuModelName, params, times, n_draws: int) -> np.ndarray:
        model = GpuModelName(model)
        p, t = _f64(params).ravel(), _f64(times).ravel()
        npar = model.n_params()
        if p.size != n_draws * npar:
            raise GpuError(f"params length {p.size} != n_draws({n_draws}) x n_params({npar})")
        out = np.empty(n_draws * t.size, dtype=np.float64)
        self._check(self._lib.lcf_eval_batch(self._h, int(model), _lib.ptr(p, C.c_double), n_draws,
                                             _lib.ptr(t, C.c_double), t.size, _lib.ptr(out, C.c_double)))
        return out

    # -- batched PSO ----------------------------------------------------------
    def fit_models(self, data: "GpuBatchData", models: Sequence[int], n_particles: int,
                   max_iters: int, stall_iters: int, *, seed: int = 42, gate: float | None = None,
                   flags: int = 0, per_model: bool = True, noalloc: bool = False) -> dict:
        """Outputs of ``lcf_batch_fit_models`` as numpy arrays (per-model
        arrays are None when ``per_model`` is false); ``noalloc`` calls
        ``lcf_batch_fit_models_noalloc`` (fixed per-device scratch, at most
        ``NOALLOC_MAX_BANDS`` sources)."""
        ms = np.ascontiguousarray(np.asarray([int(m) for m in models], dtype=np.int32))
        ns, nm = data.n_sources, ms.size
        out = {
            "model_params": np.zeros((nm, ns, MAX_PARAMS)) if per_model else None,
            "model_cost": np.zeros((nm, ns)) if per_model else None,
            "model_iters": np.zeros((nm, ns), dtype=np.int32) if per_model else None,
            "best_model": np.zeros(ns, dtype=np.int32),
            "best_params": np.zeros((ns, MAX_PARAMS)),
            "best_cost": np.zeros(ns),
        }
        fn = self._lib.lcf_batch_fit_models_noalloc if noalloc else self._lib.lcf_batch_fit_models
        rc = fn(
            self._h, data._h, _lib.ptr(ms, C.c_int32), nm, 0 if gate is None else 1,
            0.0 if gate is None else float(gate), n_particles, max_iters, stall_iters,
            C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), flags,
            _lib.ptr(out["model_params"], C.c_double), _lib.ptr(out["model_cost"], C.c_double),
            _lib.ptr(out["model_iters"], C.c_int32), _lib.ptr(out["best_model"], C.c_int32),
            _lib.ptr(out["best_params"], C.c_double), _lib.ptr(out["best_cost"], C.c_double))
        self._check(rc)
        return out

    def batch_pso(self, model: GpuModelName, data: "GpuBatchData", n_particles: int,
                  max_iters: int, stall_iters: int, seed: int) -> list[BatchPsoResult]:
        """``GpuContext::batch_pso`` (src/gpu_cuda.rs:368-567)."""
        model = GpuModelName(model)
        ns = data.n_sources
        prm = np.zeros((ns, MAX_PARAMS))
        cost = np.zeros(ns)
        self._check(self._lib.lcf_batch_pso(self._h, int(model), data._h, n_particles, max_iters,
                                            stall_iters, C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF),
                                            _lib.ptr(prm, C.c_double), _lib.ptr(cost, C.c_double)))
        npar = model.n_params()
        return [BatchPsoResult(list(prm[s, :npar]), float(cost[s])) for s in range(ns)]

    def batch_model_select(self, data: "GpuBatchData", n_particles: int, max_iters: int,
                           stall_iters: int, bazin_good_enough: float):
        """``GpuContext::batch_model_select`` (src/gpu_cuda.rs:573-620)."""
        ns = data.n_sources
        mdl = np.zeros(ns, dtype=np.int32)
        prm = np.zeros((ns, MAX_PARAMS))
        cost = np.zeros(ns)
        self._check(self._lib.lcf_batch_model_select(
            self._h, data._h, n_particles, max_iters, stall_iters, float(bazin_good_enough),
            _lib.ptr(mdl, C.c_int32), _lib.ptr(prm, C.c_double), _lib.ptr(cost, C.c_double)))
        res = []
        for s in range(ns):
            m = GpuModelName(int(mdl[s]))
            res.append((m, BatchPsoResult(list(prm[s, :m.n_params()]), float(cost[s]))))
        return res

    def batch_all_models(self, data: "GpuBatchData", n_particles: int, max_iters: int,
                         stall_iters: int):
        """``GpuContext::batch_all_models`` (src/gpu_cuda.rs:624-637)."""
        ns = data.n_sources
        prm = np.zeros((len(ALL_GPU_MODELS), ns, MAX_PARAMS))
        cost = np.zeros((len(ALL_GPU_MODELS), ns))
        self._check(self._lib.lcf_batch_all_models(self._h, data._h, n_particles, max_iters,
                                                   stall_iters, _lib.ptr(prm, C.c_double),
                                                   _lib.ptr(cost, C.c_double)))
        out = []
        for k, m in enumerate(ALL_GPU_MODELS):
            out.append((m, [BatchPsoResult(list(prm[k, s, :m.n_params()]), float(cost[k, s]))
                            for s in range(ns)]))
        return out

    # -- greedy MultiBazin -------------------------------------------------------
    def multi_bazin_arrays(self, data: "GpuBatchData", n_particles: int, max_iters: int, stall_iters: int,
                           seed: int) -> dict:
        """Array form of ``lcf_batch_multi_bazin``: best_k [n], params [n, 18],
        cost, bic [n], per_k_cost / per_k_bic [n, 4]."""
        ns = data.n_sources
        out = {"best_k": np.zeros(ns, dtype=np.int32), "params": np.zeros((ns, 18)), "cost": np.zeros(ns),
               "bic": np.zeros(ns), "per_k_cost": np.zeros((ns, 4)), "per_k_bic": np.zeros((ns, 4))}
        self._check(self._lib.lcf_batch_multi_bazin(
            self._h, data._h, int(n_particles), int(max_iters), int(stall_iters),
            C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), _lib.ptr(out["best_k"], C.c_int32),
            _lib.ptr(out["params"], C.c_double), _lib.ptr(out["cost"], C.c_double), _lib.ptr(out["bic"], C.c_double),
            _lib.ptr(out["per_k_cost"], C.c_double), _lib.ptr(out["per_k_bic"], C.c_double)))
        return out

    def batch_pso_multi_bazin(self, data: "GpuBatchData", sources: Sequence["BatchSource"] | None,
                              n_particles: int, max_iters: int, stall_iters: int,
                              seed: int) -> list["MultiBazinResult"]:
        """``GpuContext::batch_pso_multi_bazin`` (src/gpu_cuda.rs:650-963).  The
        reference also takes the sources on the host for their time ranges
        and residuals; here those come from ``data`` on the device, and
        ``sources`` is only checked for length."""
        if sources is not None and len(sources) != data.n_sources:
            raise GpuError(f"{len(sources)} sources for a batch of {data.n_sources}")
        r = self.multi_bazin_arrays(data, n_particles, max_iters, stall_iters, seed)
        out = []
        for s in range(data.n_sources):
            k = int(r["best_k"][s])
            prm = list(r["params"][s, :4 * k + 2]) if np.isfinite(r["bic"][s]) or r["bic"][s] < 0 else []
            out.append(MultiBazinResult(k, prm, float(r["cost"][s]), float(r["bic"][s]),
                                        list(r["per_k_cost"][s]), list(r["per_k_bic"][s])))
        return out

    # -- batched SVI -------------------------------------------------------------
    def batch_svi_fit(self, data: "GpuBatchData", inputs: Sequence["SviBatchInput"], n_steps: int,
                      n_samples: int, lr: float) -> list["SviBatchOutput"]:
        """``GpuContext::batch_svi_fit`` (src/gpu_cuda.rs:1315-1422)."""
        nj = len(inputs)
        if nj == 0:
            return []
        if nj != data.n_sources:
            raise GpuError(f"SVI inputs ({nj}) must match batch data size ({data.n_sources})")
        mid = np.zeros(nj, dtype=np.int32)
        se = np.zeros(nj, dtype=np.int32)
        pso, pc, pw = (np.zeros((nj, MAX_PARAMS)) for _ in range(3))
        for i, inp in enumerate(inputs):
            k = min(len(inp.pso_params), MAX_PARAMS)
            pso[i, :k] = inp.pso_params[:k]
            pc[i, :k] = inp.prior_centers[:k]
            pw[i, :k] = inp.prior_widths[:k]
            mid[i] = inp.model_id
            se[i] = inp.se_idx
        mu, ls, elbo = np.zeros((nj, MAX_PARAMS)), np.zeros((nj, MAX_PARAMS)), np.zeros(nj)
        self._check(self._lib.lcf_batch_svi_fit(
            self._h, data._h, _lib.ptr(mid, C.c_int32), _lib.ptr(pso, C.c_double), _lib.ptr(se, C.c_int32),
            _lib.ptr(pc, C.c_double), _lib.ptr(pw, C.c_double), n_steps, n_samples, float(lr),
            _lib.ptr(mu, C.c_double), _lib.ptr(ls, C.c_double), _lib.ptr(elbo, C.c_double)))
        out = []
        for i, inp in enumerate(inputs):
            k = min(len(inp.pso_params), MAX_PARAMS)
            out.append(SviBatchOutput(list(mu[i, :k]), list(ls[i, :k]), float(elbo[i])))
        return out

    def svi_fit(self, data: "GpuBatchData", model_ids, pso, prior_c, prior_w, n_steps: int,
                n_samples: int, lr: float, *, seed: int = 42, flags: int = 0) -> dict:
        """Array form of the batched SVI (``lcf_batch_svi_fit_ex``)."""
        nj = data.n_sources
        mid = np.ascontiguousarray(np.asarray(model_ids, dtype=np.int32))
        pso, pc, pw = (_rows7(x, nj) for x in (pso, prior_c, prior_w))
        mu, ls, elbo = np.zeros((nj, MAX_PARAMS)), np.zeros((nj, MAX_PARAMS)), np.zeros(nj)
        steps = np.zeros(nj, dtype=np.int32)
        self._check(self._lib.lcf_batch_svi_fit_ex(
            self._h, data._h, _lib.ptr(mid, C.c_int32), _lib.ptr(pso, C.c_double), _lib.ptr(pc, C.c_double),
            _lib.ptr(pw, C.c_double), n_steps, n_samples, float(lr), C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF),
            flags, _lib.ptr(mu, C.c_double), _lib.ptr(ls, C.c_double), _lib.ptr(elbo, C.c_double),
            _lib.ptr(steps, C.c_int32)))
        return {"mu": mu, "log_sigma": ls, "elbo": elbo, "steps": steps}

    def svi_elbo_grad(self, data: "GpuBatchData", model_ids, mu, log_sigma, prior_c, prior_w,
                      n_samples: int, *, seed: int = 42, step: int = 0, flags: int = 0):
        """One Monte-Carlo ELBO estimate per job and its gradient w.r.t.
        (mu, log_sigma), from the adjoint kernel (``lcf_svi_elbo_grad``)."""
        nj = data.n_sources
        mid = np.ascontiguousarray(np.asarray(model_ids, dtype=np.int32))
        mu, ls, pc, pw = (_rows7(x, nj) for x in (mu, log_sigma, prior_c, prior_w))
        elbo, gmu, gls = np.zeros(nj), np.zeros((nj, MAX_PARAMS)), np.zeros((nj, MAX_PARAMS))
        self._check(self._lib.lcf_svi_elbo_grad(
            self._h, data._h, _lib.ptr(mid, C.c_int32), _lib.ptr(mu, C.c_double), _lib.ptr(ls, C.c_double),
            _lib.ptr(pc, C.c_double), _lib.ptr(pw, C.c_double), n_samples,
            C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), step, flags, _lib.ptr(elbo, C.c_double),
            _lib.ptr(gmu, C.c_double), _lib.ptr(gls, C.c_double)))
        return elbo, gmu, gls


    # -- batched Laplace ------------------------------------------------------
    def laplace_fit(self, data: "GpuBatchData", model_ids, map_params, *, mode: int = 0) -> dict:
        """Batched ``laplace_fit`` (src/parametric.rs:2623-2693), FP64: one job
        per source of ``data``.  mode 0 = the reference's 4-point difference
        Hessian, 1 = central differences of the analytic gradient."""
        nj = data.n_sources
        mid = np.ascontiguousarray(np.asarray(model_ids, dtype=np.int32))
        mp = _rows7(map_params, nj)
        ls, elbo, hess = np.zeros((nj, MAX_PARAMS)), np.zeros(nj), np.zeros((nj, 36))
        self._check(self._lib.lcf_batch_laplace(self._h, data._h, _lib.ptr(mid, C.c_int32),
                                                _lib.ptr(mp, C.c_double), int(mode), _lib.ptr(ls, C.c_double),
                                                _lib.ptr(elbo, C.c_double), _lib.ptr(hess, C.c_double)))
        return {"mu": mp, "log_sigma": ls, "elbo": elbo, "hessian": hess}

    # -- CPU-path swarm fits -------------------------------------------------
    def cpu_pso(self, data: "GpuBatchData", jobs) -> dict:
        """``lcf_cpu_pso``: swarm / polish jobs (PSO_JOB_DTYPE records) on the
        bands of ``data``; returns x [n][64], cost, iters."""
        jobs = np.ascontiguousarray(jobs, dtype=PSO_JOB_DTYPE)
        nj = jobs.size
        x, cost, it = np.zeros((nj, 64)), np.zeros(nj), np.zeros(nj, np.int32)
        if nj == 0:
            return {"x": x, "cost": cost, "iters": it}
        self._check(self._lib.lcf_cpu_pso(self._h, data._h, jobs.ctypes.data_as(C.c_void_p), nj,
                                          _lib.ptr(x, C.c_double), _lib.ptr(cost, C.c_double),
                                          _lib.ptr(it, C.c_int32)))
        return {"x": x, "cost": cost, "iters": it}

    # -- batched finalize -----------------------------------------------------
    def finalize_arrays(self, data: "GpuBatchData", model_ids, pso_params, mu, log_sigma, flux, flux_err,
                        mag_obs, peak_flux, *, flags: int = FINALIZE_DEFAULT) -> dict:
        """Profile refinement and magnitude chi2 of every band of ``data``
        (profile_t0_refine, profile_refine_all and the chi2 of
        finalize_parametric_with_gpu_svi, src/parametric.rs:2705-3008,
        3366-3393), FP64.  ``flux`` / ``flux_err`` / ``mag_obs`` are per
        observation in the batch's order, ``peak_flux`` per band; returns the
        refined ``mu`` and ``log_sigma`` ([n][7]) and ``mag_chi2`` ([n], NaN
        when no term qualifies)."""
        nj = data.n_sources
        mid = np.ascontiguousarray(np.asarray(model_ids, dtype=np.int32))
        if mid.size != nj:
            raise GpuError(f"model_ids has {mid.size} entries for {nj} bands")
        pso, m0, l0 = _rows7(pso_params, nj), _rows7(mu, nj), _rows7(log_sigma, nj)
        per_obs = [_f64(a) for a in (flux, flux_err, mag_obs)]
        for name, a in zip(("flux", "flux_err", "mag_obs"), per_obs):
            if a.size != data.n_obs:
                raise GpuError(f"{name} has {a.size} entries for {data.n_obs} observations")
        pk = _f64(peak_flux)
        if pk.size != nj:
            raise GpuError(f"peak_flux has {pk.size} entries for {nj} bands")
        om, ol, chi = np.zeros((nj, MAX_PARAMS)), np.zeros((nj, MAX_PARAMS)), np.zeros(nj)
        d = lambda a: _lib.ptr(a, C.c_double)
        self._check(self._lib.lcf_batch_finalize(self._h, data._h, _lib.ptr(mid, C.c_int32), d(pso), d(m0), d(l0),
                                                 d(per_obs[0]), d(per_obs[1]), d(per_obs[2]), d(pk), int(flags),
                                                 d(om), d(ol), d(chi)))
        return {"mu": om, "log_sigma": ol, "mag_chi2": chi}

    # -- batched dense GP -----------------------------------------------------
    def batch_gp_fit(self, bands: Sequence["GpBandInput"], query_times, amp_candidates, ls_candidates,
                     max_subsample: int) -> list["GpBandOutput"]:
        """``GpuContext::batch_gp_fit`` (src/gpu_cuda.rs:976-1143).

        The reference rebuilds a CPU ``DenseGP`` per band from the returned
        state; here ``dense_gp`` is a ``DenseGpState`` carrying the same fit
        (training set, alpha, amplitude, 1/(2 l^2), mean) from the device."""
        csr = _gp_csr(bands)
        q = _f64(query_times)
        a, l = _f64(amp_candidates), _f64(ls_candidates)
        nb = len(bands)
        if nb == 0:
            return []
        state = np.zeros((nb, 54))
        pred, std = np.zeros((nb, q.size)), np.zeros((nb, q.size))
        pobs = np.zeros(csr["offsets"][-1])
        self._check(self._lib.lcf_batch_gp_fit(
            self._h, nb, _lib.ptr(csr["offsets"], C.c_int64), _lib.ptr(csr["times"], C.c_double),
            _lib.ptr(csr["mags"], C.c_double), _lib.ptr(csr["noise_var"], C.c_double), _lib.ptr(q, C.c_double),
            q.size, _lib.ptr(a, C.c_double), a.size, _lib.ptr(l, C.c_double), l.size, int(max_subsample),
            _lib.ptr(state, C.c_double), _lib.ptr(pred, C.c_double), _lib.ptr(std, C.c_double),
            _lib.ptr(pobs, C.c_double)))
        off = csr["offsets"]
        out = []
        for b in range(nb):
            m = int(state[b, 53])
            gp = DenseGpState(state[b, 25:25 + m].copy(), state[b, :m].copy(), float(state[b, 50]),
                              float(state[b, 51]), float(state[b, 52])) if 0 < m <= 25 else None
            out.append(GpBandOutput(gp, list(pred[b]), list(std[b]), list(pobs[off[b]:off[b + 1]])))
        return out

    def gp_fit(self, csr: dict, query, amp_candidates, ls_candidates, *, cap: int = 128,
               subsample_rule: int = 0, criterion: int = 0, query_per_band: bool = False) -> dict:
        """Array form (``lcf_batch_gp_fit_ex``): training-set cap up to 200,
        subsample rule (0 GPU, 1 CPU), criterion (0 NLML, 1 train RMS),
        shared or per-band query grid; returns state, pred, std, dpred, pred_obs."""
        off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
        nb = off.size - 1
        t, y, nv = _f64(csr["times"]), _f64(csr["mags"]), _f64(csr["noise_var"])
        q = _f64(query)
        nq = q.shape[-1] if q.ndim else 0
        a, l = _f64(amp_candidates), _f64(ls_candidates)
        stride = 2 * cap + 4
        out = {"state": np.zeros((nb, stride)), "pred": np.zeros((nb, nq)), "std": np.zeros((nb, nq)),
               "dpred": np.zeros((nb, nq)), "pred_obs": np.zeros(t.size)}
        self._check(self._lib.lcf_batch_gp_fit_ex(
            self._h, nb, _lib.ptr(off, C.c_int64), _lib.ptr(t, C.c_double), _lib.ptr(y, C.c_double),
            _lib.ptr(nv, C.c_double), _lib.ptr(q, C.c_double), nq, int(query_per_band), _lib.ptr(a, C.c_double),
            a.size, _lib.ptr(l, C.c_double), l.size, cap, subsample_rule, criterion,
            _lib.ptr(out["state"], C.c_double), _lib.ptr(out["pred"], C.c_double),
            _lib.ptr(out["std"], C.c_double), _lib.ptr(out["dpred"], C.c_double),
            _lib.ptr(out["pred_obs"], C.c_double)))
        return out


    def nonparametric_arrays(self, csr: dict, t_lo, t_hi, *, n_grid: int = 50,
                             amp_candidates=(0.3, 1.0, 3.0, 10.0), ls_factors=(3.0, 6.0, 12.0, 24.0, 48.0),
                             dense_max: int = 100, max_subsample: int = 25, n_inducing: int = 30) -> dict:
        """Array form of the nonparametric band pipeline
        (``lcf_batch_nonparametric``): CSR bands of raw magnitudes (offsets,
        times, mags, mag_errs), per-band grid range; returns status, features
        [n_bands, 50], pred / std / dpred [n_bands, n_grid] and state."""
        off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
        nb = off.size - 1
        t, m, e = _f64(csr["times"]), _f64(csr["mags"]), _f64(csr["mag_errs"])
        lo, hi = _f64(t_lo), _f64(t_hi)
        if lo.size != nb or hi.size != nb:
            raise GpuError("t_lo / t_hi need one value per band")
        a, lf = _f64(amp_candidates), _f64(ls_factors)
        cap = max(dense_max, max_subsample)
        out = {"status": np.zeros(nb, np.int32), "features": np.zeros((nb, NP_NFEAT)),
               "pred": np.zeros((nb, n_grid)), "std": np.zeros((nb, n_grid)), "dpred": np.zeros((nb, n_grid)),
               "state": np.zeros((nb, 2 * cap + 4))}
        self._check(self._lib.lcf_batch_nonparametric(
            self._h, nb, _lib.ptr(off, C.c_int64), _lib.ptr(t, C.c_double), _lib.ptr(m, C.c_double),
            _lib.ptr(e, C.c_double), _lib.ptr(lo, C.c_double), _lib.ptr(hi, C.c_double), n_grid,
            _lib.ptr(a, C.c_double), a.size, _lib.ptr(lf, C.c_double), lf.size, dense_max, max_subsample,
            n_inducing, _lib.ptr(out["status"], C.c_int32), _lib.ptr(out["features"], C.c_double),
            _lib.ptr(out["pred"], C.c_double), _lib.ptr(out["std"], C.c_double),
            _lib.ptr(out["dpred"], C.c_double), _lib.ptr(out["state"], C.c_double)))
        return out

    def thermal_arrays(self, pack: dict, seed: int = 42) -> dict:
        """Array form of the colour-temperature fit (``lcf_batch_thermal``):
        ``pack`` as built by ``thermal.pack_sources``; returns status, out
        [n_src, 8], col_used and the (filled) gp_state rows."""
        ro, co = (np.ascontiguousarray(np.asarray(pack[k], dtype=np.int64)) for k in ("ref_off", "col_off"))
        ns = ro.size - 1
        st = np.array(pack["gp_state"], dtype=np.float64, copy=True, order="C")
        cap = (st.shape[1] - 4) // 2
        arrs = [_f64(pack[k]) for k in ("ref_t", "ref_m", "ref_e", "col_t", "col_m", "col_e", "col_lambda",
                                        "lambda_ref", "duration")]
        out = {"status": np.zeros(ns, np.int32), "out": np.zeros((ns, 8)),
               "col_used": np.zeros(max(int(co[-1]), 1), np.int32)}
        p = [_lib.ptr(x, C.c_double) for x in arrs]
        self._check(self._lib.lcf_batch_thermal(
            self._h, ns, _lib.ptr(ro, C.c_int64), p[0], p[1], p[2], _lib.ptr(co, C.c_int64), p[3], p[4], p[5], p[6],
            p[7], p[8], _lib.ptr(st, C.c_double), cap, C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF),
            _lib.ptr(out["status"], C.c_int32), _lib.ptr(out["out"], C.c_double),
            _lib.ptr(out["col_used"], C.c_int32)))
        out["col_used"] = out["col_used"][:int(co[-1])]
        out["gp_state"] = st
        return out

    def bb_temperature(self, offsets, wavelengths, mags, errs) -> dict:
        """``lcf_batch_bb_temperature``: blackbody log10 T, chi2 and offset of
        every SED (CSR offsets); ok = 0 for sets of fewer than 2 points."""
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        ns = off.size - 1
        lam, m, e = _f64(wavelengths), _f64(mags), _f64(errs)
        out = {"log_t": np.zeros(ns), "chi2": np.zeros(ns), "offset": np.zeros(ns), "ok": np.zeros(ns, np.int32)}
        self._check(self._lib.lcf_batch_bb_temperature(
            self._h, ns, _lib.ptr(off, C.c_int64), _lib.ptr(lam, C.c_double), _lib.ptr(m, C.c_double),
            _lib.ptr(e, C.c_double), _lib.ptr(out["log_t"], C.c_double), _lib.ptr(out["chi2"], C.c_double),
            _lib.ptr(out["offset"], C.c_double), _lib.ptr(out["ok"], C.c_int32)))
        return out

def _gp2d_csr(sources) -> dict:
    off = [0]
    for s in sources:
        if not (len(s.times) == len(s.waves) == len(s.mags) == len(s.noise_var)):
            raise GpuError("Gp2dInput arrays differ in length")
        off.append(off[-1] + len(s.times))
    cat = lambda k: (np.concatenate([np.asarray(getattr(s, k), dtype=np.float64) for s in sources])
                     if sources else np.zeros(0))
    return {"offsets": np.asarray(off, dtype=np.int64), "times": cat("times"), "waves": cat("waves"),
            "mags": cat("mags"), "noise_var": cat("noise_var")}


def _batch_gp_2d(self, sources, query_times, query_waves, amp_candidates, lst_candidates, lsw_candidates,
                 max_subsample: int):
    """``GpuContext::batch_gp_2d`` (src/gpu_cuda.rs:1152-1306)."""
    ns = len(sources)
    if ns == 0:
        return []
    csr = _gp2d_csr(sources)
    qt, qw = _f64(query_times), _f64(query_waves)
    if qt.size != qw.size:
        raise GpuError("query_times and query_waves differ in length")
    a, lt, lw = _f64(amp_candidates), _f64(lst_candidates), _f64(lsw_candidates)
    state, pred, std, rms = np.zeros((ns, 125)), np.zeros((ns, qt.size)), np.zeros((ns, qt.size)), np.zeros(ns)
    self._check(self._lib.lcf_batch_gp2d_fit(
        self._h, ns, _lib.ptr(csr["offsets"], C.c_int64), _lib.ptr(csr["times"], C.c_double),
        _lib.ptr(csr["waves"], C.c_double), _lib.ptr(csr["mags"], C.c_double), _lib.ptr(csr["noise_var"], C.c_double),
        _lib.ptr(qt, C.c_double), _lib.ptr(qw, C.c_double), qt.size, _lib.ptr(a, C.c_double), a.size,
        _lib.ptr(lt, C.c_double), lt.size, _lib.ptr(lw, C.c_double), lw.size, int(max_subsample),
        _lib.ptr(state, C.c_double), _lib.ptr(pred, C.c_double), _lib.ptr(std, C.c_double),
        _lib.ptr(rms, C.c_double)))
    out = []
    for s in range(ns):
        m = int(state[s, 124])
        n_train = len(sources[s].times)
        if m >= 3:
            out.append(Gp2dOutput(list(pred[s]), list(std[s]), float(state[s, 120]),
                                  float(np.sqrt(0.5 / state[s, 121])), float(np.sqrt(0.5 / state[s, 122])),
                                  float(rms[s]), n_train, True))
        else:
            out.append(Gp2dOutput(list(pred[s]), list(std[s]), 0.0, 0.0, 0.0, float("nan"), n_train, False))
    return out


def _gp2d_fit(self, csr: dict, query_t, query_w, amp_candidates, lst_candidates, lsw_candidates, *,
              cap: int = 128, subsample_rule: int = 0, criterion: int = 0, query_per_source: bool = False) -> dict:
    """Array form of the 2D fit (``lcf_batch_gp2d_fit_ex``)."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, w, y, nv = (_f64(csr[k]) for k in ("times", "waves", "mags", "noise_var"))
    qt, qw = _f64(query_t), _f64(query_w)
    nq = qt.shape[-1] if qt.ndim else 0
    a, lt, lw = _f64(amp_candidates), _f64(lst_candidates), _f64(lsw_candidates)
    out = {"state": np.zeros((ns, 3 * cap + 5)), "pred": np.zeros((ns, nq)), "std": np.zeros((ns, nq)),
           "dpred": np.zeros((ns, nq)), "pred_obs": np.zeros(t.size)}
    self._check(self._lib.lcf_batch_gp2d_fit_ex(
        self._h, ns, _lib.ptr(off, C.c_int64), _lib.ptr(t, C.c_double), _lib.ptr(w, C.c_double),
        _lib.ptr(y, C.c_double), _lib.ptr(nv, C.c_double), _lib.ptr(qt, C.c_double), _lib.ptr(qw, C.c_double),
        nq, int(query_per_source), _lib.ptr(a, C.c_double), a.size, _lib.ptr(lt, C.c_double), lt.size,
        _lib.ptr(lw, C.c_double), lw.size, cap, subsample_rule, criterion, _lib.ptr(out["state"], C.c_double),
        _lib.ptr(out["pred"], C.c_double), _lib.ptr(out["std"], C.c_double), _lib.ptr(out["dpred"], C.c_double),
        _lib.ptr(out["pred_obs"], C.c_double)))
    return out


def _fit_gp_2d_arrays(self, csr: dict, query_t, query_w) -> dict:
    """Batched ``fit_gp_2d`` (``lcf_batch_fit_gp_2d``): csr offsets / times /
    waves / mags / noise_var; per-source queries [n][nq].  Returns hyp
    ([n][3] amp, lt, lw), train_rms, n_train, pred, std."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, w, y, nv = (_f64(csr[k]) for k in ("times", "waves", "mags", "noise_var"))
    qt = np.ascontiguousarray(np.asarray(query_t, dtype=np.float64).reshape(ns, -1))
    qw = np.ascontiguousarray(np.asarray(query_w, dtype=np.float64).reshape(ns, -1))
    nq = qt.shape[1]
    out = {"hyp": np.zeros((ns, 3)), "train_rms": np.zeros(ns), "n_train": np.zeros(ns, np.int32),
           "pred": np.zeros((ns, max(nq, 1))), "std": np.zeros((ns, max(nq, 1)))}
    d = lambda a: _lib.ptr(a, C.c_double)
    self._check(self._lib.lcf_batch_fit_gp_2d(
        self._h, ns, _lib.ptr(off, C.c_int64), d(t), d(w), d(y), d(nv), d(qt), d(qw), nq, d(out["hyp"]),
        d(out["train_rms"]), _lib.ptr(out["n_train"], C.c_int32), d(out["pred"]), d(out["std"])))
    out["pred"], out["std"] = out["pred"][:, :nq], out["std"][:, :nq]
    return out


def _gp2d_train(self, csr: dict, theta0, *, max_iters: int = 100, grad_tol: float = 1e-6, query_t=None,
                query_w=None, query_per_source: bool = False, fd_gradient: bool = False) -> dict:
    """2D GP hyper-parameter fit with the Cholesky adjoint
    (``lcf_batch_gp2d_train_ex``): csr holds offsets / times / waves (log10
    wavelength) / mags / noise_var; theta0 [n][4] = ln (a, l_t, l_w, sigma_n).
    ``fd_gradient`` runs the plain form (central finite differences of the
    NLML).  Returns theta, nlml, grad, iters, status, evals ([n][2]
    factorisations / gradient evaluations), pgrad_max (largest projected
    gradient component at the returned point) and, with queries, pred / std."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, w, y, nv = (_f64(csr[k]) for k in ("times", "waves", "mags", "noise_var"))
    th0 = np.ascontiguousarray(np.asarray(theta0, dtype=np.float64).reshape(ns, 4))
    nq = 0
    if query_t is not None:
        qt, qw = _f64(query_t), _f64(query_w)
        nq = qt.shape[-1]
    else:
        qt = qw = np.zeros(1)
    out = {"theta": np.zeros((ns, 4)), "nlml": np.zeros(ns), "grad": np.zeros((ns, 4)),
           "iters": np.zeros(ns, np.int32), "status": np.zeros(ns, np.int32), "evals": np.zeros((ns, 2), np.int32),
           "pred": np.zeros((ns, max(nq, 1))), "std": np.zeros((ns, max(nq, 1)))}
    d = lambda a: _lib.ptr(a, C.c_double)
    self._check(self._lib.lcf_batch_gp2d_train_ex(
        self._h, ns, _lib.ptr(off, C.c_int64), d(t), d(w), d(y), d(nv), d(th0), int(max_iters), float(grad_tol),
        d(qt), d(qw), nq, int(query_per_source), GP2D_FD_GRADIENT if fd_gradient else 0,
        d(out["theta"]), d(out["nlml"]), d(out["grad"]),
        _lib.ptr(out["iters"], C.c_int32), _lib.ptr(out["status"], C.c_int32), d(out["pred"]), d(out["std"]),
        _lib.ptr(out["evals"], C.c_int32)))
    if nq == 0:
        del out["pred"], out["std"]
    out["pgrad_max"] = projected_grad_max(out["theta"], out["grad"])
    return out


# box of the 2D GP hyper-parameter fit on theta = ln (a, l_t, l_w, sigma_n)
GP2D_THETA_LO = np.log([1e-6, 0.1, 1e-3, 1e-6])
GP2D_THETA_HI = np.log([1e6, 1e4, 10.0, 10.0])
GP2D_FD_GRADIENT = 1  # LCF_GP2D_FD_GRADIENT


def projected_grad_max(theta, grad) -> np.ndarray:
    """Largest |component| of the projected gradient per source: components
    at a bound whose gradient points out of the box count as zero."""
    th, g = np.atleast_2d(theta), np.atleast_2d(grad)
    pinned = ((th <= GP2D_THETA_LO) & (g > 0)) | ((th >= GP2D_THETA_HI) & (g < 0))
    return np.max(np.where(pinned, 0.0, np.abs(g)), axis=1)


def _gp2d_nlml_grad(self, csr: dict, theta) -> dict:
    """NLML and gradient of the 2D GP at theta ([n][4]), one evaluation per
    source (``lcf_gp2d_nlml_grad``)."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, w, y, nv = (_f64(csr[k]) for k in ("times", "waves", "mags", "noise_var"))
    th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(ns, 4))
    f, g, st = np.zeros(ns), np.zeros((ns, 4)), np.zeros(ns, np.int32)
    d = lambda a: _lib.ptr(a, C.c_double)
    self._check(self._lib.lcf_gp2d_nlml_grad(self._h, ns, _lib.ptr(off, C.c_int64), d(t), d(w), d(y), d(nv), d(th),
                                             d(f), d(g), _lib.ptr(st, C.c_int32)))
    return {"nlml": f, "grad": g, "status": st}


GpuContext.batch_gp_2d = _batch_gp_2d
GpuContext.gp2d_fit = _gp2d_fit
GpuContext.gp2d_train = _gp2d_train
GpuContext.fit_gp_2d_arrays = _fit_gp_2d_arrays
GpuContext.gp2d_nlml_grad = _gp2d_nlml_grad


@dataclass
class Gp2dInput:
    """One source of a batch 2D GP fit (src/gpu_types.rs:166-172)."""

    times: Sequence[float]
    waves: Sequence[float]  # log10(wavelength / Angstrom)
    mags: Sequence[float]
    noise_var: Sequence[float]


@dataclass
class Gp2dOutput:
    """Output of one source (src/gpu_types.rs:174-184)."""

    pred_grid: list
    std_grid: list
    amp: float
    ls_time: float
    ls_wave: float
    train_rms: float
    n_train: int
    success: bool


@dataclass
class GpBandInput:
    """One band of a batch GP fit (src/gpu_types.rs:143-148)."""

    times: Sequence[float]
    mags: Sequence[float]
    noise_var: Sequence[float]


@dataclass
class DenseGpState:
    """Fitted dense GP: training times, alpha, amplitude, 1/(2 l^2), mean."""

    x: np.ndarray
    alpha: np.ndarray
    amp: float
    inv_2ls2: float
    y_mean: float

    def kernel_amp(self) -> float:
        return self.amp

    def kernel_lengthscale(self) -> float:
        return float(np.sqrt(0.5 / self.inv_2ls2))

    def predict(self, query) -> np.ndarray:
        q = np.asarray(query, dtype=np.float64)
        k = self.amp * np.exp(-(q[:, None] - self.x[None, :]) ** 2 * self.inv_2ls2)
        return k @ self.alpha + self.y_mean


@dataclass
class GpBandOutput:
    """Output of one band (src/gpu_types.rs:150-160)."""

    dense_gp: DenseGpState | None
    pred_grid: list
    std_grid: list
    pred_at_obs: list


def _gp_csr(bands) -> dict:
    off = [0]
    for b in bands:
        if not (len(b.times) == len(b.mags) == len(b.noise_var)):
            raise GpuError("GpBandInput arrays differ in length")
        off.append(off[-1] + len(b.times))
    cat = lambda k: (np.concatenate([np.asarray(getattr(b, k), dtype=np.float64) for b in bands])
                     if bands else np.zeros(0))
    return {"offsets": np.asarray(off, dtype=np.int64), "times": cat("times"), "mags": cat("mags"),
            "noise_var": cat("noise_var")}


def _rows7(x, n: int) -> np.ndarray:
    """[n, MAX_PARAMS] contiguous float64 rows, zero padded."""
    a = np.asarray(x, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(n, -1)
    out = np.zeros((n, MAX_PARAMS))
    out[:, :a.shape[1]] = a[:, :MAX_PARAMS]
    return out


@dataclass
class SviBatchInput:
    """Input of one SVI job (src/gpu_types.rs:190-197)."""

    model_id: int
    pso_params: Sequence[float]
    se_idx: int
    prior_centers: Sequence[float]
    prior_widths: Sequence[float]


@dataclass
class SviBatchOutput:
    """Output of one SVI job (src/gpu_types.rs:199-205)."""

    mu: list
    log_sigma: list
    elbo: float


# prior_params (src/parametric.rs:2072-2115): (center, width) per parameter.
_PRIOR_PARAMS = {
    0: [(0.0, 2.0)] * 6,
    1: [(0.0, 2.0)] * 7,
    2: [(0.0, 2.0)] * 5 + [(1.67, 1.0), (0.0, 2.0)],
    3: [(0.0, 2.0), (0.0, 2.0), (2.3, 1.0), (0.0, 2.0), (0.0, 2.0)],
    4: [(0.0, 2.0), (0.0, 2.0), (3.0, 1.5), (2.3, 1.0), (0.0, 2.0)],
    5: [(0.0, 2.0), (0.0, 2.0), (0.5, 1.0), (1.0, 2.0), (0.0, 2.0)],
    6: [(0.0, 2.0), (0.0, 2.0), (1.0, 2.0), (0.8, 1.0), (2.0, 1.0), (0.0, 2.0)],
    7: [(-1.75, 1.25), (-1.25, 0.75), (0.5, 1.5), (-0.5, 1.5), (0.0, 2.0)],
}


def svi_model_meta(model: int) -> tuple[int, int, int]:
    """``svi_model_meta`` (src/parametric.rs:3277-3290): (model_id, n_params, se_idx)."""
    m = GpuModelName(model)
    return int(m), m.n_params(), m.n_params() - 1


def svi_prior_for_model(model: int, pso_params: Sequence[float]) -> tuple[list, list]:
    """``svi_prior_for_model`` (src/parametric.rs:3259-3273): prior widths of
    prior_params, centres moved to the PSO point except sigma_x's."""
    pri = _PRIOR_PARAMS[int(model)]
    se = len(pri) - 1
    centers = [pri[j][0] if j == se else pso_params[min(j, len(pso_params) - 1)] for j in range(len(pri))]
    widths = [w for _, w in pri]
    return centers, widths


class GpuBatchData:
    """Resident CSR observation data (``GpuBatchData``, src/gpu_cuda.rs:236-292).

    ``GpuBatchData(ctx, sources)`` uploads to ``ctx``'s device;
    ``GpuBatchData(None, sources)`` and ``GpuBatchData.new(sources)`` (the
    reference's signature) upload to the calling thread's current HIP device
    through ``lcf_batch_create_current``.  Either way any context on that
    device may use the batch."""

    def __init__(self, ctx: GpuContext | None = None, sources: Sequence[BatchSource] | None = None, *,
                 csr: dict | None = None):
        self._ctx = ctx
        self._lib = ctx._lib if ctx is not None else _lib.load()
        if csr is None:
            csr = sources_to_csr(sources or [])
        off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
        t = _f64(csr["times"])
        f = _f64(csr["flux"])
        v = _f64(csr["obs_var"])
        up = np.ascontiguousarray(np.asarray(csr["is_upper"], dtype=np.int32))
        uf = _f64(csr["upper_flux"])
        n = int(off[-1]) if off.size else 0
        for name, a in (("times", t), ("flux", f), ("obs_var", v), ("is_upper", up), ("upper_flux", uf)):
            if a.size != n:
                raise GpuError(f"{name} has {a.size} entries, offsets say {n}")
        h = C.c_void_p()
        arrays = (off.size - 1, _lib.ptr(off, C.c_int64), _lib.ptr(t, C.c_double), _lib.ptr(f, C.c_double),
                  _lib.ptr(v, C.c_double), _lib.ptr(up, C.c_int32), _lib.ptr(uf, C.c_double), C.byref(h))
        if ctx is not None:
            ctx._check(self._lib.lcf_batch_create(ctx._h, *arrays))
        else:
            rc = self._lib.lcf_batch_create_current(*arrays)
            if rc != 0:
                raise GpuError((self._lib.lcf_last_error(None) or b"").decode() or f"status {rc}")
        self._h = h
        self.n_sources = off.size - 1
        self.n_obs = n
        self._n_obs = np.diff(off)

    @classmethod
    def new(cls, sources: Sequence[BatchSource]) -> "GpuBatchData":
        return cls(None, sources)

    def n_obs_per_source(self, source_idx: int) -> int:
        return int(self._lib.lcf_batch_n_obs(self._h, int(source_idx)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lcf_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sources_to_csr(sources: Sequence[BatchSource]) -> dict:
    """Concatenate ``BatchSource``s into the CSR arrays of GpuBatchData::new."""
    off = [0]
    cols = {k: [] for k in ("times", "flux", "obs_var", "is_upper", "upper_flux")}
    for s in sources:
        n = len(s.times)
        for k in cols:
            a = getattr(s, k)
            if len(a) != n:
                raise GpuError(f"BatchSource.{k} has {len(a)} entries, times has {n}")
            cols[k].append(np.asarray(a, dtype=np.int32 if k == "is_upper" else np.float64))
        off.append(off[-1] + n)
    out = {k: (np.concatenate(v) if v else np.zeros(0, dtype=np.int32 if k == "is_upper" else np.float64))
           for k, v in cols.items()}
    out["offsets"] = np.asarray(off, dtype=np.int64)
    return out
