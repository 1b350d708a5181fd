"""Benchmark: batched ragged PSO model selection on one MI355X.

Workload ("ragged_pso_model_selection"): 100k synthetic sources with U{1,2,3}
bands each and log-uniform[10, 500] points per band; every band is fitted with
the Bazin, Villar, TDE, Arnett and Magnetar models, 64 particles x 200
iterations per model (Philox-seeded swarms, FP32 evaluation with a
Kahan-compensated chi-square), and the best model per band is selected.  One
"step" is one such pass over the whole batch, with the observations already
resident in HBM.  The metric is sources/second.

The five model families run as one persistent launch over every (band,
model) task (mixed families on one size-sorted device queue).

Also reported on the same JSON line:
  * roofline: algorithmic FP32 flop-equivalents of that launch divided by its
    HIP-event launch time, against MI355X's FP32 vector peak;
  * per_model: the same fit as one persistent launch per model, with each
    model's launch time and the slowest launch's roofline;
  * cpu_baseline: the oracle (FP64 C++ restatement, all host threads up to 16)
    on a bounded sample of the same workload, in sources/second;
  * plain: the same fit launched once per power-of-two size class and model;
  * svi: the SVI workload (10k Villar sources x 3 bands x 60 points, 16
    samples, up to 2000 Adam steps, FP32) fitted by the batched adjoint kernel
    and by the plain form (forward-difference gradients, one launch per size
    class), with the agreement of the two posteriors;
  * nonparametric: the GP workload (10k Bazin sources x 3 bands x U{10..200}
    magnitudes, 200-point grid, dense GP up to 200 detections) through the
    whole band pipeline (fit + features) on the device, with the oracle's
    process_band on a bounded sample as its CPU baseline.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG_DIR = ROOT / "lightcurve-fitting_amd"

METRIC = "sources/sec, batched+adjoint vs plain GPU fit at equal accuracy, 100k src"
MODELS = (0, 1, 2, 3, 4)
MODEL_NAMES = {0: "Bazin", 1: "Villar", 2: "Tde", 3: "Arnett", 4: "Magnetar"}
# Algorithmic FP32 flop-equivalents per model-observation evaluation (MOE:
# prediction, likelihood term and its share of the Kahan update): the fixed
# per-model operation counts of the model formulas (Bazin 12 ops + 5
# transcendentals, Villar 20 + 6, TDE 18 + 7, Arnett 18 + 6, Magnetar 15 + 5,
# including the NLL's 6 ops + log + rcp and the Kahan update's 3), a
# transcendental counted as 8 flop-equivalents (quarter-rate issue x FMA = 2).
# They do not depend on how the kernel is compiled.
FLOPEQ_ALGO = {0: 12 + 5 * 8, 1: 20 + 6 * 8, 2: 18 + 7 * 8, 3: 18 + 6 * 8, 4: 15 + 5 * 8}
# Issued-instruction flop-equivalents of the shipped check-free cost loops of
# pso_kernel<MODEL, fp32> (every VALU instruction of one loop trip, the
# update of the particle's sum included, over the trip's four observations):
# an FMA 2, a transcendental 4 (two issue slots x 2), any other VALU
# instruction 1; detections and upper limits (censored term) differ.
# Reported beside the algorithmic figure as the kernel's own issue rate.
FLOPEQ_DET = {0: 33.0, 1: 47.0, 2: 48.0, 3: 48.0, 4: 40.0}
FLOPEQ_UP = {0: 42.0, 1: 56.0, 2: 57.0, 3: 57.0, 4: 49.0}
# SVI flop-equivalents per (sample, observation, step): the Villar forward
# evaluation (20 ops + 6 transcendentals x 8) and its adjoint (about 45 ops).
SVI_FLOPEQ_FWD, SVI_FLOPEQ_ADJ = 20 + 6 * 8, 45
FP32_PEAK_TFLOPS = 157.3
# HBM traffic of the mixed-family launch on the default 100k-source workload,
# from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB per
# dispatch) over `bench.py --steps 1 --warmup 0 --no-cpu --no-plain --no-svi
# --no-np --no-gp2d --no-tune --no-accuracy`, each scaled by what the same
# counter reads for `bench.py --copy-probe 1024` (a 1 GiB device copy):
# FETCH_SIZE 524308 KiB for 1048576 KiB read (x 2.0), WRITE_SIZE 1048576 KiB
# written (x 1.0).
MIXED_FETCH_KIB, MIXED_WRITE_KIB = 917148.0, 187227.0
FETCH_SCALE, WRITE_SCALE = 1048576.0 / 524308.0, 1.0


def load_package():
    name = "lightcurve_fitting_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    """Host threads of the CPU baselines: every core the process may use, at
    most 16 (the host share that comes with one GPU on the measurement
    machines; os.cpu_count() is reported beside it)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def pso_accuracy(pkg, csr: dict, fp32_out: dict, n_src: int, particles: int, iters: int) -> dict:
    """Accuracy of the timed FP32 fit against FP64 on the bands of the first
    n_src sources of the workload: the FP64 kernels (which reproduce the
    oracle's swarm exactly) on those bands, seed 42 as in the timed pass and
    seed 43 for the method's own seed-to-seed spread.  A band's swarm depends
    only on its data, its index and the seed, so the first bands of the batch
    fitted alone follow the same streams as in the full batch."""
    nb = int(np.searchsorted(csr["band_source"], n_src))
    sub = pkg.synthetic.subset(csr, np.arange(nb))
    c64 = pkg.GpuContext(0, precision=pkg.FP64)
    d64 = pkg.GpuBatchData(c64, csr=sub)
    r64 = c64.fit_models(d64, MODELS, particles, iters, iters, seed=42)
    r64b = c64.fit_models(d64, MODELS, particles, iters, iters, seed=43)
    d64.close()
    c64.close()
    valid = np.diff(sub["offsets"]) > 0
    bm32, b32 = fp32_out["best_model"][:nb], fp32_out["best_cost"][:nb]
    gap = ((b32 - r64["best_cost"]) / np.maximum(np.abs(r64["best_cost"]), 1.0))[valid]
    same = (bm32 == r64["best_model"]) & valid
    prel = []
    for j in np.nonzero(same)[0]:
        i = MODELS.index(int(r64["best_model"][j]))
        k = pkg.GpuModelName(int(r64["best_model"][j])).n_params()
        a, b = fp32_out["model_params"][i, j, :k], r64["best_params"][j, :k]
        prel.append(float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0))))
    prel = np.asarray(prel)
    q = lambda x, qs: [float(np.quantile(x, v)) for v in qs]
    out = {"reference": "FP64 kernels (the oracle's swarm), same seed, same bands", "sources": n_src,
           "bands": int(valid.sum()),
           "best_model_agreement": round(float((bm32 == r64["best_model"])[valid].mean()), 5),
           "fp64_seed_to_seed_agreement": round(float((r64b["best_model"] == r64["best_model"])[valid].mean()), 5),
           "best_cost_rel_gap_quantiles": {"q05": None, "q50": None, "q95": None},
           "frac_fp32_worse_1e-4": round(float((gap > 1e-4).mean()), 5),
           "frac_fp32_better_1e-4": round(float((gap < -1e-4).mean()), 5),
           "frac_cost_within_1e-4": round(float((np.abs(gap) <= 1e-4).mean()), 5),
           "param_rel_err_quantiles_same_model": dict(zip(("q50", "q90", "q99"), q(prel, (0.5, 0.9, 0.99)))),
           "note": "fp64_seed_to_seed_agreement is the swarm's own spread at this budget"}
    out["best_cost_rel_gap_quantiles"] = dict(zip(("q05", "q50", "q95"), q(gap, (0.05, 0.5, 0.95))))
    log(f"accuracy: agreement {out['best_model_agreement']:.4f} (fp64 seed spread "
        f"{out['fp64_seed_to_seed_agreement']:.4f}) on {out['bands']} bands")
    return out


def cpu_baseline(pkg, n_sources: int, particles: int, iters: int, threads: int) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    csr = pkg.synthetic.ragged_selection_batch(n_sources, seed=1234)
    t0 = time.perf_counter()
    oracle.batch_model_select(csr, list(MODELS), particles, iters, iters, float("-inf"), n_threads=threads)
    dt = time.perf_counter() - t0
    return {"value": n_sources / dt, "unit": "sources/s", "cores": threads, "kind": "port",
            "sample": f"{n_sources} sources ({csr['offsets'].size - 1} bands) of the same workload, "
                      f"oracle FP64 C++, {threads} threads, {dt:.1f} s"}


def svi_leg(pkg, n_sources: int, n_steps: int, n_samples: int, lr: float, reps: int) -> dict:
    """Villar SVI workload: n_sources x 3 bands x 60 points, PSO start points
    (untimed), then the SVI fit at all n_steps Adam steps (no early stop),
    timed twice over: the batched adjoint form (analytic forward + adjoint in
    one kernel, one persistent launch) and the plain form (forward-difference
    gradients, one launch per power-of-two size class).  The reference's
    early-stopped fit is timed beside them.  Agreement of the two posteriors
    is reported with the times; the roofline counts the forward + adjoint
    flop-equivalents of every (sample, observation, step)."""
    FIXED, FD, EXACT = 8, 4, 16
    csr = pkg.synthetic.villar_svi_batch(n_sources, seed=77)
    nj = csr["offsets"].size - 1
    ctx = pkg.GpuContext(0, precision=pkg.FP32, schedule=pkg.SCHED_PERSISTENT)
    data = pkg.GpuBatchData(ctx, csr=csr)
    fit = ctx.fit_models(data, [1], 64, 200, 200, seed=42)
    pso = fit["model_params"][0]
    mids = np.full(nj, 1, dtype=np.int32)
    pc, pw = np.zeros((nj, 7)), np.zeros((nj, 7))
    for j in range(nj):
        c, w = pkg.svi_prior_for_model(1, pso[j, :7])
        pc[j], pw[j] = c, w
    log(f"svi: {nj} Villar bands, PSO start points ready")

    def run(flags, schedule):
        ctx.set_schedule(schedule)
        ctx.synchronize()
        t0 = time.perf_counter()
        out = ctx.svi_fit(data, mids, pso, pc, pw, n_steps, n_samples, lr, seed=42, flags=flags)
        ctx.synchronize()
        return out, time.perf_counter() - t0, ctx.last_kernel_ms

    run(FIXED | EXACT, pkg.SCHED_PERSISTENT)  # warm-up
    best = None
    for _ in range(reps):
        r = run(FIXED | EXACT, pkg.SCHED_PERSISTENT)
        best = r if best is None or r[1] < best[1] else best
    adj, dt_adj, kms_adj = best
    log(f"svi adjoint ({n_steps} steps): {dt_adj * 1e3:.1f} ms")
    plain, dt_plain, kms_plain = run(FIXED | EXACT | FD, pkg.SCHED_PER_CLASS)
    log(f"svi plain: {dt_plain * 1e3:.1f} ms")
    early, dt_early, kms_early = run(EXACT, pkg.SCHED_PERSISTENT)
    ctx.set_schedule(pkg.SCHED_PERSISTENT)
    log(f"svi early-stopped: {dt_early * 1e3:.1f} ms, mean {early['steps'].mean():.1f} steps")
    sig = np.exp(plain["log_sigma"][:, :7])
    dmu = np.abs(adj["mu"][:, :7] - plain["mu"][:, :7]) / sig
    dsg = np.abs(np.exp(adj["log_sigma"][:, :7]) - sig) / sig
    evals = float(adj["steps"].astype(np.float64).sum() * n_samples * 60)
    flops = evals * (SVI_FLOPEQ_FWD + SVI_FLOPEQ_ADJ)
    achieved = flops / (kms_adj * 1e-3) / 1e12
    return {"workload": "villar_svi", "sources": n_sources, "bands": nj, "points_per_band": 60,
            "samples": n_samples, "steps": n_steps, "early_stop": False, "lr": lr, "dtype": "fp32",
            "upper_limit_adjoint": "exact derivative of log_normal_cdf (LCF_SVI_EXACT_TAIL)",
            "value": round(n_sources / dt_adj, 2), "unit": "sources/s", "ms": round(dt_adj * 1e3, 3),
            "kernel_ms": round(kms_adj, 3),
            "plain": {"value": round(n_sources / dt_plain, 2), "ms": round(dt_plain * 1e3, 3),
                      "kernel_ms": round(kms_plain, 3),
                      "form": "forward-difference gradients, one launch per power-of-two size class"},
            "speedup_vs_plain": round(dt_plain / dt_adj, 3),
            "agreement": {"mu_q50_over_sigma": float(np.quantile(dmu, 0.5)),
                          "mu_q95_over_sigma": float(np.quantile(dmu, 0.95)),
                          "sigma_q95_rel": float(np.quantile(dsg, 0.95))},
            "early_stopped": {"value": round(n_sources / dt_early, 2), "ms": round(dt_early * 1e3, 3),
                              "steps_mean": float(early["steps"].mean()),
                              "rule": "EMA of the ELBO after 50 steps, 20-step stall (cuda/svi.cu:297-312)"},
            "sample_obs_step_evals": evals,
            "roofline": {"bound": "valu", "pipe": "fp32 vector ALU + transcendental unit",
                         "kernel": "svi_kernel<Villar, fp32>", "achieved": round(achieved, 3),
                         "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                         "flopeq_per_sample_obs_step": SVI_FLOPEQ_FWD + SVI_FLOPEQ_ADJ, "traffic": None}}


def np_flops(csr: dict, n_grid: int, dense_max: int, amps, lsf) -> float:
    """Algorithmic FP64 flop-equivalents of the nonparametric pipeline over a
    batch (dense bands), an FP64 exponential counted as 20: per hyper-parameter
    candidate (those with l >= the band's minimum length scale) the Cholesky
    m^3/3, the two triangular solves 2 m^2, the kernel matrix m(m+1)/2 exps
    and the train RMS from the weights (2 m: the residual at training point i
    is -D_i alpha_i); for the winner the refit, the std at n_grid + min(m, 200)
    points (m^2 + m exps each) and the means at 2 n_grid + m + fine-grid
    points (m (exp + 2) each)."""
    EXP = 20.0
    off = csr["offsets"]
    tot = 0.0
    for b in range(off.size - 1):
        t = csr["times"][off[b]:off[b + 1]]
        if t.size < 5:
            continue
        e = csr["mag_errs"][off[b]:off[b + 1]]
        det = e <= 1.0857362047581294 / 3.0
        m = int(det.sum())
        if m < 5 or m > dense_max:
            continue
        g = np.diff(np.sort(t[det]))
        g = np.sort(g[np.isfinite(g) & (g > 0)])
        min_ls = max(2.0 * (g[g.size // 2] if g.size else 1.0), 0.1)
        dur = csr["t_hi"][b] - csr["t_lo"][b]
        nc = len(amps) * sum(1 for f in lsf if max(dur / f, 0.1) >= min_ls)
        cand = m ** 3 / 3.0 + 2.0 * m * m + m * (m + 1) / 2.0 * EXP + 2.0 * m
        nf = min(int(max(dur, 1.0)) + 1, 2000)
        final = cand + (n_grid + min(m, 200)) * (m * m + m * EXP) + (2 * n_grid + m + nf) * m * (EXP + 2.0)
        tot += nc * cand + final
    return tot


def np_leg(pkg, n_sources: int, n_grid: int, dense_max: int, reps: int, cpu_sources: int, threads: int,
           do_cpu: bool) -> dict:
    """Nonparametric workload: n_sources Bazin sources x 3 bands x U{10..200}
    points (magnitudes), per-source grid of n_grid points, dense GP up to
    dense_max detections; the whole process_band (fit + features) per band
    on the device.  Time = device time of the launches (inputs resident);
    the full call with transfers is reported beside it.  CPU baseline: the
    oracle's process_band on a bounded sample of the same workload."""
    csr = pkg.synthetic.bazin_mag_batch(n_sources, seed=2024)
    nb = csr["offsets"].size - 1
    ctx = pkg.GpuContext(0)
    kw = dict(n_grid=n_grid, dense_max=dense_max)
    ctx.nonparametric_arrays(csr, csr["t_lo"], csr["t_hi"], **kw)  # warm-up
    best, best_call = None, None
    for _ in range(max(reps, 1)):
        t0 = time.perf_counter()
        r = ctx.nonparametric_arrays(csr, csr["t_lo"], csr["t_hi"], **kw)
        call = time.perf_counter() - t0
        k = ctx.last_kernel_ms
        best = k if best is None or k < best else best
        best_call = call if best_call is None or call < best_call else best_call
    log(f"nonparametric: {nb} bands, {best:.1f} ms device, {best_call * 1e3:.1f} ms with transfers")
    flops = np_flops(csr, n_grid, dense_max, (0.3, 1.0, 3.0, 10.0), (3.0, 6.0, 12.0, 24.0, 48.0))
    achieved = flops / (best * 1e-3) / 1e12
    out = {"workload": "gp_nonparametric", "sources": n_sources, "bands": nb,
           "observations": int(csr["offsets"][-1]), "points_per_band": "U{10..200}", "grid": n_grid,
           "dense_max": dense_max, "dtype": "f64", "value": round(n_sources / (best * 1e-3), 2),
           "unit": "sources/s", "ms": round(best, 3), "ms_with_transfers": round(best_call * 1e3, 3),
           "status_counts": np.bincount(r["status"], minlength=3).tolist(),
           "roofline": {"bound": "mfma", "pipe": "fp64 matrix cores (Cholesky panel tiles, predictive-std "
                                                  "substitution) + fp64 vector ALU, both 78.6 TFLOP/s",
                        "kernel": "np_band_kernel (all detection-count classes)",
                        "achieved": round(achieved, 4), "peak": 78.6, "unit": "TFLOP/s",
                        "frac": round(achieved / 78.6, 5), "flop_equivalents": flops,
                        "note": "bound by the latency of dependent chains (diagonal-block pivots, "
                                "row and weight substitutions) with four waves per band, not by "
                                "FP64 throughput"}}
    if do_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        sub = pkg.synthetic.bazin_mag_batch(cpu_sources, seed=2024)
        t0 = time.perf_counter()
        oracle.batch_nonparametric(sub, sub["t_lo"], sub["t_hi"], n_threads=threads, **kw)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cpu_sources / dt, 2), "unit": "sources/s", "cores": threads,
                               "kind": "port", "sample": f"{cpu_sources} sources of the same workload, oracle "
                                                         f"FP64 C++ process_band, {threads} threads, {dt:.1f} s"}
        log(f"nonparametric cpu baseline: {cpu_sources / dt:.1f} sources/s")
    return out


def gp2d_leg(pkg, n_sources: int, max_iters: int, n_times: int, reps: int, cpu_sources: int, threads: int,
             do_cpu: bool) -> dict:
    """2D GP hyper-parameter workload: n_sources cooling-blackbody x Bazin
    sources of U{30..512} points over g/r/i, L-BFGS (at most max_iters
    iterations) on the exact Cholesky-adjoint gradient of the log marginal
    likelihood per source, predictions on n_times x 3 bands at the trained
    hyper-parameters, then the blackbody temperature fit per (source, time).
    Time = device time of the two launches (inputs resident).  Roofline: the
    O(n^3) FP64 work per NLML evaluation (Cholesky, n^3/3) and gradient
    evaluation (inverse + X^T X, 2 n^3 / 3), counted from the kernel's own
    evaluation counters.  CPU baseline: the oracle's driver on a bounded
    sample of the same workload."""
    c = pkg.synthetic.cooling_bb_batch(n_sources, seed=2025)
    ns = n_sources
    wl = np.asarray(pkg.synthetic.GRI_WAVELENGTH)
    tq = np.linspace(0.0, 100.0, n_times)
    qt, qw = np.repeat(tq, 3), np.tile(np.log10(wl), n_times)
    ctx = pkg.GpuContext(0)
    run = lambda: ctx.gp2d_train(c, c["theta0"], max_iters=max_iters, grad_tol=1e-6, query_t=qt, query_w=qw)
    run()  # warm-up
    best = None
    for _ in range(max(reps, 1)):
        r = run()
        k = ctx.last_kernel_ms
        best = (k, r) if best is None or k < best[0] else best
    ms_train, r = best
    # plain form: the same L-BFGS on central finite differences of the NLML
    # (eight factorisations per gradient), one launch per size class
    ctx.set_schedule(pkg.SCHED_PER_CLASS)
    rp = ctx.gp2d_train(c, c["theta0"], max_iters=max_iters, grad_tol=1e-6, query_t=qt, query_w=qw,
                        fd_gradient=True)
    ms_plain = ctx.last_kernel_ms
    ctx.set_schedule(pkg.SCHED_PERSISTENT)
    log(f"gp2d plain (FD gradients, per size class): {ms_plain:.1f} ms")
    both = (r["status"] < 3) & (rp["status"] < 3)
    drel = np.abs(r["nlml"] - rp["nlml"]) / np.maximum(np.abs(r["nlml"]), 1.0)
    pg = r["pgrad_max"]
    per_status = {}
    for st_ in range(3):
        sel = r["status"] == st_
        if sel.any():
            per_status[str(st_)] = {"sources": int(sel.sum()),
                                    "pgrad_max_q50": float(np.quantile(pg[sel], 0.5)),
                                    "pgrad_max_q95": float(np.quantile(pg[sel], 0.95)),
                                    "pgrad_max_max": float(pg[sel].max())}
    ok = r["status"] < 3
    off = np.zeros(ns * n_times + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.repeat(np.where(ok, 3, 0), n_times))
    sel = np.repeat(ok, n_times)
    lam = np.tile(wl, ns * n_times).reshape(ns * n_times, 3)[sel].ravel()
    mag = r["pred"].reshape(ns * n_times, 3)[sel].ravel()
    err = r["std"].reshape(ns * n_times, 3)[sel].ravel()
    bb = ctx.bb_temperature(off, lam, mag, err)
    ms_bb = ctx.last_kernel_ms
    ms = ms_train + ms_bb
    n = np.diff(c["offsets"]).astype(np.float64)
    ev = r["evals"].astype(np.float64)
    flops = float(np.sum(ev[:, 0] * (n + 1) ** 3 / 3.0 + ev[:, 1] * 2.0 * n ** 3 / 3.0))
    achieved = flops / (ms_train * 1e-3) / 1e12
    log(f"gp2d: {ns} sources, {ms_train:.1f} ms training + {ms_bb:.2f} ms blackbody, "
        f"{achieved:.3f} TFLOP/s, evals {ev.sum(axis=0).astype(int).tolist()}")
    out = {"workload": "gp2d_hyperparameter_adjoint_lbfgs", "sources": ns, "observations": int(c["offsets"][-1]),
           "points_per_source": "U{30..512}", "lbfgs_max_iters": max_iters, "dtype": "f64",
           "value": round(ns / (ms * 1e-3), 2), "unit": "sources/s", "ms": round(ms, 3),
           "ms_training": round(ms_train, 3), "ms_blackbody": round(ms_bb, 3),
           "status_counts": np.bincount(r["status"], minlength=4).tolist(),
           "mean_iterations": round(float(r["iters"].mean()), 2),
           "evaluations": {"nlml": int(ev[:, 0].sum()), "gradient": int(ev[:, 1].sum())},
           "temperatures_fitted": int(np.sum(bb["ok"])),
           "projected_gradient_by_status": per_status,
           "plain": {"ms_training": round(ms_plain, 3), "value": round(ns / ((ms_plain + ms_bb) * 1e-3), 2),
                     "form": "L-BFGS on central finite-difference gradients (8 NLML factorisations each), "
                             "one launch per power-of-two size class",
                     "status_counts": np.bincount(rp["status"], minlength=4).tolist(),
                     "evaluations": {"factorisations": int(rp["evals"][:, 0].sum()),
                                     "gradient": int(rp["evals"][:, 1].sum())},
                     "nlml_rel_diff_q50": float(np.quantile(drel[both], 0.5)),
                     "nlml_rel_diff_q99": float(np.quantile(drel[both], 0.99))},
           "speedup_vs_plain": round((ms_plain + ms_bb) / ms, 3),
           "roofline": {"bound": "mfma", "pipe": "fp64 matrix cores (tile products) + fp64 vector ALU, both 78.6 TFLOP/s", "kernel": "gp2d_hyper_kernel",
                        "achieved": round(achieved, 4), "peak": 78.6, "unit": "TFLOP/s",
                        "frac": round(achieved / 78.6, 5), "flops": flops}}
    if do_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        sub = pkg.synthetic.cooling_bb_batch(cpu_sources, seed=2025)
        t0 = time.perf_counter()
        oracle.batch_gp2d_train(sub, sub["theta0"], max_iters, 1e-6, n_threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cpu_sources / dt, 3), "unit": "sources/s", "cores": threads,
                               "kind": "port", "sample": f"{cpu_sources} sources of the same workload, oracle FP64 "
                                                         f"C++ NLML + adjoint + L-BFGS (training only), {threads} "
                                                         f"threads, {dt:.1f} s"}
        log(f"gp2d cpu baseline: {cpu_sources / dt:.2f} sources/s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sources", type=int, default=100_000)
    ap.add_argument("--particles", type=int, default=64)
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--cpu-sources", type=int, default=192)
    ap.add_argument("--cpu-sources-1t", type=int, default=16)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--accuracy-sources", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-svi", action="store_true")
    ap.add_argument("--svi-sources", type=int, default=10_000)
    ap.add_argument("--svi-steps", type=int, default=2000)
    ap.add_argument("--no-np", action="store_true")
    ap.add_argument("--np-sources", type=int, default=10_000)
    ap.add_argument("--np-cpu-sources", type=int, default=3000)
    ap.add_argument("--no-gp2d", action="store_true")
    ap.add_argument("--gp2d-sources", type=int, default=5000)
    ap.add_argument("--gp2d-cpu-sources", type=int, default=48)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (a plain copy kernel of known "
                         "size, the scale of the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(sources are independent, see the sharding notes)", file=sys.stderr)
        return 2
    if args.copy_probe > 0:
        import torch
        x = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {args.copy_probe} MiB")
        return 0

    pkg = load_package()
    log(f"generating {args.sources} sources")
    csr = pkg.synthetic.ragged_selection_batch(args.sources, seed=2024)
    n_bands = csr["offsets"].size - 1
    n_obs = int(csr["offsets"][-1])
    ctx = pkg.GpuContext(0, precision=pkg.FP32, schedule=pkg.SCHED_PERSISTENT)
    data = pkg.GpuBatchData(ctx, csr=csr)
    log(f"resident: {n_bands} bands, {n_obs} observations, {ctx.num_cus} CUs")

    def step(per_model=False):
        return ctx.fit_models(data, MODELS, args.particles, args.iters, args.iters, seed=42,
                              per_model=per_model)

    log("first pass (untimed)")
    t_first = time.perf_counter()
    first = step(per_model=True)
    iters_exec = first["model_iters"]
    log(f"first pass done in {time.perf_counter() - t_first:.2f} s")
    for w in range(max(args.warmup - 1, 0)):
        step()
        log(f"warmup {w + 2}/{args.warmup} done")
    ctx.synchronize()
    t0 = time.perf_counter()
    mixed_ms = 0.0
    for k in range(args.steps):
        step()
        mixed_ms += ctx.last_model_ms(0)
        log(f"step {k + 1}/{args.steps}: {ctx.last_kernel_ms:.1f} ms kernels")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = args.sources / (dt / args.steps)
    mixed_ms /= args.steps

    # roofline of the one launch over every (band, model) task: the models'
    # algorithmic flop-equivalents over their executed iterations
    npts = np.diff(csr["offsets"]).astype(np.float64)
    nup = np.add.reduceat(csr["is_upper"].astype(np.float64), csr["offsets"][:-1]) if npts.size else npts
    nup = np.where(npts > 0, nup, 0.0)
    it64 = iters_exec.astype(np.float64)
    moe_up = (it64 * nup[None, :]).sum(axis=1) * args.particles
    moe = (it64 * npts[None, :]).sum(axis=1) * args.particles
    moe_det = moe - moe_up
    fe_algo = np.array([FLOPEQ_ALGO[m] for m in MODELS], dtype=np.float64)
    flops = float((moe * fe_algo).sum())
    issued = float(sum(moe_det[i] * FLOPEQ_DET[m] + moe_up[i] * FLOPEQ_UP[m] for i, m in enumerate(MODELS)))
    # measured HBM bytes per launch (default workload only); the algorithmic
    # bytes are one read of every packed record per model plus the outputs
    traffic = None
    if args.sources == 100_000 and args.particles == 64 and args.iters == 200 and MIXED_FETCH_KIB:
        traffic = round((MIXED_FETCH_KIB * FETCH_SCALE + MIXED_WRITE_KIB * WRITE_SCALE) * 1024.0)
    algo_bytes = len(MODELS) * (n_obs * 16 + n_bands * (7 * 8 + 8 + 4 + 4 + 8))
    achieved = flops / (mixed_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "pipe": "fp32 vector ALU + transcendental unit (no matrix ops)",
            "kernel": "pso_mixed_kernel<fp32> (" + ", ".join(MODEL_NAMES[m] for m in MODELS) + ")",
            "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "launch_ms": round(float(mixed_ms), 3),
            "algorithmic_bytes": int(algo_bytes),
            "moe_per_launch": float(moe.sum()), "upper_limit_moe_fraction": round(float(moe_up.sum() / moe.sum()), 4),
            "flopeq_per_moe": round(flops / float(moe.sum()), 3),
            "flopeq_convention": "model formula operation count, transcendental = 8 (per model: "
                                 + ", ".join(f"{MODEL_NAMES[m]} {FLOPEQ_ALGO[m]}" for m in MODELS) + ")",
            "issued_flopeq_per_moe": round(issued / float(moe.sum()), 3),
            "issued_tflops": round(issued / (mixed_ms * 1e-3) / 1e12, 3)}

    # the same fit as one persistent launch per model: per-model launch times
    # and the roofline of the slowest model's launch
    ctx.set_schedule(pkg.SCHED_PER_MODEL)
    step()
    ctx.synchronize()
    t_pm = time.perf_counter()
    slot_ms = np.zeros(len(MODELS))
    for _ in range(args.steps):
        step()
        slot_ms += [ctx.last_model_ms(i) for i in range(len(MODELS))]
    ctx.synchronize()
    pm_dt = (time.perf_counter() - t_pm) / args.steps
    slot_ms /= args.steps
    ctx.set_schedule(pkg.SCHED_PERSISTENT)
    dom = int(np.argmax(slot_ms))
    m_dom = MODELS[dom]
    ach_dom = moe[dom] * FLOPEQ_ALGO[m_dom] / (slot_ms[dom] * 1e-3) / 1e12
    per_model = {"value": round(args.sources / pm_dt, 2), "unit": "sources/s", "ms_per_step": round(pm_dt * 1e3, 3),
                 "schedule": "one persistent size-sorted launch per model",
                 "per_model_ms": {MODEL_NAMES[m]: round(float(slot_ms[i]), 3) for i, m in enumerate(MODELS)},
                 "slowest_kernel": {"kernel": f"pso_model_kernel<{MODEL_NAMES[m_dom]}, fp32>",
                                    "launch_ms": round(float(slot_ms[dom]), 3), "achieved": round(ach_dom, 3),
                                    "frac": round(ach_dom / FP32_PEAK_TFLOPS, 4),
                                    "flopeq_per_moe": FLOPEQ_ALGO[m_dom],
                                    "issued_flopeq_per_moe": {"detection": FLOPEQ_DET[m_dom],
                                                              "upper_limit": FLOPEQ_UP[m_dom]}}}
    log(f"per-model launches: {pm_dt * 1e3:.1f} ms per step")

    # end to end: upload + pack of the observations, the fit, the downloads
    ctx.synchronize()
    t_up = time.perf_counter()
    data_e2e = pkg.GpuBatchData(ctx, csr=csr)
    ms_upload = (time.perf_counter() - t_up) * 1e3
    t_e2e = time.perf_counter()
    ctx.fit_models(data_e2e, MODELS, args.particles, args.iters, args.iters, seed=42, per_model=True)
    ms_e2e = ms_upload + (time.perf_counter() - t_e2e) * 1e3
    data_e2e.close()
    log(f"end to end with upload: {ms_e2e:.1f} ms (upload {ms_upload:.1f} ms)")

    # autotuned launch shapes: measured per band-size bucket (untimed), then
    # the same steps under the tuned schedule (identical results)
    tuned = None
    if not args.no_tune:
        t_tune = time.perf_counter()
        table = ctx.autotune_pso(data, MODELS, args.particles, 40, 40, seed=42)
        ms_tune = (time.perf_counter() - t_tune) * 1e3
        ctx.set_schedule(pkg.SCHED_TUNED)
        step()
        ctx.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        tdt = (time.perf_counter() - t2) / args.steps
        ctx.set_schedule(pkg.SCHED_PERSISTENT)
        tuned = {"value": round(args.sources / tdt, 2), "unit": "sources/s", "ms_per_step": round(tdt * 1e3, 3),
                 "speedup_vs_default": round(ms_per_step / (tdt * 1e3), 4), "tuning_ms": round(ms_tune, 1),
                 "schedule": "one persistent launch per band-size bucket with the measured LDS tile and "
                             "waves per CU", "table": table}
        log(f"autotuned schedule: {tdt * 1e3:.1f} ms per step (tuning {ms_tune:.0f} ms)")

    accuracy = None
    if not args.no_accuracy:
        accuracy = pso_accuracy(pkg, csr, first, args.accuracy_sources, args.particles, args.iters)

    plain = None
    if not args.no_plain:
        ctx.set_schedule(pkg.SCHED_PER_CLASS)
        t1 = time.perf_counter()
        step()
        ctx.synchronize()
        pdt = time.perf_counter() - t1
        ctx.set_schedule(pkg.SCHED_PERSISTENT)
        plain = {"value": args.sources / pdt, "unit": "sources/s", "ms_per_step": pdt * 1e3,
                 "schedule": "one launch per power-of-two size class per model"}
        log(f"plain per-class schedule: {pdt * 1e3:.1f} ms")

    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        log(f"cpu baseline: {args.cpu_sources} sources on {threads} threads")
        cpu = cpu_baseline(pkg, args.cpu_sources, args.particles, args.iters, threads)
        one = cpu_baseline(pkg, args.cpu_sources_1t, args.particles, args.iters, 1)
        cpu["single_thread"] = {"value": one["value"], "unit": "sources/s", "sample": one["sample"]}
        cpu["nproc"] = os.cpu_count()

    svi = None
    if not args.no_svi:
        svi = svi_leg(pkg, args.svi_sources, args.svi_steps, 16, 0.01, reps=max(args.steps, 1))

    npl = None
    if not args.no_np:
        npl = np_leg(pkg, args.np_sources, 200, 200, max(args.steps, 1), args.np_cpu_sources,
                     cpu_threads(), not args.no_cpu)

    gp2 = None
    if not args.no_gp2d:
        gp2 = gp2d_leg(pkg, args.gp2d_sources, 100, 20, max(args.steps, 1), args.gp2d_cpu_sources,
                       cpu_threads(), not args.no_cpu)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "sources/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "ragged_pso_model_selection", "sources": args.sources, "bands": n_bands,
                   "observations": n_obs, "points_per_band": "log-uniform[10,500]",
                   "bands_per_source": "U{1,2,3}", "models": [MODEL_NAMES[m] for m in MODELS],
                   "particles": args.particles, "iterations": args.iters,
                   "schedule": "one persistent launch over every (band, model) task: mixed model "
                               "families on one size-sorted device queue",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "ms_with_upload": round(ms_e2e, 3),
        "ms_upload": round(ms_upload, 3),
        "accuracy": accuracy,
        "plain": plain,
        "autotuned": tuned,
        "per_model": per_model,
        "svi": svi,
        "gp2d": gp2,
        "nonparametric": npl,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
