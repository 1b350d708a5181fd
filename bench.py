"""Benchmark: batched ragged PSO model selection on one MI355X.

Workload ("ragged_pso_model_selection"): 100k synthetic sources with U{1,2,3}
bands each and log-uniform[10, 500] points per band; every band is fitted with
the Bazin, Villar, TDE, Arnett and Magnetar models, 64 particles x 200
iterations per model (Philox-seeded swarms, FP32 evaluation with a
Kahan-compensated chi-square), and the best model per band is selected.  One
"step" is one such pass over the whole batch, with the observations already
resident in HBM.  The metric is sources/second.

Also reported on the same JSON line:
  * roofline: algorithmic FP32 flop-equivalents of the dominant (slowest)
    model kernel divided by its HIP-event launch time, against MI355X's FP32
    vector peak;
  * cpu_baseline: the oracle (FP64 C++ restatement, all host threads up to 16)
    on a bounded sample of the same workload, in sources/second;
  * plain: the same fit launched once per power-of-two size class instead of
    one persistent size-sorted launch per model.
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
# FP32 flop-equivalents per model-observation evaluation (model + likelihood
# term + Kahan update), counted from models_fit.hpp / pso_kernels.hip: plain
# VALU ops count 1 (FMA 2), each transcendental (exp2/log2/rcp/rsq/sqrt)
# counts 8 (quarter-rate issue x FMA).
FLOPEQ_PER_MOE = {0: 64.0, 1: 91.0, 2: 94.0, 3: 86.0, 4: 74.0}
FP32_PEAK_TFLOPS = 157.3


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sources", type=int, default=100_000)
    ap.add_argument("--particles", type=int, default=64)
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--cpu-sources", type=int, default=192)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this build "
              "(sources are independent, see the sharding notes)", file=sys.stderr)
        return 2

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
    slot_ms = np.zeros(len(MODELS))
    for k in range(args.steps):
        step()
        slot_ms += [ctx.last_model_ms(i) for i in range(len(MODELS))]
        log(f"step {k + 1}/{args.steps}: {ctx.last_kernel_ms:.1f} ms kernels")
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    value = args.sources / (dt / args.steps)
    slot_ms /= args.steps

    # roofline of the dominant model kernel
    npts = np.diff(csr["offsets"]).astype(np.float64)
    moe = (iters_exec.astype(np.float64) * npts[None, :]).sum(axis=1) * args.particles
    dom = int(np.argmax(slot_ms))
    flops = moe[dom] * FLOPEQ_PER_MOE[MODELS[dom]]
    achieved = flops / (slot_ms[dom] * 1e-3) / 1e12
    roof = {"bound": "mfma", "pipe": "fp32 vector ALU + transcendental unit (no matrix ops)",
            "kernel": f"pso_model_kernel<{MODEL_NAMES[MODELS[dom]]}, fp32>",
            "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": None,
            "launch_ms": round(float(slot_ms[dom]), 3),
            "moe_per_launch": float(moe[dom]), "flopeq_per_moe": FLOPEQ_PER_MOE[MODELS[dom]]}

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
        threads = max(1, min(os.cpu_count() or 1, 16))
        log(f"cpu baseline: {args.cpu_sources} sources on {threads} threads")
        cpu = cpu_baseline(pkg, args.cpu_sources, args.particles, args.iters, threads)

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
                   "schedule": "persistent size-sorted queue, one launch per model",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "plain": plain,
        "per_model_ms": {MODEL_NAMES[m]: round(float(slot_ms[i]), 3) for i, m in enumerate(MODELS)},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
