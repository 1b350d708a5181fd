"""Build recipe for the gfx950 shared library ``liblcf_hip.so``.

Compiles every ``csrc/*.hip`` translation unit with hipcc for gfx950 and links
them into ``lightcurve-fitting_amd/liblcf_hip.so`` (in-tree, next to this
file).  Run ``python lightcurve-fitting_amd/build.py`` or call ``build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
BUILD_DIR = PKG_DIR / "_build"
LIB_PATH = PKG_DIR / "liblcf_hip.so"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# FP64 instantiations of the fitting kernels and the nonparametric band
# pipeline are built without multiply-add contraction, so that they round like
# the reference's unfused FP64 arithmetic.
NO_CONTRACT = {"pso_f64.hip", "svi_f64.hip", "np_kernels.hip", "thermal_kernels.hip", "bb_kernels.hip",
               "finalize_kernels.hip", "cpu_pso.hip"}


def _compile(src: Path) -> Path:
    obj = BUILD_DIR / (src.stem + ".o")
    extra = ["-ffp-contract=off"] if src.name in NO_CONTRACT else []
    cmd = [HIPCC, *CFLAGS, *extra, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile all HIP sources and link liblcf_hip.so; returns its path."""
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
