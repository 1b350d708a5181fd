"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/lcf_hip.h declares, and the host-side argument checks fail
cleanly (no device work is started)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "lcf_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lcf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_entry_points():
    syms = declared_symbols()
    for s in ("lcf_ctx_create", "lcf_eval_batch", "lcf_batch_create", "lcf_batch_pso",
              "lcf_batch_model_select", "lcf_batch_all_models", "lcf_batch_n_obs"):
        assert s in syms


def test_library_exports_every_declared_symbol(pkg):
    lib = C.CDLL(str(pkg.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header(pkg):
    from lightcurve_fitting_amd import _lib

    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_null_and_bad_arguments_rejected_without_device(pkg):
    from lightcurve_fitting_amd import _lib

    lib = _lib.load()
    assert lib.lcf_ctx_create(0, None) == -1
    assert lib.lcf_ctx_set_precision(None, 0) == -1
    assert lib.lcf_batch_n_obs(None, 0) == 1          # reference: 1 when out of range
    assert lib.lcf_batch_n_sources(None) == 0
    h = C.c_void_p()
    # device index beyond the device count (there is none here)
    assert lib.lcf_ctx_create(10_000, C.byref(h)) == -1 and not h.value
    # the context-free batch upload checks its CSR arrays before it looks
    # for a device, and reports through lcf_last_error(NULL)
    off = (C.c_int64 * 3)(0, 2, 1)
    assert lib.lcf_batch_create_current(2, off, None, None, None, None, None, C.byref(h)) == -1
    assert lib.lcf_last_error(None) == b"offsets not non-decreasing at 1" and not h.value
    assert lib.lcf_batch_create_current(-1, off, None, None, None, None, None, C.byref(h)) == -1
    assert b"invalid source count" in lib.lcf_last_error(None)


def test_c_driver_compiles_against_the_header(tmp_path):
    """tests/c_driver/lcf_driver.c, the C caller that makes the crate module's
    calls, is plain C99 against include/lcf_hip.h (gcc, warnings as errors)."""
    import subprocess

    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", "-c",
                        str(ROOT / "tests" / "c_driver" / "lcf_driver.c"), "-o", str(tmp_path / "d.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_model_metadata_mirrors_reference(pkg):
    M = pkg.GpuModelName
    assert [m.n_params() for m in M] == [6, 7, 7, 5, 5, 5, 6, 5]
    assert [int(m) for m in pkg.ALL_GPU_MODELS] == [0, 3, 2, 6, 1, 4, 5, 7]
    lo, hi = M.Villar.pso_bounds()
    assert lo == [-1.5, -0.02, 1.0, -60.0, -1.0, 1.0, -6.0] and hi[-1] == 0.0


def test_kernel_bounds_tables_match_oracle(orc):
    """The __constant__ tables in model_tables.hpp carry the same numbers as
    the oracle (GPU bounds; the CPU bounds differ in the Bazin baseline)."""
    src = (ROOT / "lightcurve-fitting_amd" / "csrc" / "model_tables.hpp").read_text()

    def table(name):
        m = re.search(name + r"\[8\]\[7\] = \{(.*?)\};", src, re.S)
        rows = re.findall(r"\{([^{}]*)\}", m.group(1))
        return [[eval(x.replace("LN10", "2.30258509299404568402")) for x in r.split(",")] for r in rows]

    lo, hi = table("c_lo"), table("c_hi")
    pc, pw = table("c_pc"), table("c_pw")
    for m in range(8):
        n = orc.n_params(m)
        olo, ohi = orc.bounds(m)
        assert lo[m][:n] == pytest.approx(list(olo)) and hi[m][:n] == pytest.approx(list(ohi))
        c, w = orc.population_priors(m)
        assert pc[m][:n] == pytest.approx(list(c)) and pw[m][:n] == pytest.approx(list(w))
        clo, chi = orc.bounds(m, cpu=True)
        diff = [j for j in range(n) if (clo[j], chi[j]) != (olo[j], ohi[j])]
        assert diff == ([1] if m == 0 else [])
    bb = float(re.search(r"BAZIN_B_CPU = ([0-9.]+);", src).group(1))
    assert (-bb, bb) == tuple(x[1] for x in orc.bounds(0, cpu=True))
