"""Shared fixtures: package loader, oracle binding, the `gpu` marker."""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG_DIR = ROOT / "lightcurve-fitting_amd"
sys.path.insert(0, str(ROOT / "oracle"))


def load_package():
    """Import lightcurve-fitting_amd/ (hyphenated directory) by path."""
    name = "lightcurve_fitting_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the native library")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(pkg):
    """One HIP context for the whole GPU session (FP32, persistent schedule)."""
    c = pkg.GpuContext(0)
    yield c
    c.close()


def make_reference_test_source(t0: float, amplitude: float, n_obs: int, orc):
    """Bazin source of the reference's GPU tests (tests/test_gpu.rs:54-70):
    times t0-10 + i*40/n, true params [ln amp, 0.05, t0, ln 1, ln 15, -3],
    deterministic 'noise' 0.05*(0.5 sin(1.37 i) + 0.1), obs_var 0.05^2+1e-10."""
    times = np.array([t0 - 10.0 + i * 40.0 / n_obs for i in range(n_obs)])
    true = np.array([np.log(amplitude), 0.05, t0, np.log(1.0), np.log(15.0), -3.0])
    clean = orc.eval_batch(0, true, times)
    noise = 0.05
    flux = clean + noise * (np.sin(np.arange(n_obs) * 1.37) * 0.5 + 0.1)
    return {
        "times": times, "flux": flux, "obs_var": np.full(n_obs, noise * noise + 1e-10),
        "is_upper": np.zeros(n_obs, dtype=np.int32), "upper_flux": np.zeros(n_obs),
    }


def stack_sources(srcs):
    off = [0]
    for s in srcs:
        off.append(off[-1] + len(s["times"]))
    out = {"offsets": np.asarray(off, dtype=np.int64)}
    for k in ("times", "flux", "obs_var", "is_upper", "upper_flux"):
        dt = np.int32 if k == "is_upper" else np.float64
        out[k] = np.concatenate([np.asarray(s[k], dtype=dt) for s in srcs]) if srcs else np.zeros(0, dt)
    return out
