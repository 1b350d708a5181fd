"""ctypes binding of the oracle library (liblcf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker / baseline, never as the
product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liblcf_oracle.so"

_pd = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)
_pi64 = C.POINTER(C.c_int64)
_pu32 = C.POINTER(C.c_uint32)
_i32, _i64, _u32, _u64, _f64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double

SIGNATURES = {
    "orc_n_params": (_i32, [_i32]),
    "orc_pso_bounds": (_i32, [_i32, _pd, _pd]),
    "orc_pso_bounds_cpu": (_i32, [_i32, _pd, _pd]),
    "orc_population_priors": (_i32, [_i32, _pd, _pd]),
    "orc_prior_params": (_i32, [_i32, _pd, _pd]),
    "orc_eval_batch": (_i32, [_i32, _pd, _i64, _pd, _i64, _pd]),
    "orc_metzger_kn_single": (_f64, [_pd, _f64]),
    "orc_eval_grad": (_i32, [_i32, _pd, _f64, _pd]),
    "orc_log_normal_cdf": (_f64, [_f64]),
    "orc_dlog_normal_cdf": (_f64, [_f64]),
    "orc_erf_approx": (_f64, [_f64]),
    "orc_pso_cost_gpu": (_f64, [_i32, _pd, _pd, _pd, _pd, _pi32, _pd, _i64]),
    "orc_pso_cost_cpu": (_f64, [_i32, _pd, _pd, _pd, _pd, _pi32, _pd, _i64]),
    "orc_pso_cost_grad_cpu": (_i32, [_i32, _pd, _pd, _pd, _pd, _pi32, _pd, _i64, _pd]),
    "orc_philox4x32_10": (None, [_pu32, _pu32, _pu32]),
    "orc_batch_pso": (_i32, [_i32, _i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _i32, _i32, _i32,
                             _u64, _u32, _pd, _pd, _pi32, _i32]),
    "orc_batch_model_select": (_i32, [_i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _pi32, _i32, _i32,
                                      _i32, _i32, _f64, _u32, _pi32, _pd, _pd, _i32]),
    "orc_svi_normals": (_i32, [_u64, _u32, _u32, _u32, _i32, _pd]),
    "orc_svi_elbo_grad": (_f64, [_i32, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _pd, _pd, _pi32, _pd,
                                 _i64, _i32, _pd, _pd]),
    "orc_batch_svi": (_i32, [_i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _pi32, _pd, _pd, _pd, _i32, _i32,
                             _f64, _i32, _u64, _i32, _u32, _pd, _pd, _pd, _pi32, _pd, _i32]),
    "orc_svi_prior_for_model": (_i32, [_i32, _pd, _pd, _pd]),
}

SVI_RNG_PHILOX, SVI_RNG_REFERENCE = 0, 1
SVI_CPU_STOPPING, SVI_INFLATE_SIGMA = 1, 2
SVI_FIXED_STEPS, SVI_EXACT_TAIL, SVI_SHARED_STREAM = 8, 16, 32

_LIB = None


def build() -> Path:
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def n_params(m: int) -> int:
    return lib().orc_n_params(m)


def bounds(m: int, cpu: bool = False):
    n = n_params(m)
    lo, hi = np.zeros(n), np.zeros(n)
    (lib().orc_pso_bounds_cpu if cpu else lib().orc_pso_bounds)(m, _p(lo, C.c_double), _p(hi, C.c_double))
    return lo, hi


def population_priors(m: int):
    n = n_params(m)
    c, w = np.zeros(n), np.zeros(n)
    lib().orc_population_priors(m, _p(c, C.c_double), _p(w, C.c_double))
    return c, w


def prior_params(m: int):
    n = n_params(m)
    c, w = np.zeros(n), np.zeros(n)
    lib().orc_prior_params(m, _p(c, C.c_double), _p(w, C.c_double))
    return c, w


def eval_batch(m: int, params, times) -> np.ndarray:
    p, t = f64(params).ravel(), f64(times).ravel()
    nd = p.size // n_params(m)
    out = np.empty(nd * t.size)
    rc = lib().orc_eval_batch(m, _p(p, C.c_double), nd, _p(t, C.c_double), t.size, _p(out, C.c_double))
    assert rc == 0
    return out


def eval_grad(m: int, params, t: float) -> np.ndarray:
    p = f64(params)
    out = np.zeros(n_params(m))
    rc = lib().orc_eval_grad(m, _p(p, C.c_double), float(t), _p(out, C.c_double))
    if rc != 0:
        raise ValueError("no analytic gradient")
    return out


def metzger_single(params, t: float) -> float:
    p = f64(params)
    return lib().orc_metzger_kn_single(_p(p, C.c_double), float(t))


def log_normal_cdf(x: float) -> float:
    return lib().orc_log_normal_cdf(float(x))


def dlog_normal_cdf(x: float) -> float:
    return lib().orc_dlog_normal_cdf(float(x))


def erf_approx(x: float) -> float:
    return lib().orc_erf_approx(float(x))


def _obs(src: dict, a: int = 0, b: int | None = None):
    b = src["offsets"][-1] if b is None else b
    return (f64(src["times"][a:b]), f64(src["flux"][a:b]), f64(src["obs_var"][a:b]),
            i32(src["is_upper"][a:b]), f64(src["upper_flux"][a:b]))


def pso_cost(m: int, params, t, f, v, up, uf, cpu: bool = False) -> float:
    p = f64(params)
    t, f, v, uf = f64(t), f64(f), f64(v), f64(uf)
    up = i32(up)
    fn = lib().orc_pso_cost_cpu if cpu else lib().orc_pso_cost_gpu
    return fn(m, _p(p, C.c_double), _p(t, C.c_double), _p(f, C.c_double), _p(v, C.c_double),
              _p(up, C.c_int32), _p(uf, C.c_double), t.size)


def pso_cost_grad(m: int, params, t, f, v, up, uf) -> np.ndarray:
    p = f64(params)
    t, f, v, uf = f64(t), f64(f), f64(v), f64(uf)
    up = i32(up)
    g = np.zeros(p.size)
    rc = lib().orc_pso_cost_grad_cpu(m, _p(p, C.c_double), _p(t, C.c_double), _p(f, C.c_double),
                                     _p(v, C.c_double), _p(up, C.c_int32), _p(uf, C.c_double),
                                     t.size, _p(g, C.c_double))
    if rc != 0:
        raise ValueError("no analytic gradient")
    return g


def philox(ctr, key):
    c = np.ascontiguousarray(np.asarray(ctr, dtype=np.uint32))
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32))
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def default_threads() -> int:
    return max(1, min(os.cpu_count() or 1, 16))


def batch_pso(m: int, csr: dict, n_particles: int, max_iters: int, stall_iters: int, seed: int,
              flags: int = 0, n_threads: int | None = None):
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, f, v, up, uf = _obs(csr)
    npar = n_params(m)
    prm = np.zeros((ns, npar))
    cost = np.zeros(ns)
    iters = np.zeros(ns, dtype=np.int32)
    rc = lib().orc_batch_pso(m, ns, _p(off, C.c_int64), _p(t, C.c_double), _p(f, C.c_double),
                             _p(v, C.c_double), _p(up, C.c_int32), _p(uf, C.c_double), n_particles,
                             max_iters, stall_iters, C.c_uint64(seed), flags, _p(prm, C.c_double),
                             _p(cost, C.c_double), _p(iters, C.c_int32),
                             n_threads or default_threads())
    assert rc == 0
    return prm, cost, iters


def batch_model_select(csr: dict, models, n_particles: int, max_iters: int, stall_iters: int,
                       gate: float, flags: int = 0, n_threads: int | None = None):
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, f, v, up, uf = _obs(csr)
    ms = i32(models)
    bm = np.zeros(ns, dtype=np.int32)
    prm = np.zeros((ns, 7))
    cost = np.zeros(ns)
    rc = lib().orc_batch_model_select(ns, _p(off, C.c_int64), _p(t, C.c_double), _p(f, C.c_double),
                                      _p(v, C.c_double), _p(up, C.c_int32), _p(uf, C.c_double),
                                      _p(ms, C.c_int32), ms.size, n_particles, max_iters, stall_iters,
                                      float(gate), flags, _p(bm, C.c_int32), _p(prm, C.c_double),
                                      _p(cost, C.c_double), n_threads or default_threads())
    assert rc == 0
    return bm, prm, cost


def svi_normals(seed: int, job: int, step: int, sample: int, np_: int) -> np.ndarray:
    e = np.zeros(np_)
    assert lib().orc_svi_normals(C.c_uint64(seed), job, step, sample, np_, _p(e, C.c_double)) == 0
    return e


def svi_prior_for_model(m: int, pso):
    p = f64(pso)
    n = n_params(m)
    c, w = np.zeros(n), np.zeros(n)
    lib().orc_svi_prior_for_model(m, _p(p, C.c_double), _p(c, C.c_double), _p(w, C.c_double))
    return c, w


def svi_elbo_grad(m: int, mu, ls, pc, pw, eps, t, f, v, up, uf, grad_mode: int = 0):
    """One-step MC ELBO for noise eps[S, np] and its gradient."""
    npar = n_params(m)
    mu, ls, pc, pw = (f64(x)[:npar].copy() for x in (mu, ls, pc, pw))
    eps = f64(eps).reshape(-1, npar)
    t, f, v, uf = f64(t), f64(f), f64(v), f64(uf)
    up = i32(up)
    gmu, gls = np.zeros(npar), np.zeros(npar)
    e = lib().orc_svi_elbo_grad(m, _p(mu, C.c_double), _p(ls, C.c_double), _p(pc, C.c_double),
                                _p(pw, C.c_double), _p(eps, C.c_double), eps.shape[0], _p(t, C.c_double),
                                _p(f, C.c_double), _p(v, C.c_double), _p(up, C.c_int32), _p(uf, C.c_double),
                                t.size, grad_mode, _p(gmu, C.c_double), _p(gls, C.c_double))
    return e, gmu, gls


def batch_svi(csr: dict, model_ids, pso, pc, pw, n_steps: int, n_samples: int, lr: float,
              rng_mode: int = SVI_RNG_PHILOX, seed: int = 42, grad_mode: int = 0, flags: int = 0,
              n_threads: int | None = None, with_gmax: bool = False):
    """Returns (mu, log_sigma, elbo, steps), plus the per-job largest |gradient|
    handed to Adam (inf if one was non-finite) when ``with_gmax``."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    nj = off.size - 1
    t, f, v, up, uf = _obs(csr)
    mid = i32(model_ids)

    def rows(x):
        a = np.zeros((nj, 7))
        x = np.asarray(x, dtype=np.float64).reshape(nj, -1)
        a[:, :x.shape[1]] = x[:, :7]
        return a

    pso, pc, pw = rows(pso), rows(pc), rows(pw)
    mu, ls, elbo = np.zeros((nj, 7)), np.zeros((nj, 7)), np.zeros(nj)
    steps = np.zeros(nj, dtype=np.int32)
    gmax = np.zeros(nj)
    rc = lib().orc_batch_svi(nj, _p(off, C.c_int64), _p(t, C.c_double), _p(f, C.c_double), _p(v, C.c_double),
                             _p(up, C.c_int32), _p(uf, C.c_double), _p(mid, C.c_int32), _p(pso, C.c_double),
                             _p(pc, C.c_double), _p(pw, C.c_double), n_steps, n_samples, float(lr), rng_mode,
                             C.c_uint64(seed), grad_mode, flags, _p(mu, C.c_double), _p(ls, C.c_double),
                             _p(elbo, C.c_double), _p(steps, C.c_int32), _p(gmax, C.c_double),
                             n_threads or default_threads())
    assert rc == 0
    return (mu, ls, elbo, steps, gmax) if with_gmax else (mu, ls, elbo, steps)


GP_SUBSAMPLE_GPU, GP_SUBSAMPLE_CPU = 0, 1
GP_SELECT_NLML, GP_SELECT_TRAIN_RMS = 0, 1


def dense_gp(t, y, nv, amp: float, ls: float, query):
    """DenseGP fit + (mean, std, dmean) at query, nlml, train_rms; None if not PD."""
    t, y, nv, q = f64(t), f64(y), f64(nv), f64(query)
    mean, std, dm = np.zeros(q.size), np.zeros(q.size), np.zeros(q.size)
    nl, tr = C.c_double(), C.c_double()
    L = lib()
    L.orc_dense_gp.restype = _i32
    L.orc_dense_gp.argtypes = [_pd, _pd, _pd, _i32, _f64, _f64, _pd, _i32, _pd, _pd, _pd,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rc = L.orc_dense_gp(_p(t, C.c_double), _p(y, C.c_double), _p(nv, C.c_double), t.size, amp, ls,
                        _p(q, C.c_double), q.size, _p(mean, C.c_double), _p(std, C.c_double),
                        _p(dm, C.c_double), C.byref(nl), C.byref(tr))
    if rc != 0:
        return None
    return mean, std, dm, nl.value, tr.value


def batch_gp_fit(csr: dict, query, amps, lss, cap: int, subsample_rule: int = GP_SUBSAMPLE_GPU,
                 criterion: int = GP_SELECT_NLML, n_threads: int | None = None):
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    nb = off.size - 1
    t, y, nv = f64(csr["times"]), f64(csr["mags"]), f64(csr["noise_var"])
    q, a, l = f64(query), f64(amps), f64(lss)
    stride = 2 * cap + 4
    state = np.zeros((nb, stride))
    pred, std, dpred = np.zeros((nb, q.size)), np.zeros((nb, q.size)), np.zeros((nb, q.size))
    pobs = np.zeros(t.size)
    L = lib()
    L.orc_batch_gp_fit.restype = _i32
    L.orc_batch_gp_fit.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pd, _i32, _pd, _i32, _pd, _i32, _i32, _i32,
                                   _i32, _pd, _i32, _pd, _pd, _pd, _pd, _i32]
    rc = L.orc_batch_gp_fit(nb, _p(off, C.c_int64), _p(t, C.c_double), _p(y, C.c_double), _p(nv, C.c_double),
                            _p(q, C.c_double), q.size, _p(a, C.c_double), a.size, _p(l, C.c_double), l.size,
                            cap, subsample_rule, criterion, _p(state, C.c_double), stride,
                            _p(pred, C.c_double), _p(std, C.c_double), _p(dpred, C.c_double),
                            _p(pobs, C.c_double), n_threads or default_threads())
    assert rc == 0
    return {"state": state, "pred": pred, "std": std, "dpred": dpred, "pred_obs": pobs}


LAPLACE_FD4, LAPLACE_GRAD_FD = 0, 1


def neg_log_posterior(m: int, p, t, f, v, up, uf) -> float:
    L = lib()
    L.orc_neg_log_posterior.restype = _f64
    L.orc_neg_log_posterior.argtypes = [_i32, _pd, _pd, _pd, _pd, _pi32, _pd, _i64]
    p, t, f, v, uf = f64(p), f64(t), f64(f), f64(v), f64(uf)
    up = i32(up)
    return L.orc_neg_log_posterior(m, _p(p, C.c_double), _p(t, C.c_double), _p(f, C.c_double),
                                   _p(v, C.c_double), _p(up, C.c_int32), _p(uf, C.c_double), t.size)


def symmetric_eigen(mat):
    L = lib()
    L.orc_symmetric_eigen.restype = _i32
    L.orc_symmetric_eigen.argtypes = [_pd, _i32, _pd, _pd]
    a = f64(mat)
    n = a.shape[0]
    ev, vec = np.zeros(n), np.zeros((n, n))
    assert L.orc_symmetric_eigen(_p(a, C.c_double), n, _p(ev, C.c_double), _p(vec, C.c_double)) == 0
    return ev, vec


def laplace_fit(m: int, map_params, t, f, v, up, uf, mode: int = LAPLACE_FD4):
    L = lib()
    L.orc_laplace_fit.restype = _i32
    L.orc_laplace_fit.argtypes = [_i32, _pd, _pd, _pd, _pd, _pi32, _pd, _i64, _i32, _pd,
                                  C.POINTER(C.c_double), _pd]
    p, t, f, v, uf = f64(map_params), f64(t), f64(f), f64(v), f64(uf)
    up = i32(up)
    np_ = n_params(m)
    ls, hess = np.zeros(np_), np.zeros((np_ - 1) ** 2)
    e = C.c_double()
    rc = L.orc_laplace_fit(m, _p(p, C.c_double), _p(t, C.c_double), _p(f, C.c_double), _p(v, C.c_double),
                           _p(up, C.c_int32), _p(uf, C.c_double), t.size, mode, _p(ls, C.c_double),
                           C.byref(e), _p(hess, C.c_double))
    assert rc == 0
    return ls, e.value, hess.reshape(np_ - 1, np_ - 1)


def batch_gp2d_fit(csr: dict, qt, qw, amps, lst, lsw, cap: int, subsample_rule: int = GP_SUBSAMPLE_GPU,
                   criterion: int = GP_SELECT_NLML, n_threads: int | None = None):
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, w, y, nv = (f64(csr[k]) for k in ("times", "waves", "mags", "noise_var"))
    qt, qw, a, lt, lw = f64(qt), f64(qw), f64(amps), f64(lst), f64(lsw)
    stride = 3 * cap + 5
    state = np.zeros((ns, stride))
    pred, std, dpred = np.zeros((ns, qt.size)), np.zeros((ns, qt.size)), np.zeros((ns, qt.size))
    pobs = np.zeros(t.size)
    L = lib()
    L.orc_batch_gp2d_fit.restype = _i32
    L.orc_batch_gp2d_fit.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _i32, _pd, _i32,
                                     _pd, _i32, _i32, _i32, _i32, _pd, _i32, _pd, _pd, _pd, _pd, _i32]
    rc = L.orc_batch_gp2d_fit(ns, _p(off, C.c_int64), _p(t, C.c_double), _p(w, C.c_double), _p(y, C.c_double),
                              _p(nv, C.c_double), _p(qt, C.c_double), _p(qw, C.c_double), qt.size,
                              _p(a, C.c_double), a.size, _p(lt, C.c_double), lt.size, _p(lw, C.c_double), lw.size,
                              cap, subsample_rule, criterion, _p(state, C.c_double), stride, _p(pred, C.c_double),
                              _p(std, C.c_double), _p(dpred, C.c_double), _p(pobs, C.c_double),
                              n_threads or default_threads())
    assert rc == 0
    return {"state": state, "pred": pred, "std": std, "dpred": dpred, "pred_obs": pobs}


def batch_multi_bazin(csr: dict, n_particles: int, max_iters: int, stall_iters: int, seed: int,
                      n_threads: int | None = None) -> dict:
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    t, f, v, up, uf = _obs(csr)
    out = {"best_k": np.zeros(ns, dtype=np.int32), "params": np.zeros((ns, 18)), "cost": np.zeros(ns),
           "bic": np.zeros(ns), "per_k_cost": np.zeros((ns, 4)), "per_k_bic": np.zeros((ns, 4))}
    L = lib()
    L.orc_batch_multi_bazin.restype = _i32
    L.orc_batch_multi_bazin.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _i32, _i32, _i32, _u64, _pi32,
                                        _pd, _pd, _pd, _pd, _pd, _i32]
    rc = L.orc_batch_multi_bazin(ns, _p(off, C.c_int64), _p(t, C.c_double), _p(f, C.c_double), _p(v, C.c_double),
                                 _p(up, C.c_int32), _p(uf, C.c_double), n_particles, max_iters, stall_iters,
                                 C.c_uint64(seed), _p(out["best_k"], C.c_int32), _p(out["params"], C.c_double),
                                 _p(out["cost"], C.c_double), _p(out["bic"], C.c_double),
                                 _p(out["per_k_cost"], C.c_double), _p(out["per_k_bic"], C.c_double),
                                 n_threads or default_threads())
    assert rc == 0
    return out


def multi_bazin_cost(p, k: int, t, f, v, up, uf) -> float:
    L = lib()
    L.orc_multi_bazin_cost.restype = _f64
    L.orc_multi_bazin_cost.argtypes = [_pd, _i32, _pd, _pd, _pd, _pi32, _pd, _i64]
    p, t, f, v, uf = f64(p), f64(t), f64(f), f64(v), f64(uf)
    up = i32(up)
    return L.orc_multi_bazin_cost(_p(p, C.c_double), k, _p(t, C.c_double), _p(f, C.c_double), _p(v, C.c_double),
                                  _p(up, C.c_int32), _p(uf, C.c_double), t.size)


# ---- nonparametric band pipeline (src/nonparametric.rs:1228-1525) ----------
NP_NFEAT = 50
NP_FEATURES = (
    "rise_time", "rise_halfmax", "rise_efold", "decay_efold", "dm15", "decay_halfmax", "t0", "peak_mag",
    "chi2", "baseline_chi2", "n_obs", "fwhm", "rise_rate", "decay_rate", "near_peak_rise_rate",
    "near_peak_decay_rate", "gp_dfdt_now", "gp_dfdt_next", "gp_d2fdt2_now", "gp_predicted_mag_1d",
    "gp_predicted_mag_2d", "gp_time_to_peak", "gp_extrap_slope", "gp_sigma_f", "gp_peak_to_peak",
    "gp_snr_max", "gp_dfdt_max", "gp_dfdt_min", "gp_frac_of_peak", "gp_post_var_mean", "gp_post_var_max",
    "gp_skewness", "gp_kurtosis", "gp_n_inflections", "decay_power_law_index", "decay_power_law_chi2",
    "mag_at_30d", "mag_at_60d", "mag_at_90d", "von_neumann_ratio", "pre_peak_rms",
    "rise_amplitude_over_noise", "post_peak_monotonicity", "n_local_maxima", "n_upper_limits",
    "gp_fit_amp", "gp_fit_lengthscale", "stetson_k", "excess_variance", "cusum_range")
NP_AMPS = (0.3, 1.0, 3.0, 10.0)             # src/nonparametric.rs:1110
NP_LS_FACTORS = (3.0, 6.0, 12.0, 24.0, 48.0)  # src/nonparametric.rs:1111


def batch_nonparametric(csr: dict, t_lo, t_hi, n_grid: int = 50, amps=NP_AMPS, ls_factors=NP_LS_FACTORS,
                        dense_max: int = 100, max_sub: int = 25, n_inducing: int = 30,
                        n_threads: int | None = None) -> dict:
    """process_band over CSR bands of raw magnitudes (times, mags, mag_errs)."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    nb = off.size - 1
    t, m, e = f64(csr["times"]), f64(csr["mags"]), f64(csr["mag_errs"])
    lo, hi, a, lf = f64(t_lo), f64(t_hi), f64(amps), f64(ls_factors)
    assert lo.size == nb and hi.size == nb
    cap = max(dense_max, max_sub)
    stride = 2 * cap + 4
    status = np.zeros(nb, np.int32)
    feat = np.zeros((nb, NP_NFEAT))
    pred, std, dpred = (np.zeros((nb, n_grid)) for _ in range(3))
    state = np.zeros((nb, stride))
    L = lib()
    L.orc_batch_nonparametric.restype = _i32
    L.orc_batch_nonparametric.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _i32, _pd, _i32,
                                          _i32, _i32, _i32, _pi32, _pd, _pd, _pd, _pd, _pd, _i32, _i32]
    rc = L.orc_batch_nonparametric(nb, _p(off, C.c_int64), _p(t, C.c_double), _p(m, C.c_double),
                                   _p(e, C.c_double), _p(lo, C.c_double), _p(hi, C.c_double), n_grid,
                                   _p(a, C.c_double), a.size, _p(lf, C.c_double), lf.size, dense_max, max_sub,
                                   n_inducing, _p(status, C.c_int32), _p(feat, C.c_double), _p(pred, C.c_double),
                                   _p(std, C.c_double), _p(dpred, C.c_double), _p(state, C.c_double), stride,
                                   n_threads or default_threads())
    assert rc == 0
    return {"status": status, "features": feat, "pred": pred, "std": std, "dpred": dpred, "state": state}


def np_curve_metrics(t, m, peak_idx: int) -> np.ndarray:
    t, m = f64(t), f64(m)
    out = np.zeros(11)
    L = lib()
    L.orc_np_curve_metrics.restype = _i32
    L.orc_np_curve_metrics.argtypes = [_pd, _pd, _i32, _i32, _pd]
    assert L.orc_np_curve_metrics(_p(t, C.c_double), _p(m, C.c_double), t.size, peak_idx,
                                  _p(out, C.c_double)) == 0
    return out


def np_raw_features(t, m, e, t_peak: float, peak_mag: float) -> np.ndarray:
    t, m, e = f64(t), f64(m), f64(e)
    out = np.zeros(6)
    L = lib()
    L.orc_np_raw_features.restype = _i32
    L.orc_np_raw_features.argtypes = [_pd, _pd, _pd, _i32, _f64, _f64, _pd]
    assert L.orc_np_raw_features(_p(t, C.c_double), _p(m, C.c_double), _p(e, C.c_double), t.size, t_peak,
                                 peak_mag, _p(out, C.c_double)) == 0
    return out


def np_grid_features(tg, p, sd, peak_idx: int) -> np.ndarray:
    tg, p, sd = f64(tg), f64(p), f64(sd)
    out = np.zeros(14)
    L = lib()
    L.orc_np_grid_features.restype = _i32
    L.orc_np_grid_features.argtypes = [_pd, _pd, _pd, _i32, _i32, _pd]
    assert L.orc_np_grid_features(_p(tg, C.c_double), _p(p, C.c_double), _p(sd, C.c_double), tg.size,
                                  peak_idx, _p(out, C.c_double)) == 0
    return out


def sparse_gp(t, y, nv, amp: float, ls: float, m: int, query):
    """SparseGP (FITC) fit; (mean, std) at query and approx_nlml, None on failure."""
    t, y, nv, q = f64(t), f64(y), f64(nv), f64(query)
    mean, std = np.zeros(q.size), np.zeros(q.size)
    nl = C.c_double()
    L = lib()
    L.orc_sparse_gp.restype = _i32
    L.orc_sparse_gp.argtypes = [_pd, _pd, _pd, _i32, _f64, _f64, _i32, _pd, _i32, _pd, _pd,
                                C.POINTER(C.c_double)]
    rc = L.orc_sparse_gp(_p(t, C.c_double), _p(y, C.c_double), _p(nv, C.c_double), t.size, amp, ls, m,
                         _p(q, C.c_double), q.size, _p(mean, C.c_double), _p(std, C.c_double), C.byref(nl))
    if rc != 0:
        return None
    return mean, std, nl.value


# ---- blackbody colour-temperature fit (src/thermal.rs:222-418) -------------
THERMAL_OUT = ("log_temp_peak", "log_temp_latest", "cooling_rate", "log_temp_peak_err", "cooling_rate_err",
               "chi2", "best_cost", "n_color_obs")


def bb_color_mag(lambda_ref: float, lam: float, temp: float) -> float:
    L = lib()
    L.orc_bb_color_mag.restype = _f64
    L.orc_bb_color_mag.argtypes = [_f64, _f64, _f64]
    return L.orc_bb_color_mag(lambda_ref, lam, temp)


def batch_thermal(pack: dict, seed: int = 42, n_threads: int | None = None) -> dict:
    """fit_thermal over packed sources (see lcf_oracle.h); gp_state is copied."""
    ro, co = (np.ascontiguousarray(np.asarray(pack[k], dtype=np.int64)) for k in ("ref_off", "col_off"))
    ns = ro.size - 1
    st = np.array(pack["gp_state"], dtype=np.float64, copy=True, order="C")
    status = np.zeros(ns, np.int32)
    out = np.zeros((ns, 8))
    used = np.zeros(max(int(co[-1]), 1), np.int32)
    arrs = [f64(pack[k]) for k in ("ref_t", "ref_m", "ref_e", "col_t", "col_m", "col_e", "col_lambda",
                                   "lambda_ref", "duration")]
    L = lib()
    L.orc_batch_thermal.restype = _i32
    L.orc_batch_thermal.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _i32,
                                    _u64, _pi32, _pd, _pi32, _i32]
    p = [_p(a, C.c_double) for a in arrs]
    rc = L.orc_batch_thermal(ns, _p(ro, C.c_int64), p[0], p[1], p[2], _p(co, C.c_int64), p[3], p[4], p[5], p[6],
                             p[7], p[8], _p(st, C.c_double), st.shape[1], seed, _p(status, C.c_int32),
                             _p(out, C.c_double), _p(used, C.c_int32), n_threads or default_threads())
    assert rc == 0
    return {"status": status, "out": out, "col_used": used[:int(co[-1])], "gp_state": st}


def fit_bb_temperature(lam, mag, err):
    """(log10 T, chi2, offset) or None (src/gp2d.rs:275-324)."""
    lam, mag, err = f64(lam), f64(mag), f64(err)
    out = np.zeros(3)
    L = lib()
    L.orc_fit_bb_temperature.restype = _i32
    L.orc_fit_bb_temperature.argtypes = [_pd, _pd, _pd, _i32, _pd]
    rc = L.orc_fit_bb_temperature(_p(lam, C.c_double), _p(mag, C.c_double), _p(err, C.c_double), lam.size,
                                  _p(out, C.c_double))
    return None if rc else tuple(out)


FINALIZE_T0, FINALIZE_ALL, FINALIZE_MAG_CHI2, FINALIZE_SVI_CHI2 = 1, 2, 4, 8
FINALIZE_DEFAULT = FINALIZE_T0 | FINALIZE_ALL | FINALIZE_MAG_CHI2


def finalize(csr: dict, flux_err, mag_obs, peak, model_ids, pso, mu, log_sigma, flags: int = FINALIZE_DEFAULT,
             n_threads: int | None = None) -> dict:
    """profile_t0_refine + profile_refine_all + magnitude chi2 per band
    (src/parametric.rs:2705-3008, 3366-3393); csr holds the BandFitData arrays
    (offsets, times, flux, obs_var, is_upper, upper_flux); mu / log_sigma / pso
    are [n_bands][7]."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    nb = off.size - 1
    arrs = [f64(csr[k]) for k in ("times", "flux", "obs_var")]
    up = i32(csr["is_upper"])
    uf, fe, mo, pk = f64(csr["upper_flux"]), f64(flux_err), f64(mag_obs), f64(peak)
    mids = i32(model_ids)
    pso, mu, ls = (np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(nb, 7)) for a in (pso, mu, log_sigma))
    out_mu, out_ls, chi2 = np.zeros((nb, 7)), np.zeros((nb, 7)), np.zeros(nb)
    L = lib()
    L.orc_finalize.restype = _i32
    L.orc_finalize.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _pd, _pd, _pd, _pi32, _pd, _pd, _pd, _i32,
                               _pd, _pd, _pd, _i32]
    d = lambda a: _p(a, C.c_double)
    rc = L.orc_finalize(nb, _p(off, C.c_int64), d(arrs[0]), d(arrs[1]), d(arrs[2]), _p(up, C.c_int32), d(uf), d(fe),
                        d(mo), d(pk), _p(mids, C.c_int32), d(pso), d(mu), d(ls), flags, d(out_mu), d(out_ls), d(chi2),
                        n_threads or default_threads())
    assert rc == 0
    return {"mu": out_mu, "log_sigma": out_ls, "mag_chi2": chi2}


GP2D_TH_LO = np.log([1e-6, 0.1, 1e-3, 1e-6])
GP2D_TH_HI = np.log([1e6, 1e4, 10.0, 10.0])


def gp2d_nlml_grad(t, w, y, nv, theta):
    """(NLML, gradient[4]) of the 2D GP at theta = ln(a, lt, lw, sn), or None
    when K is not positive definite."""
    t, w, y, nv, th = f64(t), f64(w), f64(y), f64(nv), f64(theta)
    g = np.zeros(4)
    f = C.c_double()
    L = lib()
    L.orc_gp2d_nlml_grad.restype = _i32
    L.orc_gp2d_nlml_grad.argtypes = [_pd, _pd, _pd, _pd, _i32, _pd, _pd, _pd]
    rc = L.orc_gp2d_nlml_grad(_p(t, C.c_double), _p(w, C.c_double), _p(y, C.c_double), _p(nv, C.c_double),
                              t.size, _p(th, C.c_double), C.byref(f), _p(g, C.c_double))
    return None if rc else (f.value, g)


def batch_gp2d_train(csr: dict, theta0, max_iters: int = 100, gtol: float = 0.0,
                     n_threads: int | None = None) -> dict:
    """L-BFGS fit of theta per source (CSR offsets / times / log_waves / mags / noise_var)."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    arrs = [f64(csr[k]) for k in ("times", "log_waves", "mags", "noise_var")]
    th0 = np.ascontiguousarray(np.asarray(theta0, dtype=np.float64).reshape(ns, 4))
    th, g = np.zeros((ns, 4)), np.zeros((ns, 4))
    f = np.zeros(ns)
    it, st, ev = np.zeros(ns, np.int32), np.zeros(ns, np.int32), np.zeros((ns, 2), np.int32)
    L = lib()
    L.orc_batch_gp2d_train.restype = _i32
    L.orc_batch_gp2d_train.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pd, _pd, _i32, _f64, _pd, _pd, _pd, _pi32,
                                       _pi32, _pi32, _i32]
    d = lambda a: _p(a, C.c_double)
    rc = L.orc_batch_gp2d_train(ns, _p(off, C.c_int64), d(arrs[0]), d(arrs[1]), d(arrs[2]), d(arrs[3]), d(th0),
                                max_iters, gtol, d(th), d(f), d(g), _p(it, C.c_int32), _p(st, C.c_int32),
                                _p(ev, C.c_int32), n_threads or default_threads())
    assert rc == 0
    return {"theta": th, "nlml": f, "grad": g, "iters": it, "status": st, "evals": ev}


def batch_fit_gp_2d(csr: dict, qt, qw, n_threads: int | None = None) -> dict:
    """fit_gp_2d per source (offsets / times / waves / mags / noise_var) with
    per-source queries qt / qw [n][nq]."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    ns = off.size - 1
    arrs = [f64(csr[k]) for k in ("times", "waves", "mags", "noise_var")]
    qt, qw = f64(qt).reshape(ns, -1), f64(qw).reshape(ns, -1)
    nq = qt.shape[1]
    hyp, rms, ntr = np.zeros((ns, 3)), np.zeros(ns), np.zeros(ns, np.int32)
    pred, sd = np.zeros((ns, nq)), np.zeros((ns, nq))
    L = lib()
    L.orc_batch_fit_gp_2d.restype = _i32
    L.orc_batch_fit_gp_2d.argtypes = [_i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _pd, _pi32, _pd, _pd,
                                      _i32]
    d = lambda a: _p(a, C.c_double)
    rc = L.orc_batch_fit_gp_2d(ns, _p(off, C.c_int64), d(arrs[0]), d(arrs[1]), d(arrs[2]), d(arrs[3]), d(qt), d(qw),
                               nq, d(hyp), d(rms), _p(ntr, C.c_int32), d(pred), d(sd), n_threads or default_threads())
    assert rc == 0
    return {"hyp": hyp, "train_rms": rms, "n_train": ntr, "pred": pred, "std": sd}


CPSO_SWARM_BAND, CPSO_SWARM_JOINT, CPSO_POLISH, CPSO_SWARM_MULTIBAZIN = 0, 1, 2, 3
PSO_JOB_DTYPE = np.dtype([("kind", "<i4"), ("model", "<i4"), ("band", "<i4"), ("n_bands", "<i4"), ("dim", "<i4"),
                          ("n_particles", "<i4"), ("max_iters", "<i4"), ("stall_iters", "<i4"),
                          ("use_prior", "<i4"), ("fixed_t0", "<i4"), ("seeded", "<i4"), ("stream", "<u4"),
                          ("seed", "<u8"), ("t0_value", "<f8"), ("start_cost", "<f8"), ("lo", "<f8", 64),
                          ("hi", "<f8", 64), ("x0", "<f8", 64)], align=True)


def cpu_pso(csr: dict, jobs, n_threads: int | None = None) -> dict:
    """Swarm / polish jobs (PSO_JOB_DTYPE records) over the BandFitData CSR
    (offsets, times, flux, obs_var, is_upper, upper_flux)."""
    off = np.ascontiguousarray(np.asarray(csr["offsets"], dtype=np.int64))
    t, f, v, uf = (f64(csr[k]) for k in ("times", "flux", "obs_var", "upper_flux"))
    up = i32(csr["is_upper"])
    jobs = np.ascontiguousarray(jobs, dtype=PSO_JOB_DTYPE)
    nj = jobs.size
    x, cost, it = np.zeros((nj, 64)), np.zeros(nj), np.zeros(nj, np.int32)
    L = lib()
    L.orc_cpu_pso.restype = _i32
    L.orc_cpu_pso.argtypes = [_pi64, _pd, _pd, _pd, _pi32, _pd, _i64, C.c_void_p, _i64, _pd, _pd, _pi32, _i32]
    d = lambda a: _p(a, C.c_double)
    rc = L.orc_cpu_pso(_p(off, C.c_int64), d(t), d(f), d(v), _p(up, C.c_int32), d(uf), t.size,
                       jobs.ctypes.data_as(C.c_void_p), nj, d(x), d(cost), _p(it, C.c_int32),
                       n_threads or default_threads())
    assert rc == 0
    return {"x": x, "cost": cost, "iters": it}


FP_RECORD = 130


def fit_parametric(sources, fit_all_models: bool = False, method: str = "Svi", forced_model: int = -1,
                   n_threads: int | None = None) -> list:
    """The CPU parametric pipeline (fit_parametric / fit_parametric_model)
    per source.  ``sources``: list of lists of (times, flux values, flux
    errors) bands, in map order.  Returns per source the result records in
    the reference's order, each a dict with the ParametricBandResult fields
    (``band`` = index of the input band) plus ``route`` (0 per band, 1
    joint)."""
    src_off = [0]
    band_off = [0]
    ts, fs, es = [], [], []
    for bands in sources:
        for t, f, e in bands:
            ts.append(f64(t))
            fs.append(f64(f))
            es.append(f64(e))
            band_off.append(band_off[-1] + len(t))
        src_off.append(len(band_off) - 1)
    cat = lambda v: np.ascontiguousarray(np.concatenate(v)) if v else np.zeros(1)
    t, f, e = cat(ts), cat(fs), cat(es)
    src_off = np.asarray(src_off, dtype=np.int64)
    band_off = np.asarray(band_off, dtype=np.int64)
    nb = int(band_off.size - 1)
    out = np.zeros((max(nb, 1), FP_RECORD))
    L = lib()
    L.orc_fit_parametric.restype = _i32
    L.orc_fit_parametric.argtypes = [_i64, _pi64, _pi64, _pd, _pd, _pd, _i32, _i32, _i32, _pd, _i32]
    rc = L.orc_fit_parametric(len(sources), _p(src_off, C.c_int64), _p(band_off, C.c_int64), _p(t, C.c_double),
                              _p(f, C.c_double), _p(e, C.c_double), int(fit_all_models),
                              0 if method == "Svi" else 1, int(forced_model), _p(out, C.c_double),
                              n_threads or default_threads())
    assert rc == 0
    res = []
    for s in range(len(sources)):
        recs = []
        for b in range(src_off[s], src_off[s + 1]):
            r = out[b]
            if r[0] < 0:
                continue
            m = int(r[1])
            np_ = n_params(m)
            nan_none = lambda v: None if np.isnan(v) else float(v)
            rec = {"band": int(b - src_off[s]), "rank": int(r[0]), "model": m, "pso_params": list(r[2:2 + np_]),
                   "pso_chi2": nan_none(r[9]), "svi_mu": list(r[10:10 + np_]), "svi_log_sigma": list(r[17:17 + np_]),
                   "svi_elbo": nan_none(r[24]), "n_obs": int(r[25]), "mag_chi2": nan_none(r[26]),
                   "per_model_chi2": {k: nan_none(r[27 + k]) for k in range(8) if r[35 + k] > 0.5},
                   "route": int(r[43]), "multi_bazin": None,
                   "per_model_params": ({k: list(r[74 + 7 * k:74 + 7 * k + n_params(k)]) for k in range(8)
                                         if r[35 + k] > 0.5} if fit_all_models else {})}
            if r[44] > 0.5:
                k = int(r[45])
                rec["multi_bazin"] = {"best_k": k, "params": list(r[46:46 + 4 * k + 2]), "cost": float(r[64]),
                                      "bic": float(r[65]), "per_k_cost": list(r[66:70]), "per_k_bic": list(r[70:74])}
            recs.append(rec)
        recs.sort(key=lambda x: x["rank"])
        res.append(recs)
    return res
