"""The C ABI driven from C (needs an MI355X).

tests/c_driver/lcf_driver (built by build() with gcc against
include/lcf_hip.h) makes the calls of the crate's HIP backend module
(lightcurve-fitting_amd/rust/gpu_hip.rs) in that module's order and argument
packing: context, context-free batch upload (the reference's
GpuBatchData::new(&[BatchSource]) signature), n_obs_per_source, eval_batch,
batch_pso, batch_model_select, batch_all_models, batch_pso_multi_bazin,
batch_gp_fit, batch_gp_2d and batch_svi_fit.  Every output equals the
Python binding's on the same inputs (one library, deterministic kernels:
bit-identical), and eval_batch equals the oracle to 1e-12.

Also: a batch uploaded without a context serves every context on its device,
with results identical to a batch uploaded through a context.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "c_driver" / "lcf_driver"

ARGS = dict(model=1, n_particles=32, max_iters=30, stall=10, seed=42, gate=2.0, svi_steps=120, svi_samples=4,
            lr=0.01, gp_sub=25, g2_sub=40)


def _put(d: Path, name: str, a, dtype):
    np.ascontiguousarray(np.asarray(a, dtype=dtype)).tofile(d / name)


def _get(d: Path, name: str, dtype):
    return np.fromfile(d / f"out_{name}", dtype=dtype)


def _inputs(pkg, orc):
    rng = np.random.default_rng(11)
    csr = pkg.synthetic.ragged_selection_batch(10, seed=21, n_min=12, n_max=90)
    ns = csr["offsets"].size - 1
    lo, hi = orc.bounds(ARGS["model"])
    ep = rng.uniform(lo, hi, size=(3, len(lo)))
    et = np.linspace(-20.0, 120.0, 37)
    bands, g2 = [], []
    for n in (6, 18, 40, 75):
        t = np.sort(rng.uniform(0.0, 100.0, n))
        bands.append(pkg.GpBandInput(t, 19.0 + np.sin(t / 17.0) + 0.05 * rng.normal(size=n), np.full(n, 0.0025)))
    for n in (9, 30, 64):
        t = np.sort(rng.uniform(0.0, 80.0, n))
        w = np.log10(rng.choice([4770.0, 6231.0, 7625.0], n))
        g2.append(pkg.Gp2dInput(t, w, 18.5 + np.cos(t / 20.0) + 0.3 * (w - 3.8), np.full(n, 0.004)))
    gq = np.linspace(0.0, 100.0, 50)
    g2qt = np.repeat(np.linspace(0.0, 80.0, 10), 3)
    g2qw = np.tile(np.log10([4770.0, 6231.0, 7625.0]), 10)
    svi = []
    for s in range(ns):
        m = 0
        p = [0.0, 0.0, 30.0, 1.0, 3.0, -3.0]
        c, w = pkg.svi_prior_for_model(m, p)
        svi.append(pkg.SviBatchInput(m, p, 5, c, w))
    return csr, ep, et, bands, gq, g2, g2qt, g2qw, svi


def _write(d: Path, csr, ep, et, bands, gq, g2, g2qt, g2qw, svi):
    _put(d, "off.i64", csr["offsets"], np.int64)
    for k, n, dt in (("times", "t.f64", np.float64), ("flux", "f.f64", np.float64), ("obs_var", "v.f64", np.float64),
                     ("is_upper", "up.i32", np.int32), ("upper_flux", "uf.f64", np.float64)):
        _put(d, n, csr[k], dt)
    _put(d, "eval_params.f64", ep.ravel(), np.float64)
    _put(d, "eval_times.f64", et, np.float64)
    goff = np.concatenate([[0], np.cumsum([len(b.times) for b in bands])])
    _put(d, "gp_off.i64", goff, np.int64)
    _put(d, "gp_t.f64", np.concatenate([b.times for b in bands]), np.float64)
    _put(d, "gp_m.f64", np.concatenate([b.mags for b in bands]), np.float64)
    _put(d, "gp_nv.f64", np.concatenate([b.noise_var for b in bands]), np.float64)
    _put(d, "gp_q.f64", gq, np.float64)
    _put(d, "gp_amp.f64", [0.3, 1.0, 3.0, 10.0], np.float64)
    _put(d, "gp_ls.f64", [4.0, 8.0, 16.0, 32.0], np.float64)
    hoff = np.concatenate([[0], np.cumsum([len(s.times) for s in g2])])
    _put(d, "g2_off.i64", hoff, np.int64)
    for k, n in (("times", "g2_t.f64"), ("waves", "g2_w.f64"), ("mags", "g2_m.f64"), ("noise_var", "g2_nv.f64")):
        _put(d, n, np.concatenate([getattr(s, k) for s in g2]), np.float64)
    _put(d, "g2_qt.f64", g2qt, np.float64)
    _put(d, "g2_qw.f64", g2qw, np.float64)
    _put(d, "g2_amp.f64", [0.3, 1.0, 3.0], np.float64)
    _put(d, "g2_lst.f64", [10.0, 30.0], np.float64)
    _put(d, "g2_lsw.f64", [0.1, 0.3], np.float64)
    ns = len(svi)
    pso, pc, pw = np.zeros((ns, 7)), np.zeros((ns, 7)), np.zeros((ns, 7))
    for i, inp in enumerate(svi):
        pso[i, :6], pc[i, :6], pw[i, :6] = inp.pso_params, inp.prior_centers, inp.prior_widths
    _put(d, "svi_ids.i32", [i.model_id for i in svi], np.int32)
    _put(d, "svi_se.i32", [i.se_idx for i in svi], np.int32)
    _put(d, "svi_pso.f64", pso, np.float64)
    _put(d, "svi_pc.f64", pc, np.float64)
    _put(d, "svi_pw.f64", pw, np.float64)
    a = ARGS
    (d / "args.txt").write_text(f"{a['model']} {a['n_particles']} {a['max_iters']} {a['stall']} {a['seed']} "
                                f"{a['gate']!r} {a['svi_steps']} {a['svi_samples']} {a['lr']!r} {a['gp_sub']} "
                                f"{a['g2_sub']}\n")


def _same(a, b):
    a, b = np.asarray(a, dtype=np.float64).ravel(), np.asarray(b, dtype=np.float64).ravel()
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def test_c_driver_matches_binding_and_oracle(pkg, orc, tmp_path):
    assert DRIVER.exists(), "tests/c_driver/lcf_driver is built by build()"
    csr, ep, et, bands, gq, g2, g2qt, g2qw, svi = _inputs(pkg, orc)
    _write(tmp_path, csr, ep, et, bands, gq, g2, g2qt, g2qw, svi)
    print("c driver: running", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path)], capture_output=True, text=True, timeout=180)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("done"), r.stderr
    a = ARGS
    ctx = pkg.GpuContext(0)
    data = pkg.GpuBatchData.new([pkg.BatchSource(*(csr[k][csr["offsets"][s]:csr["offsets"][s + 1]]
                                                   for k in ("times", "flux", "obs_var", "is_upper", "upper_flux")))
                                 for s in range(csr["offsets"].size - 1)])
    ns = data.n_sources
    # n_obs_per_source, last entry out of range -> 1
    assert list(_get(tmp_path, "n_obs.i32", np.int32)) == [data.n_obs_per_source(s) for s in range(ns + 1)]
    assert _get(tmp_path, "n_obs.i32", np.int32)[-1] == 1
    # eval_batch: binding and oracle
    m = pkg.GpuModelName(a["model"])
    ev = ctx.eval_batch(m, ep.ravel(), et, 3)
    assert _same(_get(tmp_path, "eval.f64", np.float64), ev)
    assert np.allclose(_get(tmp_path, "eval.f64", np.float64), orc.eval_batch(a["model"], ep, et),
                       rtol=1e-12, atol=1e-12)
    # batch_pso / batch_model_select / batch_all_models
    pso = ctx.batch_pso(m, data, a["n_particles"], a["max_iters"], a["stall"], a["seed"])
    pp = _get(tmp_path, "pso_p.f64", np.float64).reshape(ns, 7)
    assert _same(_get(tmp_path, "pso_c.f64", np.float64), [r.cost for r in pso])
    assert all(_same(pp[s, :m.n_params()], pso[s].params) for s in range(ns))
    sel = ctx.batch_model_select(data, a["n_particles"], a["max_iters"], a["stall"], a["gate"])
    assert list(_get(tmp_path, "sel_m.i32", np.int32)) == [int(x[0]) for x in sel]
    assert _same(_get(tmp_path, "sel_c.f64", np.float64), [x[1].cost for x in sel])
    allm = ctx.batch_all_models(data, a["n_particles"], a["max_iters"], a["stall"])
    ac = _get(tmp_path, "all_c.f64", np.float64).reshape(len(pkg.ALL_GPU_MODELS), ns)
    for k, (mm, rows) in enumerate(allm):
        assert _same(ac[k], [r.cost for r in rows]), mm
    # batch_pso_multi_bazin
    mb = ctx.batch_pso_multi_bazin(data, None, a["n_particles"], a["max_iters"], a["stall"], a["seed"])
    assert list(_get(tmp_path, "mb_k.i32", np.int32)) == [r.best_k for r in mb]
    assert _same(_get(tmp_path, "mb_b.f64", np.float64), [r.bic for r in mb])
    assert _same(_get(tmp_path, "mb_kc.f64", np.float64), [r.per_k_cost for r in mb])
    # batch_gp_fit
    gp = ctx.batch_gp_fit(bands, gq, [0.3, 1.0, 3.0, 10.0], [4.0, 8.0, 16.0, 32.0], a["gp_sub"])
    assert _same(_get(tmp_path, "gp_pred.f64", np.float64), [b.pred_grid for b in gp])
    assert _same(_get(tmp_path, "gp_std.f64", np.float64), [b.std_grid for b in gp])
    assert _same(_get(tmp_path, "gp_po.f64", np.float64), np.concatenate([b.pred_at_obs for b in gp]))
    # batch_gp_2d
    g2o = ctx.batch_gp_2d(g2, g2qt, g2qw, [0.3, 1.0, 3.0], [10.0, 30.0], [0.1, 0.3], a["g2_sub"])
    assert _same(_get(tmp_path, "g2_pred.f64", np.float64), [s.pred_grid for s in g2o])
    st = _get(tmp_path, "g2_state.f64", np.float64).reshape(len(g2), 125)
    assert [st[s, 124] >= 3 for s in range(len(g2))] == [o.success for o in g2o]
    assert _same(st[:, 120][[o.success for o in g2o]], [o.amp for o in g2o if o.success])
    # batch_svi_fit
    so = ctx.batch_svi_fit(data, svi, a["svi_steps"], a["svi_samples"], a["lr"])
    mu = _get(tmp_path, "svi_mu.f64", np.float64).reshape(ns, 7)
    assert all(_same(mu[s, :6], so[s].mu) for s in range(ns))
    assert _same(_get(tmp_path, "svi_elbo.f64", np.float64), [o.elbo for o in so])
    ctx.close()


def test_context_free_batch_serves_every_context_on_its_device(pkg):
    csr = pkg.synthetic.ragged_selection_batch(16, seed=4, n_min=10, n_max=120)
    a, b = pkg.GpuContext(0), pkg.GpuContext(0)
    free = pkg.GpuBatchData(None, csr=csr)      # lcf_batch_create_current
    bound = pkg.GpuBatchData(a, csr=csr)        # lcf_batch_create through context a
    ra = a.fit_models(free, [0, 1], 32, 40, 15, seed=7)
    rb = b.fit_models(free, [0, 1], 32, 40, 15, seed=7)
    rc = b.fit_models(bound, [0, 1], 32, 40, 15, seed=7)
    for k in ("best_model", "best_cost", "model_cost"):
        assert np.array_equal(ra[k], rb[k]) and np.array_equal(ra[k], rc[k]), k
    assert free.n_obs_per_source(3) == bound.n_obs_per_source(3) == int(np.diff(csr["offsets"])[3])
    free.close(), bound.close(), a.close(), b.close()
