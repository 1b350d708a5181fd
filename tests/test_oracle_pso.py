"""Pins the oracle's PSO cost and batched swarm (CPU only).

  * the GPU-path cost against a brute-force sum on tiny inputs (including
    upper limits, an empty band and a non-finite prediction);
  * the CPU-path cost's Villar constraint branch and prior scaling;
  * swarm invariants: bounds, reported cost = cost(reported params), the
    reference GPU stall rule (a swarm always runs min(stall, max_iters)
    iterations because its baseline starts at +inf, src/gpu_cuda.rs:418-552),
    determinism, independence from other sources' data;
  * the reference's own GPU PSO acceptance tests (tests/test_gpu.rs:72-146)
    run against the oracle;
  * model-selection gate semantics (src/gpu_cuda.rs:573-620).
"""
import math

import numpy as np
import pytest

from conftest import make_reference_test_source, stack_sources


def brute_cost(orc, m, p, src):
    """sum[(pred-f)^2/v + ln v] - sum_up ln Phi(z), / n, + prior / n^2.

    Returns (cost, tol) where tol bounds the effect of the oracle's
    Abramowitz & Stegun erfc (|error| <= 1.5e-7 absolute) on ln Phi."""
    t, f, v, up, uf = (np.asarray(src[k]) for k in ("times", "flux", "obs_var", "is_upper", "upper_flux"))
    n = t.size
    np_ = orc.n_params(m)
    tol = 0.0
    if n == 0:
        nll = 1e99
    else:
        pred = np.array([orc.metzger_single(p, x) for x in t]) if m == 7 else orc.eval_batch(m, p, t)
        if not np.all(np.isfinite(pred)):
            nll = 1e99
        else:
            se2 = math.exp(p[np_ - 1]) ** 2
            s = 0.0
            for i in range(n):
                tv = v[i] + se2
                if up[i]:
                    z = (uf[i] - pred[i]) / math.sqrt(tv)
                    phi = 0.5 * math.erfc(-z / math.sqrt(2))
                    if phi < 1e-250:
                        return None, None  # deep tail: A&S floors Phi, no exact comparison
                    s -= math.log(phi)
                    tol += 7.5e-8 / phi / n
                else:
                    s += (pred[i] - f[i]) ** 2 / tv + math.log(tv)
            nll = s / n
    c, w = orc.population_priors(m)
    pri = sum(0.5 * ((p[j] - c[j]) / w[j]) ** 2 for j in range(np_) if w[j] > 0)
    nn = max(n, 1)
    return nll + pri / nn ** 2, tol


def tiny_source(seed, n=5, upper=(1, 3)):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 40, n))
    up = np.zeros(n, dtype=np.int32)
    for i in upper:
        if i < n:
            up[i] = 1
    return {"times": t, "flux": rng.uniform(0.1, 1.0, n), "obs_var": rng.uniform(1e-3, 1e-2, n),
            "is_upper": up, "upper_flux": rng.uniform(0.05, 0.3, n)}


@pytest.mark.parametrize("m", range(8))
def test_gpu_cost_brute_force(orc, m):
    rng = np.random.default_rng(m)
    lo, hi = orc.bounds(m)
    for trial in range(6):
        src = tiny_source(trial + 10 * m, n=3 + trial)
        p = lo + rng.uniform(0.2, 0.8, lo.size) * (hi - lo)
        got = orc.pso_cost(m, p, src["times"], src["flux"], src["obs_var"], src["is_upper"], src["upper_flux"])
        want, tol = brute_cost(orc, m, p, src)
        if want is None:
            continue
        assert got == pytest.approx(want, rel=1e-12, abs=1e-12 + 1.01 * tol), (m, trial)


def test_cost_empty_and_nonfinite(orc):
    p = np.array([0.0, 0.0, 10.0, 1.0, 2.0, -3.0])
    e = np.zeros(0)
    c = orc.pso_cost(0, p, e, e, e, np.zeros(0, np.int32), e)
    assert c == 1e99
    # Bazin with t0 = 100, tau_fall = e^-2 at t = 0: e^{740} overflows in FP64
    p = np.array([0.0, 0.0, 100.0, -2.0, -2.0, -3.0])
    src = tiny_source(1, n=4, upper=())
    src["times"][0] = 0.0
    c = orc.pso_cost(0, p, src["times"], src["flux"], src["obs_var"], src["is_upper"], src["upper_flux"])
    assert c == 1e99


def test_cpu_cost_villar_constraint_and_prior_scaling(orc):
    src = tiny_source(3, n=6, upper=())
    # beta * gamma > 1 violates c1 = max(gamma*beta - 1, 0)
    p = np.array([0.0, 0.05, 4.0, 0.0, 1.0, 2.0, -3.0])
    gamma = math.exp(4.0)
    c = orc.pso_cost(1, p, src["times"], src["flux"], src["obs_var"], src["is_upper"], src["upper_flux"], cpu=True)
    assert c >= 1e12 + 1e4 * (gamma * 0.05 - 1.0) - 1e-3
    # GPU path has no constraint
    g = orc.pso_cost(1, p, src["times"], src["flux"], src["obs_var"], src["is_upper"], src["upper_flux"])
    assert g < 1e6
    # Bazin inside bounds: CPU cost == GPU cost (same NLL/n + prior/n^2)
    p = np.array([0.0, 0.1, 10.0, 1.0, 2.5, -3.0])
    a = orc.pso_cost(0, p, src["times"], src["flux"], src["obs_var"], src["is_upper"], src["upper_flux"], cpu=True)
    b = orc.pso_cost(0, p, src["times"], src["flux"], src["obs_var"], src["is_upper"], src["upper_flux"])
    assert a == pytest.approx(b, rel=1e-14)


def test_cost_gradient_fd(orc):
    src = tiny_source(5, n=12, upper=(2, 7))
    args = [src[k] for k in ("times", "flux", "obs_var", "is_upper", "upper_flux")]
    for m, p in [(0, [0.1, 0.05, 12.0, 1.0, 2.5, -2.0]), (1, [0.1, 0.01, 2.0, 8.0, 1.0, 2.5, -2.0]),
                 (4, [0.2, 3.0, 2.5, 2.0, -2.0])]:
        p = np.asarray(p, float)
        g = orc.pso_cost_grad(m, p, *args)
        # the gradient omits the prior term (src/parametric.rs:1383-1446)
        c, w = orc.population_priors(m)
        n = len(args[0])
        gp = (p - c) / w ** 2 / n ** 2
        for j in range(p.size):
            h = 1e-6 * max(abs(p[j]), 1.0)
            pp, pm = p.copy(), p.copy()
            pp[j] += h
            pm[j] -= h
            fd = (orc.pso_cost(m, pp, *args, cpu=True) - orc.pso_cost(m, pm, *args, cpu=True)) / (2 * h)
            assert g[j] + gp[j] == pytest.approx(fd, rel=2e-3, abs=2e-6), (m, j)


@pytest.fixture(scope="module")
def ref_batch(orc):
    srcs = [make_reference_test_source(10.0 + i, 1.0, 50, orc) for i in range(10)]
    return stack_sources(srcs)


def test_reference_pso_acceptance(orc, ref_batch):
    # gpu_batch_pso_finds_reasonable_cost: Bazin, 40 particles, 50 iters, stall 10
    prm, cost, iters = orc.batch_pso(0, ref_batch, 40, 50, 10, 42)
    assert np.all(np.isfinite(cost)) and np.all(cost < 10.0)
    # reference stall rule: exactly stall_iters iterations
    assert np.all(iters == 10)
    lo, hi = orc.bounds(0)
    assert np.all(prm >= lo) and np.all(prm <= hi)


def test_reference_pso_many_sources(orc):
    srcs = [make_reference_test_source(5.0 + i * 0.3, 0.5 + i * 0.01, 30 + i % 20, orc) for i in range(100)]
    prm, cost, _ = orc.batch_pso(0, stack_sources(srcs), 40, 50, 10, 42)
    assert (np.isfinite(cost) & (cost < 100.0)).sum() > 90


def test_reported_cost_is_cost_of_reported_params(orc, ref_batch):
    for m in (0, 1, 3):
        prm, cost, _ = orc.batch_pso(m, ref_batch, 24, 30, 30, 7)
        off = ref_batch["offsets"]
        for s in range(off.size - 1):
            a, b = off[s], off[s + 1]
            c = orc.pso_cost(m, prm[s], *(ref_batch[k][a:b] for k in ("times", "flux", "obs_var", "is_upper", "upper_flux")))
            assert c == cost[s]


def test_stall_rule_variants(orc, ref_batch):
    _, _, it = orc.batch_pso(0, ref_batch, 16, 40, 100, 3)
    assert np.all(it == 40)
    _, _, it = orc.batch_pso(0, ref_batch, 16, 40, 0, 3)
    assert np.all(it == 1)
    _, c_ref, it = orc.batch_pso(0, ref_batch, 16, 200, 5, 3, flags=1)
    assert np.all(it <= 200) and np.all(it >= 6) and np.any(it < 200)


def test_determinism_and_source_independence(orc, ref_batch):
    a = orc.batch_pso(2, ref_batch, 20, 15, 15, 11)
    b = orc.batch_pso(2, ref_batch, 20, 15, 15, 11, n_threads=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # change every other source's data: source 3's result must not move
    other = {k: v.copy() for k, v in ref_batch.items()}
    off = other["offsets"]
    for s in range(off.size - 1):
        if s != 3:
            other["flux"][off[s]:off[s + 1]] *= 1.7
    c = orc.batch_pso(2, other, 20, 15, 15, 11)
    assert np.array_equal(a[0][3], c[0][3]) and a[1][3] == c[1][3]
    # a different seed gives a different swarm
    d = orc.batch_pso(2, ref_batch, 20, 15, 15, 12)
    assert not np.array_equal(a[0], d[0])


def test_model_select_gate(orc, ref_batch):
    models = [0, 3, 2, 6, 1, 4, 5]
    per = {m: orc.batch_pso(m, ref_batch, 16, 12, 12, 42) for m in models}
    # gate = +inf: nobody needs more, Bazin everywhere
    bm, prm, cost = orc.batch_model_select(ref_batch, models, 16, 12, 12, math.inf)
    assert np.all(bm == 0) and np.array_equal(cost, per[0][1])
    # gate = -inf: everyone continues; best = first strict minimum in list order
    bm, prm, cost = orc.batch_model_select(ref_batch, models, 16, 12, 12, -math.inf)
    allc = np.stack([per[m][1] for m in models])
    first = np.argmin(allc, axis=0)
    assert np.array_equal(bm, np.asarray(models)[first])
    assert np.array_equal(cost, allc.min(axis=0))
