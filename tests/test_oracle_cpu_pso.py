"""Pins the oracle's CPU-path swarm engine (CPU only):

  * the per-band cost in the device's summation order against the
    sequentially summed PsoCost oracle (itself pinned by brute force), every
    model, with and without upper limits; the Villar constraint branch;
  * the joint multi-band cost in closed form: B identical bands with zero
    offsets give (B nll + prior(ref) + (B - 1) sum_j (mean_j / sigma_j)^2 / 2)
    / (B n);
  * swarm invariants: positions inside the box, the returned cost is the
    cost of the returned position, a seeded run never ends above the seed's
    cost, determinism, the iteration cap and stall rule;
  * recovery: a noise-free Bazin band fitted to within its noise floor;
  * the polish never raises the cost and lowers it from a perturbed start;
    on the prior-free likelihood it ends where scipy's L-BFGS-B ends.
"""
import math
import numpy as np
import pytest

from conftest import make_reference_test_source

OBS = ("times", "flux", "obs_var", "is_upper", "upper_flux")


def single(src):
    n = src["times"].size
    return {"offsets": np.array([0, n]), **{k: src[k] for k in OBS}}


def job(orc, **kw):
    j = np.zeros((), dtype=orc.PSO_JOB_DTYPE)
    for k, v in kw.items():
        if k in ("lo", "hi", "x0"):
            j[k][:len(v)] = v
        else:
            j[k] = v
    return j


def evaluate(orc, csr, m, p, prior=1):
    """PsoCost at p via a polish job with no iterations."""
    lo, hi = orc.bounds(m, cpu=True)
    lo, hi = np.minimum(lo, p) - 1.0, np.maximum(hi, p) + 1.0
    j = job(orc, kind=orc.CPSO_POLISH, model=m, band=0, dim=len(p), max_iters=0, use_prior=prior,
            lo=lo, hi=hi, x0=p, start_cost=np.inf)
    return orc.cpu_pso(csr, np.array([j]))["cost"][0]


@pytest.mark.parametrize("m", range(7))
def test_cost_matches_sequential_oracle(orc, m):
    src = make_reference_test_source(12.0, 1.3, 70, orc)
    src["is_upper"][::9] = 1
    lo, hi = orc.bounds(m, cpu=True)
    rng = np.random.default_rng(m)
    for _ in range(5):
        p = lo + (hi - lo) * rng.uniform(0.2, 0.8, lo.size)
        want = orc.pso_cost(m, p, *(src[k] for k in OBS), cpu=True)
        got = evaluate(orc, single(src), m, p)
        assert got == pytest.approx(want, rel=1e-12)


def test_villar_constraint_branch(orc):
    src = make_reference_test_source(12.0, 1.3, 30, orc)
    p = np.array([0.0, 0.04, np.log(4.0), 10.0, 0.0, np.log(80.0), -3.0])  # beta tau_fall > 1
    got = evaluate(orc, single(src), 1, p)
    assert got > 1e12 and got == orc.pso_cost(1, p, *(src[k] for k in OBS), cpu=True)


@pytest.mark.parametrize("B", [1, 2, 3])
def test_joint_cost_closed_form(orc, B):
    src = make_reference_test_source(12.0, 1.3, 40, orc)
    n = src["times"].size
    csr = {"offsets": np.arange(B + 1) * n, **{k: np.tile(src[k], B) for k in OBS}}
    ref = np.array([0.05, 0.01, np.log(15.0), 11.0, np.log(2.0), np.log(25.0), np.log(0.05)])
    x0 = np.r_[ref, np.zeros(7 * (B - 1))]
    lo, hi = np.full(7 * B, -1e3), np.full(7 * B, 1e3)
    j = job(orc, kind=orc.CPSO_SWARM_JOINT, model=1, band=0, n_bands=B, dim=7 * B, n_particles=1, max_iters=0,
            seeded=1, lo=lo, hi=hi, x0=x0, seed=1)
    got = orc.cpu_pso(csr, np.array([j]))["cost"][0]
    pc, pw = orc.population_priors(1)
    pop = sum(0.5 * ((ref[k] - pc[k]) / pw[k]) ** 2 for k in range(7))
    nll = n * orc.pso_cost(1, ref, *(src[k] for k in OBS), cpu=True) - pop / n
    ln10 = np.log(10.0)
    mean = np.array([-0.015 * ln10, -0.001, -0.095 * ln10, -1.181, -0.195 * ln10, -0.323 * ln10, -0.025 * ln10])
    sig = np.array([0.08 * ln10, 0.008, 0.25 * ln10, 1.5, 0.35 * ln10, 0.25 * ln10, 0.30 * ln10])
    rel = (B - 1) * np.sum(0.5 * (mean / sig) ** 2)
    assert got == pytest.approx((B * nll + pop + rel) / (B * n), rel=1e-11)


def test_swarm_invariants_and_recovery(orc):
    src = make_reference_test_source(15.0, 1.0, 60, orc)
    csr = single(src)
    lo, hi = orc.bounds(0, cpu=True)
    jobs = []
    truth = np.array([0.0, 0.05, 15.0, 0.0, np.log(15.0), -3.0])
    for seed in (42, 137, 271):
        for seeded in (0, 1):
            jobs.append(job(orc, kind=orc.CPSO_SWARM_BAND, model=0, band=0, dim=6, n_particles=30, max_iters=60,
                            stall_iters=12, use_prior=1, seeded=seeded, lo=lo, hi=hi, x0=truth, seed=seed, stream=3))
    r = orc.cpu_pso(csr, np.array(jobs))
    r2 = orc.cpu_pso(csr, np.array(jobs))
    assert np.array_equal(r["x"], r2["x"]) and np.array_equal(r["cost"], r2["cost"])
    c_truth = evaluate(orc, csr, 0, truth)
    for k in range(len(jobs)):
        x = r["x"][k, :6]
        assert np.all(x >= lo) and np.all(x <= hi)
        assert evaluate(orc, csr, 0, x) == r["cost"][k]
        assert 1 <= r["iters"][k] <= 60
        if jobs[k]["seeded"]:
            assert r["cost"][k] <= c_truth
    assert r["cost"].min() < c_truth + 0.05  # within the noise of the synthetic band


def test_polish(orc):
    src = make_reference_test_source(15.0, 1.0, 60, orc)
    csr = single(src)
    lo, hi = orc.bounds(0, cpu=True)
    p = np.array([0.1, 0.0, 14.0, 0.2, np.log(12.0), -2.5])
    c0 = evaluate(orc, csr, 0, p)
    j = job(orc, kind=orc.CPSO_POLISH, model=0, band=0, dim=6, max_iters=50, use_prior=1, lo=lo, hi=hi, x0=p,
            start_cost=c0)
    r = orc.cpu_pso(csr, np.array([j]))
    assert r["cost"][0] < c0 - 1e-3
    assert evaluate(orc, csr, 0, r["x"][0, :6]) == r["cost"][0]


def _exp(x):
    """IEEE exp: +inf on overflow (math.exp raises instead)."""
    return math.exp(x) if x < 709.0 else (math.exp(x) if x <= 709.78 else math.inf)


def multi_bazin_cost_py(p, k, t, f, v, up, uf):
    """multi_bazin_cost (src/parametric.rs:983-1006) transcribed in plain
    Python: K Bazin components plus the baseline, the censored term for
    upper limits, 1e99 on a non-finite prediction, / max(n, 1)."""
    se2 = math.exp(p[4 * k + 1]) ** 2
    nll = 0.0
    for i in range(len(t)):
        pred = 0.0
        for c in range(k):
            a, tr, tf = math.exp(p[4 * c]), math.exp(p[4 * c + 2]), math.exp(p[4 * c + 3])
            dt = t[i] - p[4 * c + 1]
            pred += a * _exp(-dt / tf) * (1.0 / (1.0 + _exp(-dt / tr)))
        pred += p[4 * k]
        if not math.isfinite(pred):
            return 1e99
        tv = v[i] + se2
        if up[i]:
            z = (uf[i] - pred) / math.sqrt(tv)
            # log_normal_cdf, A&S 7.1.26 erfc (src/parametric.rs:882-901)
            if z > 8.0:
                lz = 0.0
            elif z < -30.0:
                lz = -0.5 * z * z - 0.5 * math.log(2 * math.pi) - math.log(-z)
            else:
                zz = -z / math.sqrt(2.0)
                tt = 1.0 / (1.0 + 0.3275911 * abs(zz))
                poly = tt * (0.254829592 + tt * (-0.284496736 + tt * (1.421413741 + tt * (-1.453152027 + tt * 1.061405429))))
                ec = poly * math.exp(-zz * zz)
                lz = math.log(max(0.5 * ec if zz >= 0.0 else 1.0 - 0.5 * ec, 1e-300))
            nll -= lz
        else:
            d = pred - f[i]
            nll += d * d / tv + math.log(tv)
    return nll / max(len(t), 1)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_multi_bazin_job_cost_brute_force(orc, pkg, k):
    """The MultiBazin swarm job (fit_multi_bazin's swarms): the cost it
    reports for its returned position equals the plain-Python
    multi_bazin_cost there, and the swarm respects its box."""
    rng = np.random.default_rng(40 + k)
    c = pkg.synthetic.ragged_selection_batch(3, seed=60 + k, n_min=20, n_max=60)
    csr = pkg.synthetic.subset(c, np.arange(1))
    t = csr["times"]
    lo, hi = [], []
    for _ in range(k):
        lo += [-3.0, t.min() - 30.0, -2.0, -2.0]
        hi += [3.0, t.max() + 30.0, 5.0, 6.0]
    lo, hi = np.array(lo + [-0.3, -5.0]), np.array(hi + [0.3, 0.0])
    j = job(orc, kind=orc.CPSO_SWARM_MULTIBAZIN, model=k, band=0, dim=4 * k + 2, n_particles=20, max_iters=15,
            stall_iters=15, lo=lo, hi=hi, seed=42)
    o = orc.cpu_pso(csr, np.array([j]))
    x = o["x"][0, :4 * k + 2]
    assert np.all(x >= lo) and np.all(x <= hi)
    want = multi_bazin_cost_py(list(x), k, *(csr[q] for q in OBS))
    assert o["cost"][0] == pytest.approx(want, rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("t0,amp,seed", [(15.0, 1.0, 1), (30.0, 2.0, 2), (8.0, 3.0, 3)])
def test_polish_end_point_matches_scipy_lbfgsb(orc, t0, amp, seed):
    """The polish (lbfgs_refine, src/parametric.rs:1453-1497, restated with a
    backtracking Armijo line search in place of argmin's More-Thuente) on the
    likelihood alone (its gradient, PsoCost::gradient, omits the prior, so
    the prior-free cost is the consistent objective) ends where scipy's
    L-BFGS-B ends on the same objective, gradient and box, from a start
    inside the basin: the same minimum to 1e-9 in cost and 1e-4 in the
    parameters."""
    so = pytest.importorskip("scipy.optimize")
    src = make_reference_test_source(t0, amp, 60, orc)
    csr = single(src)
    obs = [src[k] for k in OBS]
    lo, hi = orc.bounds(0, cpu=True)
    rng = np.random.default_rng(seed)
    truth = np.array([math.log(amp), 0.05, t0, 0.0, math.log(15.0), -3.0])  # the fixture's parameters
    p0 = np.clip(truth + rng.normal(0.0, 0.05, 6), lo + 1e-3, hi - 1e-3)
    c0 = evaluate(orc, csr, 0, p0, prior=0)
    j = job(orc, kind=orc.CPSO_POLISH, model=0, band=0, dim=6, max_iters=200, use_prior=0, lo=lo, hi=hi, x0=p0,
            start_cost=c0)
    r = orc.cpu_pso(csr, np.array([j]))
    x_pol, f_pol = r["x"][0, :6], r["cost"][0]

    def fg(p):
        return evaluate(orc, csr, 0, p, prior=0), orc.pso_cost_grad(0, p, *obs)

    ref = so.minimize(fg, p0, jac=True, method="L-BFGS-B", bounds=list(zip(lo, hi)),
                      options={"ftol": 1e-15, "gtol": 1e-10, "maxiter": 1000})
    assert f_pol < c0
    assert f_pol == pytest.approx(ref.fun, rel=1e-9, abs=1e-12), (f_pol, ref.fun)
    assert np.allclose(x_pol, ref.x, rtol=1e-4, atol=1e-4), (x_pol, ref.x)
