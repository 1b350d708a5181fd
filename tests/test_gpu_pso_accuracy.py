"""Accuracy of the FP32 model-selection path at the shape of the headline
workload (64 particles x 200 iterations, the Bazin / Villar / TDE / Arnett /
Magnetar set, bands of 10-500 points log-uniform, 1-3 bands per source)
against FP64.

The FP64 kernels reproduce the oracle's swarm exactly (test_gpu_pso.py), so
the FP64 kernel run stands in for the oracle on 2000 sources; the FP32 costs
are re-evaluated by the oracle itself.  Measured on MI355X (3932 bands):

  * FP32 and FP64 (same seed) pick the same model for 97.4 % of bands; two
    FP64 runs that differ only in the seed agree on 67.7 % -- the swarm at
    this budget is far from converged, and its own seed-to-seed spread is
    what limits agreement, not the FP32 arithmetic;
  * the FP32 best cost is worse than FP64's by more than 1e-4 relative on
    12.0 % of bands and better on 10.5 %: a symmetric spread (FP32 finds
    optima as good as FP64's), which a polish of both end points by L-BFGS-B
    does not remove (different basins, not unconverged minima);
  * the cost FP32 reports for its returned parameters is the reference's
    FP64 cost there (the swarm re-costs its best position in FP64; before
    that, its FP32 cost agreed to 1.1e-6, except where the reference's cost
    jumps at a phase of 709.78 d, which FP32 times cannot place).
The selection rule is the reference's: strict '<' over models in order
(src/gpu_cuda.rs:595-616).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODELS = [0, 1, 2, 3, 4]
KEYS = ("times", "flux", "obs_var", "is_upper", "upper_flux")


@pytest.fixture(scope="module")
def runs(pkg):
    csr = pkg.synthetic.ragged_selection_batch(2000, seed=4242, n_min=10, n_max=500)
    out = {}
    for name, prec, seed in (("fp32", pkg.FP32, 42), ("fp64", pkg.FP64, 42), ("fp64_seed43", pkg.FP64, 43)):
        c = pkg.GpuContext(0, precision=prec)
        d = pkg.GpuBatchData(c, csr=csr)
        out[name] = c.fit_models(d, MODELS, 64, 200, 200, seed=seed)
        d.close()
        c.close()
    return csr, out


def _valid(csr):
    return np.diff(csr["offsets"]) > 0


def test_fp32_best_model_agreement(runs):
    csr, r = runs
    v = _valid(csr)
    agree = (r["fp32"]["best_model"] == r["fp64"]["best_model"])[v].mean()
    floor = (r["fp64_seed43"]["best_model"] == r["fp64"]["best_model"])[v].mean()
    assert agree >= 0.965, agree
    # FP32's departure from FP64 is far inside the method's own seed spread
    assert (1.0 - agree) * 8.0 <= (1.0 - floor), (agree, floor)


def test_fp32_cost_gap_is_symmetric(runs):
    csr, r = runs
    v = _valid(csr)
    b32, b64 = r["fp32"]["best_cost"][v], r["fp64"]["best_cost"][v]
    gap = (b32 - b64) / np.maximum(np.abs(b64), 1.0)
    worse, better = (gap > 1e-4).mean(), (gap < -1e-4).mean()
    assert worse <= 0.14 and abs(worse - better) <= 0.04, (worse, better)
    assert abs(np.median(gap)) <= 1e-6, np.median(gap)
    assert np.quantile(gap, 0.95) <= 3e-3 and np.quantile(gap, 0.05) >= -3e-3


def test_fp32_reported_cost_is_the_fp64_cost_at_its_parameters(runs, orc):
    csr, r = runs
    off = csr["offsets"]
    a = r["fp32"]
    worst = 0.0
    for i, m in enumerate(MODELS):
        k = orc.n_params(m)
        for j in range(off.size - 1):
            s, e = off[j], off[j + 1]
            if e == s:
                continue
            want = orc.pso_cost(m, a["model_params"][i, j, :k], *(csr[q][s:e] for q in KEYS))
            got = a["model_cost"][i, j]
            if want >= 1e99:
                assert got >= 1e99
                continue
            worst = max(worst, abs(got - want) / max(abs(want), 1.0))
    # the FP32 swarms re-cost their best positions in FP64
    assert worst <= 1e-12, worst
