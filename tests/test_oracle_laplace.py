"""Pins the oracle's Laplace approximation (CPU only).

  * the Jacobi eigensolver against numpy.linalg.eigh on random SPD matrices;
  * neg_log_posterior against the PSO cost (detections only):
    f = n/2 (cost_nll + ln 2 pi) + sum (p - c)^2 / (2 w^2) with the
    prior_params priors;
  * the 4-point difference Hessian against an independent numpy 5-point
    stencil of the same objective, and log sigma = 1/2 ln diag(H^-1) by
    numpy.linalg.inv when H is positive definite;
  * the forward-over-adjoint Hessian against the 4-point one.
"""
import math

import numpy as np
import pytest

from conftest import make_reference_test_source

OBS = ("times", "flux", "obs_var", "is_upper", "upper_flux")


def test_jacobi_against_eigh(orc):
    rng = np.random.default_rng(0)
    for n in range(1, 7):
        a = rng.normal(size=(n, n))
        h = a @ a.T + n * np.eye(n)
        ev, vec = orc.symmetric_eigen(h)
        assert np.allclose(np.sort(ev), np.linalg.eigvalsh(h), rtol=1e-10)
        assert np.allclose(vec @ np.diag(ev) @ vec.T, h, atol=1e-9)


def test_nlp_against_cost(orc):
    src = make_reference_test_source(12.0, 1.2, 35, orc)
    p = np.array([0.1, 0.05, 12.0, 0.1, 2.6, -2.8])
    f = orc.neg_log_posterior(0, p, *(src[k] for k in OBS))
    n = 35
    cost = orc.pso_cost(0, p, *(src[k] for k in OBS))
    c, w = orc.population_priors(0)
    nll = cost - sum(0.5 * ((p[j] - c[j]) / w[j]) ** 2 for j in range(6)) / n ** 2
    pc, pw = orc.prior_params(0)
    want = 0.5 * n * (nll + math.log(2 * math.pi)) + sum(0.5 * (p - pc) ** 2 / pw ** 2)
    assert f == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("m,map_", [
    (0, [0.1, 0.05, 12.0, 0.1, 2.6, -2.8]),
    (1, [0.1, 0.01, 2.0, 10.0, 0.5, 2.7, -2.8]),
    (4, [0.2, 6.0, 2.5, 2.2, -2.8]),
])
def test_hessian_and_log_sigma(orc, m, map_):
    src = make_reference_test_source(12.0, 1.2, 35, orc)
    src["is_upper"][[0, 1]] = 1
    src["upper_flux"][[0, 1]] = 0.2
    args = [src[k] for k in OBS]
    p0 = np.asarray(map_, float)
    ls, elbo, H = orc.laplace_fit(m, p0, *args)
    assert elbo == pytest.approx(-orc.neg_log_posterior(m, p0, *args), rel=1e-14)
    ns = p0.size - 1
    f = lambda p: orc.neg_log_posterior(m, p, *args)
    Hn = np.zeros((ns, ns))
    for i in range(ns):
        for j in range(ns):
            hi, hj = 2e-3 * max(abs(p0[i]), 1.0), 2e-3 * max(abs(p0[j]), 1.0)

            def fij(a, b):
                p = p0.copy()
                p[i] += a * hi
                p[j] += b * hj
                return f(p)
            Hn[i, j] = (fij(1, 1) - fij(1, -1) - fij(-1, 1) + fij(-1, -1)) / (4 * hi * hj)
    scale = np.sqrt(np.outer(np.abs(np.diag(Hn)), np.abs(np.diag(Hn)))) + 1e-9
    assert np.all(np.abs(H - Hn) <= 1e-3 * scale + 1e-6), np.abs(H - Hn) / scale
    ev = np.linalg.eigvalsh(H)
    if ev.min() > 1e-6:
        want = 0.5 * np.log(np.diag(np.linalg.inv(H)))
        assert np.allclose(ls[:ns], want, atol=1e-8)
    pc, pw = orc.prior_params(m)
    assert ls[-1] == pytest.approx(math.log(pw[-1]))
    ls2, _, H2 = orc.laplace_fit(m, p0, *args, mode=orc.LAPLACE_GRAD_FD)
    assert np.all(np.abs(H2 - H) <= 1e-4 * scale + 1e-6)
    assert np.allclose(ls2, ls, atol=1e-4)
