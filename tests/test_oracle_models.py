"""Pins the oracle's model layer (CPU only).

Checks are chosen so a plausible restatement mistake (swapped parameter,
dropped term, wrong sign) fails one of them:
  * closed-form values at special points (t = t0, phase = gamma, equal
    afterglow slopes, asymptotes);
  * analytic gradients against FP64 central finite differences;
  * log_normal_cdf / erf_approx against math.erfc to the 1.5e-7 bound of
    Abramowitz & Stegun 7.1.26;
  * Philox4x32-10 against the published Random123 known-answer vectors;
  * bound and prior tables against the values of src/gpu_types.rs:64-103 and
    src/parametric.rs:1541-1621, 2072-2115.
"""
import math

import numpy as np
import pytest

LN2 = math.log(2.0)

# Parameter vectors of the reference's CPU/GPU cross-check
# (tests/test_gpu.rs:29-37), evaluated there on t = -10..39.
REF_CASES = [
    (0, [0.5, 0.1, 10.0, 1.0, 3.0, -3.0]),
    (1, [0.5, 0.01, 2.0, 10.0, 1.0, 3.0, -3.0]),
    (2, [0.5, 0.1, 5.0, 1.0, 3.0, 1.67, -3.0]),
    (3, [0.5, 5.0, 2.3, 0.0, -3.0]),
    (4, [0.5, 5.0, 3.0, 2.3, -3.0]),
    (5, [0.5, 5.0, 0.5, 1.0, -3.0]),
    (6, [0.5, 5.0, 2.0, 0.8, 2.2, -3.0]),
]


def f(orc, m, p, t):
    return orc.eval_batch(m, np.asarray(p, float), np.atleast_1d(np.asarray(t, float)))


def test_bazin_special_points(orc):
    la, b, t0, ltr, ltf = 0.5, 0.1, 10.0, 1.0, 3.0
    p = [la, b, t0, ltr, ltf, -3.0]
    a = math.exp(la)
    # logistic(0) = 1/2, decay(0) = 1
    assert f(orc, 0, p, t0)[0] == pytest.approx(a / 2 + b, rel=1e-14)
    # one e-fold of the decay and rise: f(t0 + tf) = A e^-1 s(tf/tr) + B
    tf, tr = math.exp(ltf), math.exp(ltr)
    assert f(orc, 0, p, t0 + tf)[0] == pytest.approx(a * math.exp(-1) / (1 + math.exp(-tf / tr)) + b, rel=1e-13)
    # long after the peak only the baseline remains
    assert f(orc, 0, p, t0 + 5000.0)[0] == pytest.approx(b, abs=1e-12)


def test_villar_at_transition(orc):
    la, beta, lg, t0, ltr, ltf = 0.5, 0.01, 2.0, 10.0, 1.0, 3.0
    p = [la, beta, lg, t0, ltr, ltf, -3.0]
    g = math.exp(lg)
    # at phase = gamma: w = 1/2 and both pieces equal (1 - beta*gamma)
    want = math.exp(la) / (1 + math.exp(-g / math.exp(ltr))) * (1 - beta * g)
    assert f(orc, 1, p, t0 + g)[0] == pytest.approx(want, rel=1e-13)
    # well before gamma the plateau is linear: A s(ph/tr) (1 - beta ph)
    ph = 1.0
    w = 1 / (1 + math.exp(-10 * (ph - g)))
    want = math.exp(la) / (1 + math.exp(-ph / math.exp(ltr))) * ((1 - w) * (1 - beta * ph)
                                                                   + w * (1 - beta * g) * math.exp((g - ph) / math.exp(ltf)))
    assert f(orc, 1, p, t0 + ph)[0] == pytest.approx(want, rel=1e-13)


def test_tde_arnett_magnetar_shock_at_t0(orc):
    ps = LN2 + 1e-6  # softplus(0) + 1e-6
    la = 0.5
    a = math.exp(la)
    # TDE: s(0) = 1/2
    p = [la, 0.1, 5.0, 1.0, 3.0, 1.67, -3.0]
    assert f(orc, 2, p, 5.0)[0] == pytest.approx(a / 2 * (1 + ps / math.exp(3.0)) ** -1.67 + 0.1, rel=1e-13)
    # Arnett, logit f = 0 -> f = 1/2
    p = [la, 5.0, 2.3, 0.0, -3.0]
    heat = 0.5 * math.exp(-ps / 8.8) + 0.5 * math.exp(-ps / 111.3)
    assert f(orc, 3, p, 5.0)[0] == pytest.approx(a * heat * (1 - math.exp(-(ps / math.exp(2.3)) ** 2)), rel=1e-12)
    # Magnetar
    p = [la, 5.0, 3.0, 2.3, -3.0]
    w = 1 + ps / math.exp(3.0)
    assert f(orc, 4, p, 5.0)[0] == pytest.approx(a / w ** 2 * (1 - math.exp(-(ps / math.exp(2.3)) ** 2)), rel=1e-12)
    # ShockCooling: logistic(5*0) = 1/2
    p = [la, 5.0, 0.5, 1.0, -3.0]
    assert f(orc, 5, p, 5.0)[0] == pytest.approx(a / 2 * ps ** -0.5 * math.exp(-(ps / math.e) ** 2), rel=1e-12)


def test_afterglow_equal_slopes(orc):
    # alpha1 = alpha2 = al: f = A (2 r^{2 al})^{-1/2} = A r^{-al} / sqrt(2)
    la, t0, ltb, al = 0.3, 2.0, 1.5, 1.2
    p = [la, t0, ltb, al, al, -3.0]
    t = 40.0
    ps = math.log1p(math.exp(t - t0)) + 1e-6
    r = ps / math.exp(ltb)
    assert f(orc, 6, p, t)[0] == pytest.approx(math.exp(la) * r ** -al / math.sqrt(2), rel=1e-12)


def test_metzger_properties(orc):
    p = [-2.0, -0.7, 0.5, 0.0, -3.0]
    t = np.linspace(-5, 30, 71)
    y = orc.eval_batch(7, p, t)
    assert np.all(y[t <= 0] == 0.0)
    assert np.all(np.isfinite(y)) and y.max() <= 1.0 + 1e-12 and y.max() > 0.5
    # per-observation solve agrees with the batch grid to interpolation accuracy
    single = np.array([orc.metzger_single(p, tt) for tt in t])
    assert np.allclose(single[t > 0.5], y[t > 0.5], rtol=0.15, atol=0.02)


@pytest.mark.parametrize("m,p", REF_CASES)
def test_gradients_match_finite_differences(orc, m, p):
    p = np.asarray(p, float)
    for t in np.arange(-10.0, 40.0, 3.0):
        g = orc.eval_grad(m, p, t)
        for j in range(p.size):
            h = 1e-5 * max(abs(p[j]), 1.0)
            pp, pm = p.copy(), p.copy()
            pp[j] += h
            pm[j] -= h
            fd = (f(orc, m, pp, t)[0] - f(orc, m, pm, t)[0]) / (2 * h)
            # log(1+e^x) in the model loses ~1e-10 relative at x << 0, so the
            # difference quotient is only good to ~1e-5 there
            assert g[j] == pytest.approx(fd, rel=1e-4, abs=1e-8), (m, t, j)


def test_ref_vectors_finite(orc):
    t = np.arange(-10.0, 40.0)
    for m, p in REF_CASES:
        y = orc.eval_batch(m, p, t)
        assert np.all(np.isfinite(y)), m


def test_log_normal_cdf_against_erfc(orc):
    for x in np.linspace(-29.5, 7.9, 400):
        exact = math.log(0.5 * math.erfc(-x / math.sqrt(2)))
        phi = math.exp(orc.log_normal_cdf(x))
        # A&S 7.1.26: |erf error| <= 1.5e-7 absolute in erfc
        assert abs(phi - math.exp(exact)) <= 1.5e-7 * 0.5 + 1e-15 + 1e-6 * math.exp(exact), x
    assert orc.log_normal_cdf(9.0) == 0.0
    x = -40.0
    assert orc.log_normal_cdf(x) == pytest.approx(-0.5 * x * x - 0.5 * math.log(2 * math.pi) - math.log(-x), rel=1e-15)
    for x in np.linspace(-4, 4, 81):
        assert abs(orc.erf_approx(x) - math.erf(x)) <= 1.5e-7


def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for c, k, want in cases:
        assert list(orc.philox(c, k)) == want


def test_tables(orc):
    lo, hi = orc.bounds(1)
    assert list(lo) == [-1.5, -0.02, 1.0, -60.0, -1.0, 1.0, -6.0]
    assert list(hi) == [1.5, 0.05, 5.0, 40.0, 3.0, 5.5, 0.0]
    lo, hi = orc.bounds(0)
    assert lo[1] == -1.0 and hi[1] == 1.0          # GPU Bazin baseline
    lo, hi = orc.bounds(0, cpu=True)
    assert lo[1] == -0.3 and hi[1] == 0.3          # CPU Bazin baseline
    lo, hi = orc.bounds(7)
    assert list(hi) == [-0.5, -0.5, 2.0, 1.0, 0.0]
    c, w = orc.population_priors(1)
    ln10 = math.log(10)
    assert c[2] == pytest.approx(1.327 * ln10) and w[3] == 15.0 and c[1] == 0.009
    c, w = orc.population_priors(0)
    assert list(c) == [0.0, 0.0, 0.0, 1.5, 3.0, -3.0] and list(w) == [2.0, 0.2, 30.0, 2.5, 2.5, 2.0]
    c, w = orc.population_priors(3)
    assert list(c) == [0.0] * 5 and list(w) == [5.0] * 5
    c, w = orc.prior_params(6)
    assert list(c) == [0.0, 0.0, 1.0, 0.8, 2.0, 0.0] and list(w) == [2.0, 2.0, 2.0, 1.0, 1.0, 2.0]
    assert [orc.n_params(m) for m in range(8)] == [6, 7, 7, 5, 5, 5, 6, 5]


@pytest.mark.parametrize("m", [0, 1, 2, 3, 4])
def test_eval_batch_against_numpy_transcription(orc, pkg, m):
    """The oracle's closed-form models against the package's independent
    numpy transcription of the model formulas (synthetic.model_flux, written
    from the formulas, not from the oracle), on random parameters of the
    synthetic workloads' boxes and times before, across and after the peak."""
    lo, hi = (np.asarray(x) for x in pkg.synthetic._BOXES[m])
    rng = np.random.default_rng(100 + m)
    t = np.linspace(-20.0, 200.0, 221)
    for _ in range(25):
        p = lo + rng.uniform(size=lo.size) * (hi - lo)
        want = pkg.synthetic.model_flux(m, p, t)
        got = orc.eval_batch(m, p, t)
        fin = np.isfinite(want)
        assert np.array_equal(fin, np.isfinite(got))
        assert np.allclose(got[fin], want[fin], rtol=1e-11, atol=1e-14), np.max(np.abs(got - want))
