"""Pins the oracle's blackbody grid fit fit_bb_temperature (CPU only):
magnitudes of an exact Planck SED (up to an offset) at a grid temperature
give back that temperature with chi2 ~ 0 and the offset; fewer than 2 points
give no fit; the chi2 normalisation is 1 / max(n - 2, 1)."""
import numpy as np
import pytest

HC_K = 1.4387769e8


def planck_mag(lam, temp):
    return -2.5 * (-5.0 * np.log(lam) - np.log(np.expm1(HC_K / (lam * temp)))) / np.log(10.0)


def test_exact_sed_recovers_grid_temperature(orc):
    lam = np.array([3671.0, 4770.0, 6231.0, 7625.0])
    for i in (10, 77, 150):
        lt = 3.0 + i * 2.5 / 199.0
        m = planck_mag(lam, 10 ** lt) + 17.3
        got = orc.fit_bb_temperature(lam, m, np.full(4, 0.05))
        assert got[0] == pytest.approx(lt, abs=1e-12)
        assert got[1] < 1e-12 and got[2] == pytest.approx(17.3, abs=1e-7)


def test_no_fit_and_normalisation(orc):
    assert orc.fit_bb_temperature([4770.0], [18.0], [0.1]) is None
    lam = np.array([4770.0, 6231.0])
    lt = 3.0 + 120 * 2.5 / 199.0
    m = planck_mag(lam, 10 ** lt) + np.array([0.01, -0.01])
    got = orc.fit_bb_temperature(lam, m, np.array([0.1, 0.1]))
    # two points: the offset absorbs the mean, chi2 / max(0, 1)
    assert got[1] >= 0 and np.isfinite(got[1])
