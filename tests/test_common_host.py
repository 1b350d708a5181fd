"""Host-side input preparation (CPU only): the reference's photometry helpers
(src/common.rs:9-150) against its own known-answer tests
(tests/test_common.rs:13-116), and the band normalisation feeding the
fits (src/parametric.rs:3581-3632; the extractor's twin
src/bin/extract_features.rs:548-576) against properties that fix it."""
import math

import numpy as np
import pytest


def test_mag2flux_known_answers(pkg):
    c = pkg.common
    flux, _ = c.mag2flux(23.9, 0.1, 23.9)
    assert abs(flux - 1.0) < 1e-6                      # tests/test_common.rs:13-18
    flux, _ = c.mag2flux(21.4, 0.1, 23.9)
    assert abs(flux - 10.0) < 0.01                     # :20-28
    flux, err = c.mag2flux(23.9, 0.1, 23.9)
    assert abs(err - 0.1 / 1.0857362047581294 * flux) < 1e-8  # :30-39
    # and the exact decade steps of the 10^(-0.4 (m - zp)) form
    assert c.mag2flux(18.9, 0.0)[0] == pytest.approx(100.0, rel=1e-14)


def test_build_mag_bands(pkg):
    c = pkg.common
    times, mags, errs, bands = pkg.synthetic.reference_test_source(10, 42)
    r = c.build_mag_bands(times, mags, errs, bands)
    assert set(r) == {"g", "r", "i"} and len(r) == 3     # :41-50
    assert min(min(b.times) for b in r.values()) == pytest.approx(0.0, abs=1e-10)  # :52-65
    assert all(len(b.times) == len(b.values) == len(b.errors) == 10 for b in r.values())  # :67-76
    assert c.build_mag_bands([], [], [], []) == {}     # :78-82
    # non-finite magnitudes or errors are dropped, times stay relative to the earliest input
    r = c.build_mag_bands([5.0, 6.0, 7.0], [20.0, math.nan, 21.0], [0.1, 0.1, math.inf], ["g", "g", "r"])
    assert r["g"].times == [0.0] and "r" not in r


def test_build_flux_bands(pkg):
    c = pkg.common
    times, mags, errs, bands = pkg.synthetic.reference_test_source(10, 42)
    r = c.build_flux_bands(times, mags, errs, bands)
    assert set(r) == {"g", "r", "i"}                   # :100-108
    for b in r.values():                                # :86-98
        assert all(v > 0.0 for v in b.values) and all(e > 0.0 for e in b.errors)
    assert c.build_flux_bands([], [], [], []) == {}    # :110-114
    # a non-positive error is dropped
    r = c.build_flux_bands([0.0, 1.0], [20.0, 20.0], [0.0, 0.1], ["g", "g"])
    assert len(r["g"].values) == 1 and r["g"].times == [1.0]
    # values are mag2flux of the inputs
    f, e = c.mag2flux(mags[0], errs[0])
    one = c.build_flux_bands(times[:1], mags[:1], errs[:1], bands[:1])[bands[0]]
    assert one.values == [f] and one.errors == [e]


def test_band_normalisation(pkg):
    """band_fit_data: divided by the largest finite flux (NaN ignored), so the
    peak observation is exactly 1; obs_var = err'^2 + 1e-10; an upper limit
    where err > 0 and flux / err < 3, at 3 err'; magnitudes of the raw flux;
    None for an empty band or a non-positive peak."""
    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 50, 30))
    f = 100.0 * np.exp(-((t - 20) / 8) ** 2) + rng.normal(0, 3, 30)
    e = np.full(30, 3.0)
    f[4] = math.nan
    d, mags = pkg.parametric.band_fit_data(pkg.BandData(list(t), list(f), list(e)))
    k = int(np.nanargmax(f))
    peak = f[k]
    assert d.flux[k] == 1.0 and d.peak_flux_obs == peak
    assert np.array_equal(d.flux_err, e / peak)
    assert np.array_equal(d.obs_var, (e / peak) ** 2 + 1e-10)
    with np.errstate(invalid="ignore"):
        assert np.array_equal(d.is_upper.astype(bool), (e > 0) & (f / e < 3.0))
    assert np.array_equal(d.upper_flux, 3.0 * e / peak)
    fin = f > 0
    assert np.allclose(mags[fin], 23.9 - 2.5 * np.log10(f[fin]), rtol=0, atol=1e-12)
    assert pkg.parametric.band_fit_data(pkg.BandData([], [], [])) is None
    assert pkg.parametric.band_fit_data(pkg.BandData([0.0, 1.0], [-1.0, 0.0], [1.0, 1.0])) is None
    # the extractor's twin over a CSR of bands gives the same normalised arrays
    off = np.array([0, 30])
    g = np.where(np.isnan(f), -1e300, f)  # the CSR twin takes finite fluxes
    n = pkg.synthetic.normalise_bands(off, t, g, e)
    assert np.allclose(n["flux"][fin], d.flux[fin]) and np.array_equal(n["obs_var"], d.obs_var)
    assert np.array_equal(n["upper_flux"], d.upper_flux)


def test_prepare_bands_order(pkg):
    """Bands sorted by size, largest first, stable in the map's order."""
    B = pkg.BandData
    mk = lambda n: B(list(np.arange(n, dtype=float)), [1.0 + i for i in range(n)], [0.1] * n)
    out = pkg.parametric.prepare_bands({"a": mk(5), "b": mk(9), "c": mk(5), "d": mk(0), "e": mk(9)})
    assert [name for name, _, _ in out] == ["b", "e", "a", "c"]
