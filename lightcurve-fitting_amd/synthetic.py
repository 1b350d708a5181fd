"""Seeded synthetic light-curve batches for the benchmark and parity tests.

Each generator returns per-band ``BatchSource`` data in the CSR layout of
``GpuBatchData`` (offsets, times, flux, obs_var, is_upper, upper_flux) plus
bookkeeping (source id of each band, true model).  Band normalisation is the
reference's (src/bin/extract_features.rs:548-576): flux and errors divided by
the band's peak observed flux, ``obs_var = err^2 + 1e-10``, an observation is
an upper limit when ``flux/err < 3`` and then ``upper_flux = 3 err / peak``.
"""
from __future__ import annotations

import math

import numpy as np

# Models of the ragged model-selection workload: Bazin, Villar, Tde, Arnett,
# Magnetar (model ids of GpuModelName).
SELECTION_MODELS = (0, 1, 2, 3, 4)

# Sampling boxes (internal parameterisation, last entry ln sigma_x unused)
# chosen so that every curve peaks inside a [0, 150] d window.
_BOXES = {
    0: ((-0.5, -0.05, 20.0, 0.0, 2.3, -5.0), (0.5, 0.05, 80.0, 2.3, 4.1, -5.0)),
    1: ((-0.5, 0.0, 1.5, 10.0, 0.0, 2.0, -5.0), (0.5, 0.02, 3.5, 60.0, 2.0, 4.0, -5.0)),
    2: ((-0.5, -0.05, 10.0, 0.5, 1.0, 1.0, -5.0), (0.5, 0.05, 60.0, 2.5, 3.0, 2.5, -5.0)),
    3: ((-0.5, 5.0, 2.0, -2.0, -5.0), (0.5, 50.0, 3.5, 2.0, -5.0)),
    4: ((-0.5, 5.0, 2.0, 1.5, -5.0), (0.5, 50.0, 4.0, 3.0, -5.0)),
}


def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def _sp(x):
    return np.logaddexp(0.0, x) + 1e-6


def model_flux(model: int, p: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Closed-form model flux; p is [n_params] or broadcastable rows [.., n_params]."""
    p = np.asarray(p, dtype=np.float64)
    c = [p[..., j] for j in range(p.shape[-1])]
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        if model == 0:
            dt = t - c[2]
            return np.exp(c[0]) * np.exp(-dt / np.exp(c[4])) * _sig(dt / np.exp(c[3])) + c[1]
        if model == 1:
            a, beta, g, ph = np.exp(c[0]), c[1], np.exp(c[2]), t - c[3]
            w = _sig(10.0 * (ph - g))
            return a * _sig(ph / np.exp(c[4])) * ((1 - w) * (1 - beta * ph)
                                                  + w * (1 - beta * g) * np.exp((g - ph) / np.exp(c[5])))
        if model == 2:
            ph = t - c[2]
            return (np.exp(c[0]) * _sig(ph / np.exp(c[3]))
                    * (1.0 + _sp(ph) / np.exp(c[4])) ** (-c[5]) + c[1])
        if model == 3:
            ps = _sp(t - c[1])
            f = _sig(c[3])
            heat = f * np.exp(-ps / 8.8) + (1 - f) * np.exp(-ps / 111.3)
            return np.exp(c[0]) * heat * (1 - np.exp(-(ps / np.exp(c[2])) ** 2))
        if model == 4:
            ps = _sp(t - c[1])
            w = 1 + ps / np.exp(c[2])
            return np.exp(c[0]) / (w * w) * (1 - np.exp(-(ps / np.exp(c[3])) ** 2))
    raise ValueError(f"no generator for model {model}")


def ragged_selection_batch(n_sources: int, seed: int = 0, n_min: int = 10, n_max: int = 500,
                           t_max: float = 150.0) -> dict:
    """Ragged model-selection workload.

    ``n_sources`` sources with U{1,2,3} bands each and log-uniform[n_min, n_max]
    points per band; the true model is drawn uniformly from SELECTION_MODELS
    with parameters uniform in a per-model box; Gaussian noise with peak
    SNR ~ U(5, 50); times uniform on [0, t_max], sorted.
    """
    rng = np.random.Generator(np.random.Philox(seed))
    nb = rng.integers(1, 4, size=n_sources)
    band_src = np.repeat(np.arange(n_sources), nb)
    n_bands = band_src.size
    src_model = rng.choice(np.asarray(SELECTION_MODELS), size=n_sources)
    band_model = src_model[band_src]
    npts = np.floor(np.exp(rng.uniform(np.log(n_min), np.log(n_max + 1), size=n_bands))).astype(np.int64)
    npts = np.clip(npts, n_min, n_max)
    off = np.zeros(n_bands + 1, dtype=np.int64)
    np.cumsum(npts, out=off[1:])
    n = int(off[-1])
    band_of = np.repeat(np.arange(n_bands), npts)
    t = rng.uniform(0.0, t_max, size=n)
    # sort times within each band
    order = np.lexsort((t, band_of))
    t = t[order]
    # true parameters per source (shared shape across its bands)
    params = np.zeros((n_sources, 7))
    for m in SELECTION_MODELS:
        sel = src_model == m
        lo, hi = (np.asarray(x) for x in _BOXES[m])
        params[sel, :lo.size] = rng.uniform(lo, hi, size=(int(sel.sum()), lo.size))
    clean = np.empty(n)
    for m in SELECTION_MODELS:
        bm = band_model[band_of] == m
        npar = len(_BOXES[m][0])
        clean[bm] = model_flux(m, params[band_src[band_of[bm]], :npar], t[bm])
    clean = np.nan_to_num(clean, nan=0.0, posinf=0.0, neginf=0.0)
    peak_true = np.maximum.reduceat(np.abs(clean), off[:-1]) if n else np.zeros(0)
    peak_true = np.where(peak_true > 0, peak_true, 1.0)
    snr = rng.uniform(5.0, 50.0, size=n_bands)
    sigma_b = peak_true / snr
    sig = sigma_b[band_of]
    fobs = clean + rng.normal(size=n) * sig
    csr = normalise_bands(off, t, fobs, sig)
    kept = csr["kept_bands"]
    csr["band_source"] = band_src[kept]
    csr["true_model"] = band_model[kept]
    csr["true_params"] = params
    csr["n_sources"] = n_sources
    return csr


def normalise_bands(off: np.ndarray, t: np.ndarray, flux: np.ndarray, err: np.ndarray) -> dict:
    """Reference band normalisation (src/bin/extract_features.rs:556-566)."""
    nb = off.size - 1
    npts = np.diff(off)
    band_of = np.repeat(np.arange(nb), npts)
    peak = np.full(nb, -np.inf)
    if flux.size:
        np.maximum.at(peak, band_of, flux)
    good = peak > 0
    # bands whose peak is not positive are dropped, as the reference does
    keep = good[band_of]
    new_np = npts[good]
    new_off = np.zeros(new_np.size + 1, dtype=np.int64)
    np.cumsum(new_np, out=new_off[1:])
    pk = peak[band_of][keep]
    f = flux[keep] / pk
    e = err[keep] / pk
    is_up = (err[keep] > 0) & (flux[keep] / err[keep] < 3.0)
    return {
        "offsets": new_off,
        "times": t[keep].copy(),
        "flux": f,
        "obs_var": e * e + 1e-10,
        "is_upper": is_up.astype(np.int32),
        "upper_flux": 3.0 * e,
        "kept_bands": np.nonzero(good)[0],
    }


def subset(csr: dict, bands: np.ndarray) -> dict:
    """CSR restricted to the given band indices (in that order)."""
    off = csr["offsets"]
    bands = np.asarray(bands, dtype=np.int64)
    npts = off[bands + 1] - off[bands]
    new_off = np.zeros(bands.size + 1, dtype=np.int64)
    np.cumsum(npts, out=new_off[1:])
    idx = np.concatenate([np.arange(off[b], off[b + 1]) for b in bands]) if bands.size else np.zeros(0, np.int64)
    out = {"offsets": new_off}
    for k in ("times", "flux", "obs_var", "is_upper", "upper_flux"):
        out[k] = csr[k][idx]
    return out


def villar_svi_batch(n_sources: int, seed: int = 0, n_bands: int = 3, n_per_band: int = 60,
                     t_max: float = 120.0) -> dict:
    """SVI workload: ``n_sources`` Villar light curves with ``n_bands`` bands of
    ``n_per_band`` points each, times uniform on [0, t_max] (sorted).  True
    parameters: A ~ U(100, 1000) uJy scaled per band by U(0.7, 1.3), beta ~
    U(0, 0.02), gamma ~ U(5, 30) d, t0 ~ U(10, 60) d, tau_rise ~ U(1, 10) d,
    tau_fall ~ U(10, 60) d; Gaussian noise sigma = 0.05 peak + 2 uJy with peak
    the band's largest true flux.  Returned normalised per band as above,
    with ``band_source`` and the true internal parameters (amplitude of the
    normalised band)."""
    rng = np.random.Generator(np.random.Philox(seed))
    nb = n_sources * n_bands
    band_src = np.repeat(np.arange(n_sources), n_bands)
    off = np.arange(nb + 1, dtype=np.int64) * n_per_band
    t = np.sort(rng.uniform(0.0, t_max, size=(nb, n_per_band)), axis=1).ravel()
    a = rng.uniform(100.0, 1000.0, size=n_sources)[band_src] * rng.uniform(0.7, 1.3, size=nb)
    beta = rng.uniform(0.0, 0.02, size=n_sources)[band_src]
    gam = rng.uniform(5.0, 30.0, size=n_sources)[band_src]
    t0 = rng.uniform(10.0, 60.0, size=n_sources)[band_src]
    tr = rng.uniform(1.0, 10.0, size=n_sources)[band_src]
    tf = rng.uniform(10.0, 60.0, size=n_sources)[band_src]
    p = np.stack([np.log(a), beta, np.log(gam), t0, np.log(tr), np.log(tf), np.full(nb, -5.0)], axis=1)
    band_of = np.repeat(np.arange(nb), n_per_band)
    clean = model_flux(1, p[band_of], t)
    peak = clean.reshape(nb, n_per_band).max(axis=1)
    sig = (0.05 * peak + 2.0)[band_of]
    fobs = clean + rng.normal(size=t.size) * sig
    csr = normalise_bands(off, t, fobs, sig)
    kept = csr["kept_bands"]
    csr["band_source"] = band_src[kept]
    p[:, 0] -= np.log(peak)
    csr["true_params"] = p[kept]
    csr["n_sources"] = n_sources
    return csr


def bazin_mag_batch(n_sources: int, seed: int = 0, n_bands: int = 3, n_min: int = 10, n_max: int = 200,
                    t_span: float = 150.0, sort_times: bool = True) -> dict:
    """Nonparametric workload in magnitudes: ``n_sources`` Bazin light curves
    with ``n_bands`` bands of U{n_min..n_max} points each, times uniform on
    [0, t_span] d.  Flux F = A e^{-(t - t0)/tau_fall} / (1 + e^{-(t - t0)/tau_rise})
    + B with A ~ U(100, 1000) uJy scaled per band by U(0.7, 1.3), t0 ~ U(20, 80),
    tau_rise ~ U(1, 10), tau_fall ~ U(10, 60), B ~ N(0, 5); Gaussian noise of
    sigma = 0.03 A + 2 uJy; mag = 23.9 - 2.5 log10(max(F, 1)), magerr = 1.0857
    sigma / max(F, 1).  Times are made relative to the source's earliest
    observation (as build_mag_bands does), and ``t_lo`` / ``t_hi`` hold each
    band's source time range (the span of the prediction grid)."""
    rng = np.random.Generator(np.random.Philox(seed))
    nb = n_sources * n_bands
    band_src = np.repeat(np.arange(n_sources), n_bands)
    sizes = rng.integers(n_min, n_max + 1, size=nb)
    off = np.zeros(nb + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    band_of = np.repeat(np.arange(nb), sizes)
    t = rng.uniform(0.0, t_span, size=int(off[-1]))
    if sort_times:
        t = t + band_of * (2.0 * t_span)
        t = np.sort(t) - band_of * (2.0 * t_span)
    a = rng.uniform(100.0, 1000.0, size=n_sources)[band_src] * rng.uniform(0.7, 1.3, size=nb)
    t0 = rng.uniform(20.0, 80.0, size=n_sources)[band_src]
    tr = rng.uniform(1.0, 10.0, size=n_sources)[band_src]
    tf = rng.uniform(10.0, 60.0, size=n_sources)[band_src]
    b = rng.normal(0.0, 5.0, size=n_sources)[band_src]
    dt = t - t0[band_of]
    with np.errstate(over="ignore"):
        flux = a[band_of] * np.exp(-dt / tf[band_of]) / (1.0 + np.exp(-dt / tr[band_of])) + b[band_of]
    sig = 0.03 * a[band_of] + 2.0
    flux = flux + rng.normal(size=t.size) * sig
    fc = np.maximum(flux, 1.0)
    mags = 23.9 - 2.5 * np.log10(fc)
    errs = 1.0857 * sig / fc
    src_of = band_src[band_of]
    tmin = np.full(n_sources, np.inf)
    np.minimum.at(tmin, src_of, t)
    t = t - tmin[src_of]
    tmax = np.full(n_sources, -np.inf)
    np.maximum.at(tmax, src_of, t)
    return {"offsets": off, "times": t, "mags": mags, "mag_errs": errs, "band_source": band_src,
            "t_lo": np.zeros(nb), "t_hi": tmax[band_src]}


def parametric_sources(n_sources: int = 100, seed: int = 0, n_per_band: int = 30, bands=("g", "r", "i")) -> list:
    """The plain parametric workload as the reference's fit_parametric takes
    it: ``n_sources`` Bazin sources (A ~ U(100, 1000) uJy scaled per band by
    U(0.7, 1.3), t0 ~ U(20, 80), tau_rise ~ U(1, 10), tau_fall ~ U(10, 60),
    B ~ N(0, 5); noise sigma = 0.03 A + 2 uJy) with ``n_per_band`` points per
    band, times uniform on [0, 100] d, sorted, given as magnitudes (23.9 -
    2.5 log10 max(F, 1), error 1.0857 sigma / max(F, 1)) and turned into flux
    bands by build_flux_bands (src/common.rs:79-111).  Returns one
    {band name: BandData} map per source, in g, r, i order."""
    from .common import build_flux_bands

    c = bazin_mag_batch(n_sources, seed=seed, n_bands=len(bands), n_min=n_per_band, n_max=n_per_band,
                        t_span=100.0)
    off = c["offsets"]
    out = []
    for s in range(n_sources):
        ts, ms, es, bs = [], [], [], []
        for k, name in enumerate(bands):
            b = s * len(bands) + k
            ts += list(c["times"][off[b]:off[b + 1]])
            ms += list(c["mags"][off[b]:off[b + 1]])
            es += list(c["mag_errs"][off[b]:off[b + 1]])
            bs += [name] * int(off[b + 1] - off[b])
        out.append(build_flux_bands(ts, ms, es, bs))
    return out


GRI_WAVELENGTH = (4770.0, 6231.0, 7625.0)  # effective wavelengths of g, r, i (src/gp2d.rs:84-90), Angstrom


def cooling_bb_batch(n_sources: int, seed: int = 0, n_min: int = 30, n_max: int = 512,
                     t_span: float = 100.0, noise: float = 0.03) -> dict:
    """2D GP workload: ``n_sources`` sources of U{n_min..n_max} points spread
    uniformly over g / r / i, times uniform on [0, t_span] d (sorted per
    source).  Flux = Bazin(t) x B_lambda(T(t)) / B_6231(T(t)) with the
    temperature cooling exponentially from 10^4 K at t = 0 to 6000 K at
    t = t_span, Bazin A ~ U(100, 1000) uJy, t0 ~ U(20, 60), tau_rise ~ U(1, 10),
    tau_fall ~ U(10, 60); multiplicative Gaussian noise of 3 %;
    mag = 23.9 - 2.5 log10(max(F, 1)), noise variance (1.0857 x 0.03)^2.
    Returns the 2D GP CSR (offsets, times, waves = log10 wavelength, mags,
    noise_var) and ``theta0`` = ln (var(mag), span / 4, 0.2 dex, 0.01) per
    source."""
    rng = np.random.Generator(np.random.Philox(seed))
    sizes = rng.integers(n_min, n_max + 1, size=n_sources)
    off = np.zeros(n_sources + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    src = np.repeat(np.arange(n_sources), sizes)
    t = rng.uniform(0.0, t_span, size=int(off[-1]))
    t = np.sort(t + src * (2.0 * t_span)) - src * (2.0 * t_span)
    lam = np.asarray(GRI_WAVELENGTH)[rng.integers(0, 3, size=t.size)]
    a = rng.uniform(100.0, 1000.0, n_sources)[src]
    t0 = rng.uniform(20.0, 60.0, n_sources)[src]
    tr = rng.uniform(1.0, 10.0, n_sources)[src]
    tf = rng.uniform(10.0, 60.0, n_sources)[src]
    temp = 1e4 * (0.6 ** (t / t_span))
    hck = 1.4387769e8  # h c / k in Angstrom K
    with np.errstate(over="ignore"):
        bb = (6231.0 / lam) ** 5 * np.expm1(hck / (6231.0 * temp)) / np.expm1(hck / (lam * temp))
        f = a * np.exp(-(t - t0) / tf) / (1.0 + np.exp(-(t - t0) / tr)) * bb
    f = f * (1.0 + noise * rng.normal(size=t.size))
    mags = 23.9 - 2.5 * np.log10(np.maximum(f, 1.0))
    nv = np.full(t.size, (1.0857 * noise) ** 2)
    w = np.log10(lam)
    th0 = np.zeros((n_sources, 4))
    for s in range(n_sources):
        y = mags[off[s]:off[s + 1]]
        th0[s] = [np.log(max(float(np.var(y)), 1e-4)), np.log(max(float(np.ptp(t[off[s]:off[s + 1]])), 1.0) / 4.0),
                  np.log(0.2), np.log(0.01)]
    return {"offsets": off, "times": t, "waves": w, "log_waves": w, "mags": mags, "noise_var": nv, "theta0": th0}


class _XorShift64:
    """xorshift64 with 53-bit uniforms and Box-Muller normals, the generator
    of the reference's test fixtures (tests/synthetic.rs:17-46)."""

    M = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = max(int(seed), 1) & self.M

    def next_u64(self) -> int:
        x = self.s
        x ^= (x << 13) & self.M
        x ^= x >> 7
        x ^= (x << 17) & self.M
        self.s = x
        return x

    def uniform(self) -> float:
        return (self.next_u64() >> 11) / float(1 << 53)

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-15)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def reference_test_source(n_per_band: int, seed: int):
    """generate_bazin_source (tests/synthetic.rs:53-91): g, r, i bands of
    n_per_band points on [0, 100) d after JD 2460000, Bazin flux A = 1,
    B = 0.01, t0 = 30, tau_rise = 3, tau_fall = 25 as magnitudes (25 where the
    flux is not positive), colour offsets -0.5 / 0 / +0.3, errors U(0.05,
    0.08) and Gaussian noise; returns (times, mags, mag_errs, bands)."""
    rng = _XorShift64(seed)
    jd = 2460000.0
    times, mags, errs, bands = [], [], [], []
    for name, off in (("g", -0.5), ("r", 0.0), ("i", 0.3)):
        for j in range(n_per_band):
            t = jd + j * 100.0 / n_per_band
            dt = (t - jd) - 30.0
            flux = 1.0 * math.exp(-dt / 25.0) / (1.0 + math.exp(-dt / 3.0)) + 0.01
            mag = (-2.5 * math.log10(flux) + 23.9 if flux > 0.0 else 25.0) + off
            sigma = 0.05 + 0.03 * rng.uniform()
            times.append(t)
            mags.append(mag + sigma * rng.normal())
            errs.append(sigma)
            bands.append(name)
    return times, mags, errs, bands
