"""Nonparametric GP light-curve features on the GPU: the reference's
``fit_nonparametric`` / ``fit_nonparametric_batch_gpu`` interface
(src/nonparametric.rs:74-160, 780-1225) over ``lcf_batch_nonparametric``.

Every band of every source runs ``process_band`` on the device in one pass
(grid, GP fit, features); this module only flattens the per-source band maps
into CSR arrays and turns the feature rows back into
``NonparametricBandResult`` records.  Sources whose bands give no result
(all below 5 points) and hold at least 5 observations in total are fitted
again as one pooled, mean-subtracted "combined" band (fit_combined_band,
src/nonparametric.rs:1158-1225).
"""
from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Mapping, Optional, Sequence

import numpy as np

from .gpu import NP_NFEAT, DenseGpState, GpuContext

# src/nonparametric.rs:1110-1111
AMP_CANDIDATES = (0.3, 1.0, 3.0, 10.0)
LS_FACTORS = (3.0, 6.0, 12.0, 24.0, 48.0)


@dataclass
class BandData:
    """Per-band time / value / error triplet (src/common.rs:9-15)."""

    times: Sequence[float]
    values: Sequence[float]
    errors: Sequence[float]


@dataclass
class NonparametricBandResult:
    """Result of one band (src/nonparametric.rs:74-160); None for NaN / inf."""

    band: str
    rise_time: Optional[float] = None
    rise_halfmax: Optional[float] = None
    rise_efold: Optional[float] = None
    decay_efold: Optional[float] = None
    dm15: Optional[float] = None
    decay_halfmax: Optional[float] = None
    t0: Optional[float] = None
    peak_mag: Optional[float] = None
    chi2: Optional[float] = None
    baseline_chi2: Optional[float] = None
    n_obs: int = 0
    fwhm: Optional[float] = None
    rise_rate: Optional[float] = None
    decay_rate: Optional[float] = None
    near_peak_rise_rate: Optional[float] = None
    near_peak_decay_rate: Optional[float] = None
    gp_dfdt_now: Optional[float] = None
    gp_dfdt_next: Optional[float] = None
    gp_d2fdt2_now: Optional[float] = None
    gp_predicted_mag_1d: Optional[float] = None
    gp_predicted_mag_2d: Optional[float] = None
    gp_time_to_peak: Optional[float] = None
    gp_extrap_slope: Optional[float] = None
    gp_sigma_f: Optional[float] = None
    gp_peak_to_peak: Optional[float] = None
    gp_snr_max: Optional[float] = None
    gp_dfdt_max: Optional[float] = None
    gp_dfdt_min: Optional[float] = None
    gp_frac_of_peak: Optional[float] = None
    gp_post_var_mean: Optional[float] = None
    gp_post_var_max: Optional[float] = None
    gp_skewness: Optional[float] = None
    gp_kurtosis: Optional[float] = None
    gp_n_inflections: Optional[float] = None
    decay_power_law_index: Optional[float] = None
    decay_power_law_chi2: Optional[float] = None
    mag_at_30d: Optional[float] = None
    mag_at_60d: Optional[float] = None
    mag_at_90d: Optional[float] = None
    von_neumann_ratio: Optional[float] = None
    pre_peak_rms: Optional[float] = None
    rise_amplitude_over_noise: Optional[float] = None
    post_peak_monotonicity: Optional[float] = None
    n_local_maxima: Optional[float] = None
    n_upper_limits: int = 0
    gp_fit_amp: Optional[float] = None
    gp_fit_lengthscale: Optional[float] = None
    stetson_k: Optional[float] = None
    excess_variance: Optional[float] = None
    cusum_range: Optional[float] = None


FEATURE_NAMES = tuple(f.name for f in fields(NonparametricBandResult))[1:]
assert len(FEATURE_NAMES) == NP_NFEAT
_INT_FIELDS = {"n_obs", "n_upper_limits"}
# fields in the magnitude system of the pooled curve (src/nonparametric.rs:1213-1222)
_COMBINED_SHIFT = ("peak_mag", "gp_predicted_mag_1d", "gp_predicted_mag_2d", "mag_at_30d", "mag_at_60d",
                   "mag_at_90d")


def _result(name: str, row: np.ndarray) -> NonparametricBandResult:
    kw = {}
    for k, v in zip(FEATURE_NAMES, row):
        if k in _INT_FIELDS:
            kw[k] = int(v) if np.isfinite(v) else 0
        else:
            kw[k] = float(v) if np.isfinite(v) else None
    return NonparametricBandResult(band=name, **kw)


def _state_gp(st: np.ndarray) -> Optional[DenseGpState]:
    cap = (st.size - 4) // 2
    m = int(st[2 * cap + 3])
    if m <= 0:
        return None
    return DenseGpState(x=st[cap:cap + m].copy(), alpha=st[:m].copy(), amp=float(st[2 * cap]),
                        inv_2ls2=float(st[2 * cap + 1]), y_mean=float(st[2 * cap + 2]))


def _combined_band(bands: Mapping[str, BandData]):
    """Pooled band of fit_combined_band (src/nonparametric.rs:1164-1203):
    per-band means subtracted, finite points only, sorted by time; returns
    (times, values, errors, grand mean) or None below 5 points."""
    ts, vs, es = [], [], []
    gsum, gn = 0.0, 0
    for bd in bands.values():
        t = np.asarray(bd.times, dtype=np.float64)
        if t.size == 0:
            continue
        v = np.asarray(bd.values, dtype=np.float64)
        e = np.asarray(bd.errors, dtype=np.float64)
        mean = v.sum() / v.size
        ok = np.isfinite(v) & np.isfinite(e)
        ts.append(t[ok])
        vs.append(v[ok] - mean)
        es.append(e[ok])
        gsum += float(np.sum(v[ok]))
        gn += int(ok.sum())
    if gn < 5:
        return None
    t, v, e = np.concatenate(ts), np.concatenate(vs), np.concatenate(es)
    o = np.argsort(t, kind="stable")
    return t[o], v[o], e[o], gsum / gn


def fit_nonparametric_batch(gpu: GpuContext, sources: Sequence[Mapping[str, BandData]], *,
                            max_dense_subsample: int = 25, n_grid: int = 50, dense_max: int = 100,
                            n_inducing: int = 30, amp_candidates=AMP_CANDIDATES, ls_factors=LS_FACTORS):
    """fit_nonparametric_with_opts (src/nonparametric.rs:1085-1149) for every
    source, batched on the GPU.  Returns, per source, the list of band
    results (input band order) and the dense GPs kept for thermal reuse.
    ``n_grid`` and ``dense_max`` are the reference's 50 and 100 by default."""
    band_src, names, t, m, e, lo, hi = [], [], [], [], [], [], []
    for s, bands in enumerate(sources):
        tmin, tmax = np.inf, -np.inf
        for bd in bands.values():
            tt = np.asarray(bd.times, dtype=np.float64)
            if tt.size:
                tmin, tmax = min(tmin, tt.min()), max(tmax, tt.max())
        for name, bd in bands.items():
            band_src.append(s)
            names.append(name)
            t.append(np.asarray(bd.times, dtype=np.float64))
            m.append(np.asarray(bd.values, dtype=np.float64))
            e.append(np.asarray(bd.errors, dtype=np.float64))
            lo.append(tmin)
            hi.append(tmax)
    out = [([], {}) for _ in sources]
    kw = dict(n_grid=n_grid, amp_candidates=amp_candidates, ls_factors=ls_factors, dense_max=dense_max,
              n_inducing=n_inducing)

    def run(ts, ms, es, los, his, max_sub):
        off = np.zeros(len(ts) + 1, dtype=np.int64)
        off[1:] = np.cumsum([x.size for x in ts])
        cat = lambda a: np.concatenate(a) if a else np.zeros(0)
        csr = {"offsets": off, "times": cat(ts), "mags": cat(ms), "mag_errs": cat(es)}
        return gpu.nonparametric_arrays(csr, los, his, max_subsample=max_sub, **kw)

    have = np.zeros(len(sources), dtype=bool)
    if t:
        r = run(t, m, e, lo, hi, max_dense_subsample)
        for b, s in enumerate(band_src):
            if r["status"][b] > 0:
                out[s][0].append(_result(names[b], r["features"][b]))
                gp = _state_gp(r["state"][b])
                if gp is not None:
                    out[s][1][names[b]] = gp
                have[s] = True
    # pooled fallback for sources without any band result
    ct, cm, ce, clo, chi, csrc, cmean = [], [], [], [], [], [], []
    for s, bands in enumerate(sources):
        if have[s] or not bands:
            continue
        tt = [np.asarray(bd.times, dtype=np.float64) for bd in bands.values()]
        if sum(x.size for x in tt) < 5:
            continue
        comb = _combined_band(bands)
        if comb is None:
            continue
        tmin = min(x.min() for x in tt if x.size)
        tmax = max(x.max() for x in tt if x.size)
        if not tmax - tmin > 0:
            continue
        ct.append(comb[0])
        cm.append(comb[1])
        ce.append(comb[2])
        clo.append(tmin)
        chi.append(tmax)
        csrc.append(s)
        cmean.append(comb[3])
    if ct:
        r = run(ct, cm, ce, clo, chi, 25)
        for b, s in enumerate(csrc):
            if r["status"][b] > 0:
                res = _result("combined", r["features"][b])
                for k in _COMBINED_SHIFT:
                    v = getattr(res, k)
                    if v is not None:
                        setattr(res, k, v + cmean[b])
                out[s][0].append(res)
    return out


def fit_nonparametric(gpu: GpuContext, bands: Mapping[str, BandData], max_dense_subsample: int = 25, **kw):
    """fit_nonparametric / fit_nonparametric_with_opts for one source."""
    return fit_nonparametric_batch(gpu, [bands], max_dense_subsample=max_dense_subsample, **kw)[0]


def fit_nonparametric_batch_gpu(gpu: GpuContext, sources: Sequence[Mapping[str, BandData]],
                                max_dense_subsample: int = 25):
    """fit_nonparametric_batch_gpu[_with_opts] (src/nonparametric.rs:790-1081):
    same signature and result shape; the bands run the CPU path's
    process_band semantics (per-source grid and length scales, train-RMS /
    FITC selection, fine-grid features) on the device."""
    return fit_nonparametric_batch(gpu, sources, max_dense_subsample=max_dense_subsample)
