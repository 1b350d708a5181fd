"""Consolidated feature map of one source (src/features.rs).

``extract_features_from_results`` flattens the nonparametric, parametric,
thermal and 2D-GP results of a source into one ``FeatureMap``: a dict of
about 120 named features, ``None`` where a value is missing or not finite
(src/features.rs:477-494).  The keys, the reference band fallback (the named
band, else the first result), the model codes and every derived quantity
follow the reference's extract_*_features helpers; ``extract_features`` and
``extract_features_batch`` run the device fitters first
(src/features.rs:498-528).
"""
from __future__ import annotations

import math
from typing import Mapping, Optional, Sequence

from .gp2d import Gp2dResult, Gp2dThermalResult
from .gpu import GpuContext, GpuModelName, MultiBazinResult
from .nonparametric import NonparametricBandResult
from .parametric import ParametricBandResult
from .thermal import ThermalResult

FeatureMap = dict  # str -> Optional[float]

# model codes of the feature map (src/features.rs:26-37): not GpuModelName ids
MODEL_CODE = {GpuModelName.Bazin: 0.0, GpuModelName.Villar: 1.0, GpuModelName.MetzgerKN: 2.0,
              GpuModelName.Tde: 3.0, GpuModelName.Arnett: 4.0, GpuModelName.Magnetar: 5.0,
              GpuModelName.ShockCooling: 6.0, GpuModelName.Afterglow: 7.0}
# per-model chi2 keys, in the reference's order (src/features.rs:39-48)
FEATURE_MODELS = (GpuModelName.Bazin, GpuModelName.Villar, GpuModelName.MetzgerKN, GpuModelName.Tde,
                  GpuModelName.Arnett, GpuModelName.Magnetar, GpuModelName.ShockCooling, GpuModelName.Afterglow)

# nonparametric fields copied from the reference band, with their keys
# (src/features.rs:94-163)
_NP_FIELDS = (
    "rise_time", "rise_halfmax", "rise_efold", "t0", "peak_mag", "chi2", "baseline_chi2", "fwhm", "rise_rate",
    "decay_rate", "decay_efold", "decay_halfmax", "dm15", "near_peak_rise_rate", "near_peak_decay_rate",
    "gp_dfdt_now", "gp_dfdt_next", "gp_d2fdt2_now", "gp_predicted_mag_1d", "gp_predicted_mag_2d",
    "gp_time_to_peak", "gp_extrap_slope", "gp_sigma_f", "gp_peak_to_peak", "gp_snr_max", "gp_dfdt_max",
    "gp_dfdt_min", "gp_frac_of_peak", "gp_post_var_mean", "gp_post_var_max", "gp_skewness", "gp_kurtosis",
    "gp_n_inflections", "decay_power_law_index", "decay_power_law_chi2", "von_neumann_ratio", "pre_peak_rms",
    "rise_amplitude_over_noise", "post_peak_monotonicity", "n_local_maxima", "gp_fit_amp",
    "gp_fit_lengthscale", "mag_at_30d", "mag_at_60d", "mag_at_90d")


def _opt(v) -> Optional[float]:
    """``opt`` (src/features.rs:67-69): finite values only."""
    if v is None:
        return None
    v = float(v)
    return v if math.isfinite(v) else None


def _fin(*vs) -> bool:
    return all(v is not None and math.isfinite(float(v)) for v in vs)


def _fold_min(vs) -> float:
    """fold(inf, f64::min): NaN operands are dropped."""
    m = math.inf
    for v in vs:
        if not math.isnan(v) and v < m:
            m = v
    return m


def _fold_max(vs) -> float:
    m = -math.inf
    for v in vs:
        if not math.isnan(v) and v > m:
            m = v
    return m


def _find(results, band):
    for r in results:
        if r.band == band:
            return r
    return None


def extract_np_features(np_results: Sequence[NonparametricBandResult], ref_band: str) -> FeatureMap:
    """src/features.rs:83-261."""
    f: FeatureMap = {}
    ref = _find(np_results, ref_band) or (np_results[0] if np_results else None)
    if ref is None:
        return f
    for name in _NP_FIELDS:
        f[f"np_{name}"] = _opt(getattr(ref, name))
    f["np_n_det"] = float(ref.n_obs)
    f["np_n_upper_limits"] = float(ref.n_upper_limits)
    peak = ref.peak_mag
    for key, val in (("30d", ref.mag_at_30d), ("60d", ref.mag_at_60d), ("90d", ref.mag_at_90d)):
        f[f"np_delta_mag_{key}"] = val - peak if _fin(peak, val) else None
    m30, m60, m90 = ref.mag_at_30d, ref.mag_at_60d, ref.mag_at_90d
    dec = None
    if _fin(m30, m60, m90):
        early = m60 - m30
        dec = (m90 - m60) / early if abs(early) > 1e-6 else None
    f["np_decay_deceleration"] = dec
    # cross-band (g, r)
    g, r = _find(np_results, "g"), _find(np_results, "r")
    both = g is not None and r is not None
    f["np_peak_color_g_r"] = g.peak_mag - r.peak_mag if both and _fin(g.peak_mag, r.peak_mag) else None
    f["np_peak_time_offset_g_r"] = g.t0 - r.t0 if both and _fin(g.t0, r.t0) else None
    for off in ("30d", "60d", "90d"):
        col = None
        if both:
            a, b = getattr(g, f"mag_at_{off}"), getattr(r, f"mag_at_{off}")
            col = a - b if _fin(a, b) else None
        f[f"np_color_g_r_{off}"] = col
    c30, c60 = f["np_color_g_r_30d"], f["np_color_g_r_60d"]
    f["np_color_evolution_rate"] = (c60 - c30) / 30.0 if c30 is not None and c60 is not None else None
    # aggregates over all bands
    vn = [x.von_neumann_ratio for x in np_results if _fin(x.von_neumann_ratio)]
    f["np_von_neumann_mean"] = sum(vn) / len(vn) if vn else None
    rise = [x.rise_amplitude_over_noise for x in np_results if _fin(x.rise_amplitude_over_noise)]
    f["np_rise_significance_max"] = max(rise) if rise else None
    mono = [x.post_peak_monotonicity for x in np_results if _fin(x.post_peak_monotonicity)]
    f["np_monotonicity_mean"] = sum(mono) / len(mono) if mono else None
    f["np_n_bands_fit"] = float(len(np_results))
    return f


def extract_mb_features(mb: Optional[MultiBazinResult], f: FeatureMap) -> None:
    """src/features.rs:341-366."""
    if mb is None:
        for k in ("mb_best_k", "mb_cost", "mb_bic", "mb_bic_improvement_k2"):
            f[k] = None
        return
    f["mb_best_k"] = float(mb.best_k)
    f["mb_cost"] = _opt(mb.cost)
    f["mb_bic"] = _opt(mb.bic)
    kb = list(mb.per_k_bic)
    f["mb_bic_improvement_k2"] = kb[0] - kb[1] if len(kb) >= 2 and _fin(kb[0], kb[1]) else None


def extract_param_features(param_results: Sequence[ParametricBandResult], ref_band: str) -> FeatureMap:
    """src/features.rs:267-339."""
    f: FeatureMap = {}
    ref = _find(param_results, ref_band) or (param_results[0] if param_results else None)
    if ref is None:
        return f
    model = GpuModelName(int(ref.model))
    f["param_best_model"] = MODEL_CODE[model]
    f["param_pso_chi2"] = _opt(ref.pso_chi2)
    f["param_svi_elbo"] = _opt(ref.svi_elbo)
    f["param_mag_chi2"] = _opt(ref.mag_chi2)
    f["param_n_obs"] = float(ref.n_obs)
    for i in range(4):
        f[f"param_svi_mu_{i}"] = _opt(ref.svi_mu[i]) if i < len(ref.svi_mu) else None
    ls = list(ref.svi_log_sigma)
    f["param_svi_sigma_mean"] = sum(math.exp(v) for v in ls) / len(ls) if ls else None
    if model == GpuModelName.Bazin and len(ref.pso_params) >= 5:
        f["param_bazin_rise_tau"] = math.exp(ref.pso_params[3])
        f["param_bazin_decay_tau"] = math.exp(ref.pso_params[4])
    else:
        f["param_bazin_rise_tau"] = f["param_bazin_decay_tau"] = None
    chi2s = sorted(float(r.pso_chi2) for r in param_results if _fin(r.pso_chi2))
    f["param_delta_chi2"] = chi2s[1] - chi2s[0] if len(chi2s) >= 2 else None
    best = ref.pso_chi2
    pm = {GpuModelName(int(k)): v for k, v in ref.per_model_chi2.items()}
    for m in FEATURE_MODELS:
        mc = pm.get(m)
        f[f"param_chi2_{m.name}"] = _opt(mc)
        f[f"param_dchi2_{m.name}"] = mc - best if _fin(mc, best) else None
    f["param_n_bands_fit"] = float(len(param_results))
    extract_mb_features(ref.multi_bazin, f)
    return f


_THERMAL_KEYS = ("thermal_log_temp_peak", "thermal_cooling_rate", "thermal_log_temp_peak_err",
                 "thermal_cooling_rate_err", "thermal_chi2", "thermal_n_color_obs", "thermal_n_bands_used")


def extract_thermal_features(thermal: Optional[ThermalResult]) -> FeatureMap:
    """src/features.rs:370-392."""
    if thermal is None:
        return {k: None for k in _THERMAL_KEYS}
    t = thermal
    vals = (_opt(t.log_temp_peak), _opt(t.cooling_rate), _opt(t.log_temp_peak_err), _opt(t.cooling_rate_err),
            _opt(t.chi2), float(t.n_color_obs), float(t.n_bands_used))
    return dict(zip(_THERMAL_KEYS, vals))


_GP2D_KEYS = ("gp2d_train_rms", "gp2d_n_train", "gp2d_n_bands", "gp2d_amp", "gp2d_ls_time", "gp2d_ls_wave")
_GP2D_THERMAL_KEYS = ("gp2d_log_temp_peak", "gp2d_log_temp_latest", "gp2d_temp_range", "gp2d_cooling_rate",
                      "gp2d_thermal_chi2_mean", "gp2d_thermal_chi2_min")


def extract_gp2d_features(gp2d: Optional[Gp2dResult], thermal: Optional[Gp2dThermalResult]) -> FeatureMap:
    """src/features.rs:397-468.  Values are taken as they are (the reference
    does not filter these through ``opt``)."""
    f: FeatureMap = {}
    if gp2d is None:
        f.update({k: None for k in _GP2D_KEYS})
    else:
        f.update(zip(_GP2D_KEYS, (float(gp2d.train_rms), float(gp2d.n_train), float(gp2d.n_bands),
                                  float(gp2d.amp), float(gp2d.ls_time), float(gp2d.ls_wave))))
    if thermal is None or not thermal.log_temps:
        f.update({k: None for k in _GP2D_THERMAL_KEYS})
        return f
    lt, ts, c2 = [float(x) for x in thermal.log_temps], [float(x) for x in thermal.times], \
        [float(x) for x in thermal.chi2s]
    n = float(len(lt))
    # peak temperature at the smallest chi2 (first of equal values; NaN compares equal)
    best = 0
    for i in range(1, len(c2)):
        if c2[i] < c2[best]:
            best = i
    f["gp2d_log_temp_peak"] = lt[best] if best < len(lt) else lt[0]
    f["gp2d_log_temp_latest"] = lt[-1]
    f["gp2d_temp_range"] = _fold_max(lt) - _fold_min(lt)
    mt, ml = sum(ts) / n, sum(lt) / n
    cov = sum((ts[i] - mt) * (lt[i] - ml) for i in range(len(ts)))
    var = sum((ts[i] - mt) ** 2 for i in range(len(ts)))
    f["gp2d_cooling_rate"] = cov / var if var > 1e-10 else 0.0
    f["gp2d_thermal_chi2_mean"] = sum(c2) / n
    f["gp2d_thermal_chi2_min"] = _fold_min(c2)
    return f


def extract_features_from_results(np_results: Sequence[NonparametricBandResult],
                                  param_results: Sequence[ParametricBandResult], thermal: Optional[ThermalResult],
                                  gp2d_result: Optional[Gp2dResult], gp2d_thermal: Optional[Gp2dThermalResult],
                                  ref_band: str = "r") -> FeatureMap:
    """``extract_features_from_results`` (src/features.rs:477-494)."""
    f: FeatureMap = {}
    f.update(extract_np_features(np_results, ref_band))
    f.update(extract_param_features(param_results, ref_band))
    f.update(extract_thermal_features(thermal))
    f.update(extract_gp2d_features(gp2d_result, gp2d_thermal))
    return f


def extract_features_batch(gpu: GpuContext, sources_mag: Sequence[Mapping], sources_flux: Sequence[Mapping],
                           ref_band: str = "r") -> list:
    """``extract_features_batch`` (src/features.rs:520-528): every fitter on
    the device for all sources at once (nonparametric + thermal on its GPs,
    parametric with all models and SVI, 2D GP with thermal), then the feature
    map of each source."""
    from . import gp2d, nonparametric, parametric, thermal

    np_out = nonparametric.fit_nonparametric_batch(gpu, sources_mag)
    th = thermal.fit_thermal_batch(gpu, sources_mag, [gps for _, gps in np_out])
    par = parametric.fit_parametric_batch(gpu, sources_flux, True, parametric.UncertaintyMethod.Svi)
    g2 = gp2d.fit_gp_2d_with_thermal_batch(gpu, sources_mag, 50)
    out = []
    for k in range(len(sources_mag)):
        r2, t2 = g2[k] if g2[k] is not None else (None, None)
        out.append(extract_features_from_results(np_out[k][0], par[k], th[k], r2, t2, ref_band))
    return out


def extract_features(gpu: GpuContext, mag_bands: Mapping, flux_bands: Mapping, ref_band: str = "r") -> FeatureMap:
    """``extract_features`` (src/features.rs:498-515) for one source."""
    return extract_features_batch(gpu, [mag_bands], [flux_bands], ref_band)[0]
