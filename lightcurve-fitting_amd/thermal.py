"""Blackbody colour-temperature fit on the GPU: the reference's
``fit_thermal`` (src/thermal.rs:222-418) over ``lcf_batch_thermal``.

The host part is the reference's band bookkeeping: the reference band is the
bluest of u > g > r with at least 5 points (src/thermal.rs:226-229), every
other band with a known effective wavelength (src/thermal.rs:46-63)
contributes colours; the GP of the reference band comes from the prefit map
(e.g. the trained GPs of ``nonparametric.fit_nonparametric_batch``) or is
fitted on the device.  Everything numeric runs in the kernel.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Mapping, Optional, Sequence

import numpy as np

from .gpu import DenseGpState, GpuContext

# band_wavelength (src/thermal.rs:46-63), Angstrom
BAND_WAVELENGTH = {
    "u": 3671.0, "lsstu": 3671.0, "U": 3671.0, "B": 4361.0,
    "g": 4770.0, "ztfg": 4770.0, "lsstg": 4770.0, "c": 5330.0, "V": 5448.0,
    "r": 6231.0, "ztfr": 6231.0, "lsstr": 6231.0, "R": 6555.0, "o": 6790.0,
    "i": 7625.0, "ztfi": 7625.0, "lssti": 7625.0, "I": 8060.0,
    "z": 8691.0, "lsstz": 8691.0, "y": 9712.0, "lssty": 9712.0,
    "J": 12350.0, "H": 16620.0, "K": 21590.0, "Ks": 21590.0,
}
# reference-band preference (src/thermal.rs:226-229)
REF_ALIASES = (("u", "lsstu"), ("g", "ztfg", "lsstg"), ("r", "ztfr", "lsstr"))
STATE_CAP_MIN = 32


@dataclass
class ThermalResult:
    """src/thermal.rs:148-160."""

    log_temp_peak: Optional[float]
    log_temp_latest: Optional[float]
    cooling_rate: Optional[float]
    log_temp_peak_err: Optional[float]
    cooling_rate_err: Optional[float]
    chi2: Optional[float]
    n_color_obs: int
    n_bands_used: int
    ref_band: str


def find_ref_band(bands: Mapping) -> Optional[str]:
    """find_band_name over REF_ALIASES (src/thermal.rs:30-39)."""
    for aliases in REF_ALIASES:
        for a in aliases:
            bd = bands.get(a)
            if bd is not None and len(bd.times) >= 5:
                return a
    return None


def pack_sources(sources: Sequence[Mapping], prefit: Sequence[Optional[Mapping]] | None = None):
    """Flattens sources (band name -> BandData) into the arrays of
    ``lcf_batch_thermal``.  Returns (pack, meta); meta[s] is None for a source
    without a reference band or time range (the reference returns None), else
    (ref band name, [band name of each colour observation])."""
    prefit = prefit or [None] * len(sources)
    caps = [STATE_CAP_MIN]
    for pf in prefit:
        for gp in (pf or {}).values():
            caps.append(len(gp.x))
    cap = max(caps)
    ref_off, col_off = [0], [0]
    rt, rm, re_, ct, cm, ce, cl, lref, dur, meta, states = [], [], [], [], [], [], [], [], [], [], []
    for s, bands in enumerate(sources):
        st = np.zeros(2 * cap + 4)
        ref = find_ref_band(bands)
        tmin = min((float(np.min(bd.times)) for bd in bands.values() if len(bd.times)), default=np.inf)
        tmax = max((float(np.max(bd.times)) for bd in bands.values() if len(bd.times)), default=-np.inf)
        d = tmax - tmin
        names = []
        if ref is None or not d > 0:
            meta.append(None)
            d = 0.0
        else:
            bd = bands[ref]
            rt.append(np.asarray(bd.times, float))
            rm.append(np.asarray(bd.values, float))
            re_.append(np.asarray(bd.errors, float))
            gp = (prefit[s] or {}).get(ref)
            if gp is not None:
                m = len(gp.x)
                st[:m] = gp.alpha
                st[cap:cap + m] = gp.x
                st[2 * cap:2 * cap + 4] = [gp.amp, gp.inv_2ls2, gp.y_mean, m]
            for name, b in bands.items():
                if name == ref or name not in BAND_WAVELENGTH:
                    continue
                n = len(b.times)
                ct.append(np.asarray(b.times, float))
                cm.append(np.asarray(b.values, float))
                ce.append(np.asarray(b.errors, float))
                cl.append(np.full(n, BAND_WAVELENGTH[name]))
                names += [name] * n
            meta.append((ref, names))
        ref_off.append(ref_off[-1] + (len(rt[-1]) if meta[-1] is not None else 0))
        col_off.append(col_off[-1] + len(names))
        lref.append(BAND_WAVELENGTH.get(ref, 0.0) if ref else 0.0)
        dur.append(d)
        states.append(st)
    cat = lambda a: np.concatenate(a) if a else np.zeros(0)
    pack = {"ref_off": np.asarray(ref_off, np.int64), "ref_t": cat(rt), "ref_m": cat(rm), "ref_e": cat(re_),
            "col_off": np.asarray(col_off, np.int64), "col_t": cat(ct), "col_m": cat(cm), "col_e": cat(ce),
            "col_lambda": cat(cl), "lambda_ref": np.asarray(lref), "duration": np.asarray(dur),
            "gp_state": np.asarray(states).reshape(len(sources), 2 * cap + 4)}
    return pack, meta


def _opt(v: float) -> Optional[float]:
    return float(v) if np.isfinite(v) else None


def results_from_arrays(r: dict, pack: dict, meta) -> list:
    out = []
    co = pack["col_off"]
    for s, mt in enumerate(meta):
        st = int(r["status"][s])
        if mt is None or st == 0:
            out.append(None)
            continue
        ref, names = mt
        used = r["col_used"][co[s]:co[s + 1]]
        nb = len({n for n, u in zip(names, used) if u})
        o = r["out"][s]
        if st == 1:
            out.append(ThermalResult(None, None, None, None, None, None, int(o[7]), nb, ref))
        else:
            out.append(ThermalResult(_opt(o[0]), _opt(o[1]), _opt(o[2]), _opt(o[3]), _opt(o[4]), _opt(o[5]),
                                     int(o[7]), nb, ref))
    return out


def fit_thermal_batch(gpu: GpuContext, sources: Sequence[Mapping],
                      prefit_gps: Sequence[Optional[Mapping[str, DenseGpState]]] | None = None,
                      seed: int = 42) -> list:
    """fit_thermal for every source on the GPU; None where the reference
    returns None."""
    if not sources:
        return []
    pack, meta = pack_sources(sources, prefit_gps)
    r = gpu.thermal_arrays(pack, seed)
    return results_from_arrays(r, pack, meta)


def fit_thermal(gpu: GpuContext, bands: Mapping, prefit_gps: Optional[Mapping[str, DenseGpState]] = None,
                seed: int = 42) -> Optional[ThermalResult]:
    """fit_thermal (src/thermal.rs:222-418) for one source."""
    return fit_thermal_batch(gpu, [bands], [prefit_gps], seed)[0]
