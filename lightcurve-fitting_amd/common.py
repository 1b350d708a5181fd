"""Photometry helpers of the reference's common module (src/common.rs:9-160):
``BandData`` groups, ``mag2flux`` and the band builders that feed the GPU
pipelines (host-side bookkeeping, no numerics beyond the unit conversion)."""
from __future__ import annotations

import math
from typing import Dict, Sequence

from .nonparametric import BandData

FACTOR = 1.0857362047581294  # 2.5 / ln 10 (src/common.rs:30)


def mag2flux(mag: float, mag_err: float, zp: float = 23.9) -> tuple[float, float]:
    """src/common.rs:32-36."""
    flux = 10.0 ** (-0.4 * (mag - zp))
    return flux, mag_err / FACTOR * flux


def _group(times, vals, errs, bands, keep) -> Dict[str, BandData]:
    if len(times) == 0:
        return {}
    jd_min = min(times)
    out: Dict[str, BandData] = {}
    for i in range(len(times)):
        v, e = keep(vals[i], errs[i])
        if v is None:
            continue
        bd = out.setdefault(bands[i], BandData([], [], []))
        bd.times.append(times[i] - jd_min)
        bd.values.append(v)
        bd.errors.append(e)
    return out


def build_mag_bands(times: Sequence[float], mags: Sequence[float], mag_errs: Sequence[float],
                    bands: Sequence[str]) -> Dict[str, BandData]:
    """src/common.rs:41-73: group by band, times relative to the earliest,
    non-finite magnitudes or errors dropped."""
    return _group(times, mags, mag_errs, bands,
                  lambda m, e: (m, e) if math.isfinite(m) and math.isfinite(e) else (None, None))


def build_flux_bands(times: Sequence[float], mags: Sequence[float], mag_errs: Sequence[float],
                     bands: Sequence[str]) -> Dict[str, BandData]:
    """src/common.rs:79-111: as build_mag_bands, converted to flux (ZP 23.9),
    non-finite or non-positive fluxes and errors dropped."""
    def keep(m, e):
        f, fe = mag2flux(m, e, 23.9)
        ok = math.isfinite(f) and math.isfinite(fe) and f > 0.0 and fe > 0.0
        return (f, fe) if ok else (None, None)
    return _group(times, mags, mag_errs, bands, keep)


def build_raw_flux_bands(times: Sequence[float], fluxes: Sequence[float], flux_errs: Sequence[float],
                         bands: Sequence[str]) -> Dict[str, BandData]:
    """src/common.rs:117-150: linear fluxes grouped as they are (non-finite
    or non-positive fluxes and errors dropped)."""
    return _group(times, fluxes, flux_errs, bands,
                  lambda f, e: (f, e) if math.isfinite(f) and math.isfinite(e) and f > 0.0 and e > 0.0
                  else (None, None))
