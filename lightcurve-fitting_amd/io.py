"""Photometry readers and the per-source JSON record of the feature
extractor (src/bin/extract_features.rs).

Readers return ``(times, mags, mag_errs, bands)`` lists, or ``None`` when a
file is missing, unreadable or yields no usable row, with the reference's row
rules:

* ``read_photometry_npz`` -- AppleCider NPZ, array ``data.npy`` of rows
  [dt, ., band_id, logflux, logflux_err]: mag = -2.5 logflux + 23.9,
  err = 2.5 logflux_err, band 0/1/2 -> g/r/i; rows with a non-finite flux or
  error, err <= 0 or another band id are dropped (:91-130);
* ``parse_datacube_csv`` / ``build_datacube_index`` -- GOPREAUX
  ``<id>_datacube_mangled.csv`` (columns Phase, Filter, Mag, Magerr, optional
  Nondetection): non-detections (True / true / 1) and rows without a finite
  phase and magnitude or a positive finite error are dropped, filter names map
  to g r i z y u c o (ztfg, ps1::g, ... included), other filters are dropped
  (:133-238);
* ``read_photometry_ztf`` / ``discover_ztf_sources`` -- a source directory
  holding ``photometry.csv`` (jd, fid 1/2/3 -> g/r/i, magpsf, sigmapsf,
  diffmaglim): a row with a parseable magpsf is a detection (finite mag,
  positive finite sigmapsf, jd > 0), else a finite positive diffmaglim enters
  as an upper limit with error 1.0 (SNR < 3); directories sorted by name
  (:243-340);
* source lists: ``read_npz_source_list`` (obj_id, split) and
  ``read_datacube_source_list`` (Name, Type, Subtype) (:346-392).

``source_record`` / ``write_source_json`` build the output record
``{source_id, nonparametric, parametric, thermal, gp2d, gp2d_thermal,
features}`` (:759-767) with serde's encoding of the result types: struct
fields by name, enum variants as their names, maps keyed by variant name,
``None`` and non-finite floats as null.
"""
from __future__ import annotations

import csv
import dataclasses
import enum
import json
import math
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

Photometry = Optional[tuple]  # (times, mags, mag_errs, bands)

_DATACUBE_BANDS = {"g": "g", "ztfg": "g", "ps1::g": "g", "r": "r", "ztfr": "r", "ps1::r": "r",
                   "i": "i", "ztfi": "i", "ps1::i": "i", "z": "z", "ps1::z": "z", "y": "y", "ps1::y": "y",
                   "u": "u", "c": "c", "o": "o"}
_NPZ_BANDS = {0: "g", 1: "r", 2: "i"}
_ZTF_BANDS = {1: "g", 2: "r", 3: "i"}


@dataclass
class SourceEntry:
    """One source of a source list (:80-85)."""

    id: str
    split: str = ""
    subtype: str = ""


def _float(s) -> Optional[float]:
    """Rust's str::parse::<f64>: None when the field does not parse (inf and
    nan parse; surrounding whitespace does not)."""
    if s is None:
        return None
    try:
        if s != s.strip() or s == "" or "_" in s:
            return None
        return float(s)
    except ValueError:
        return None


def _int(s) -> Optional[int]:
    try:
        if s is None or s != s.strip() or "_" in s:
            return None
        return int(s)
    except ValueError:
        return None


def read_photometry_npz(path) -> Photometry:
    try:
        with np.load(path, allow_pickle=False) as z:
            name = "data" if "data" in z.files else ("data.npy" if "data.npy" in z.files else None)
            if name is None:
                return None
            data = np.asarray(z[name])
    except Exception:  # missing, not a zip archive, or not loadable without pickles
        return None
    if data.ndim != 2 or data.shape[1] < 5 or data.dtype != np.float32:
        return None
    times, mags, errs, bands = [], [], [], []
    for row in data:
        dt, lf, lfe = float(row[0]), float(row[3]), float(row[4])
        b = float(row[2])  # Rust's saturating f32 -> i32 cast: NaN -> 0, truncation toward zero
        band_id = 0 if math.isnan(b) else int(max(min(b, 2147483647.0), -2147483648.0))
        if not math.isfinite(lf) or not math.isfinite(lfe) or lfe <= 0.0:
            continue
        band = _NPZ_BANDS.get(band_id)
        if band is None:
            continue
        times.append(dt)
        mags.append(-2.5 * lf + 23.9)
        errs.append(2.5 * lfe)
        bands.append(band)
    return (times, mags, errs, bands) if times else None


def build_datacube_index(directory) -> dict:
    """``<obj_id>_datacube_mangled.csv`` files under ``directory`` (walked
    once) by object id."""
    index = {}
    suffix = "_datacube_mangled.csv"
    for root, _dirs, files in os.walk(directory):
        for name in files:
            if name.endswith(suffix):
                index[name[: -len(suffix)]] = Path(root) / name
    return index


def _rows(path):
    with open(path, newline="") as fh:
        rd = csv.reader(fh)
        header = next(rd, None)
        if header is None:
            return None, []
        return header, list(rd)


def parse_datacube_csv(path) -> Photometry:
    try:
        header, rows = _rows(path)
    except OSError:
        return None
    if header is None:
        return None
    try:
        ip, iflt, im, ie = (header.index(h) for h in ("Phase", "Filter", "Mag", "Magerr"))
    except ValueError:
        return None
    ind = header.index("Nondetection") if "Nondetection" in header else None
    times, mags, errs, bands = [], [], [], []
    get = lambda r, i: r[i] if i < len(r) else None
    for r in rows:
        if ind is not None:
            v = get(r, ind)
            if v is not None and v.strip() in ("True", "true", "1"):
                continue
        ph, mg, me = _float(get(r, ip)), _float(get(r, im)), _float(get(r, ie))
        if ph is None or not math.isfinite(ph) or mg is None or not math.isfinite(mg):
            continue
        if me is None or not (math.isfinite(me) and me > 0.0):
            continue
        flt = get(r, iflt)
        if flt is None:
            continue
        band = _DATACUBE_BANDS.get(flt.strip())
        if band is None:
            continue
        times.append(ph)
        mags.append(mg)
        errs.append(me)
        bands.append(band)
    return (times, mags, errs, bands) if times else None


def read_photometry_datacube(index: dict, obj_id: str) -> Photometry:
    p = index.get(obj_id)
    return parse_datacube_csv(p) if p is not None else None


def read_photometry_ztf(directory) -> Photometry:
    try:
        header, rows = _rows(Path(directory) / "photometry.csv")
    except OSError:
        return None
    if header is None:
        return None
    try:
        ij, ifd, im, ie = (header.index(h) for h in ("jd", "fid", "magpsf", "sigmapsf"))
    except ValueError:
        return None
    il = header.index("diffmaglim") if "diffmaglim" in header else None
    times, mags, errs, bands = [], [], [], []
    get = lambda r, i: r[i] if i < len(r) else None
    for r in rows:
        jd = _float(get(r, ij))
        if jd is None or not jd > 0.0:
            continue
        fid = _int(get(r, ifd))
        band = _ZTF_BANDS.get(fid) if fid is not None else None
        if band is None:
            continue
        mag = _float((get(r, im) or "").strip())
        if mag is not None:
            if not math.isfinite(mag):
                continue
            err = _float(get(r, ie))
            if err is None or not (err > 0.0 and math.isfinite(err)):
                continue
            times.append(jd)
            mags.append(mag)
            errs.append(err)
            bands.append(band)
        elif il is not None:
            lim = _float(get(r, il))
            if lim is not None and math.isfinite(lim) and lim > 0.0:
                times.append(jd)
                mags.append(lim)
                errs.append(1.0)  # SNR < 3: an upper limit downstream
                bands.append(band)
    return (times, mags, errs, bands) if times else None


def discover_ztf_sources(directory) -> list:
    out = []
    try:
        entries = list(os.scandir(directory))
    except OSError:
        return out
    for e in entries:
        if e.name and not e.name.startswith(".") and e.is_dir(follow_symlinks=True):
            out.append(SourceEntry(e.name))
    out.sort(key=lambda s: s.id)
    return out


def read_npz_source_list(path) -> list:
    header, rows = _rows(path)
    ii = header.index("obj_id") if "obj_id" in header else 0
    isp = header.index("split") if "split" in header else 1
    out = []
    for r in rows:
        oid = r[ii] if ii < len(r) else ""
        if oid:
            out.append(SourceEntry(oid, r[isp] if isp < len(r) else ""))
    return out


def read_datacube_source_list(path) -> list:
    header, rows = _rows(path)
    inm = header.index("Name") if "Name" in header else 0
    ity = header.index("Type") if "Type" in header else None
    ist = header.index("Subtype") if "Subtype" in header else None
    out = []
    for r in rows:
        name = r[inm] if inm < len(r) else ""
        typ = r[ity] if ity is not None and ity < len(r) else ""
        sub = r[ist] if ist is not None and ist < len(r) else ""
        if name:
            out.append(SourceEntry(name, typ, sub))
    return out


# ---- JSON output --------------------------------------------------------------

def to_json_value(v):
    """serde_json's encoding of the result types: dataclasses as objects of
    their fields, enums by variant name, maps keyed by variant name, None
    and non-finite floats as null, numpy scalars and arrays as numbers and
    lists."""
    if v is None or isinstance(v, (bool, str)):
        return v
    if isinstance(v, enum.Enum):
        return v.name
    if isinstance(v, (int, np.integer)):
        return int(v)
    if isinstance(v, (float, np.floating)):
        f = float(v)
        return f if math.isfinite(f) else None
    if dataclasses.is_dataclass(v):
        out = {}
        for f in dataclasses.fields(v):
            x = getattr(v, f.name)
            # ParametricBandResult skips an empty per_model_params map and an
            # absent multi_bazin (src/parametric.rs:3063, 3070)
            if f.name == "per_model_params" and not x:
                continue
            if f.name == "multi_bazin" and x is None:
                continue
            out[f.name] = to_json_value(x)
        return out
    if isinstance(v, dict):
        return {(k.name if isinstance(k, enum.Enum) else str(k)): to_json_value(x) for k, x in v.items()}
    if isinstance(v, (list, tuple, np.ndarray)):
        return [to_json_value(x) for x in v]
    raise TypeError(f"no JSON encoding for {type(v).__name__}")


def source_record(source_id: str, np_bands, param_bands, thermal, gp2d, gp2d_thermal, features) -> dict:
    """The per-source output object (:759-767)."""
    return {"source_id": source_id, "nonparametric": to_json_value(np_bands),
            "parametric": to_json_value(param_bands), "thermal": to_json_value(thermal),
            "gp2d": to_json_value(gp2d), "gp2d_thermal": to_json_value(gp2d_thermal),
            "features": to_json_value(features)}


def write_source_json(output_dir, source_id: str, np_bands, param_bands, thermal, gp2d, gp2d_thermal,
                      features) -> Path:
    """``<output_dir>/<source_id>.json``; an existing file is what the
    reference's --skip-existing treats as done (:428-434)."""
    p = Path(output_dir) / f"{source_id}.json"
    rec = source_record(source_id, np_bands, param_bands, thermal, gp2d, gp2d_thermal, features)
    p.write_text(json.dumps(rec, allow_nan=False, separators=(",", ":")))
    return p
