"""2D (time x wavelength) GP with blackbody thermal extraction on the GPU:
the reference's ``fit_gp_2d_batch_gpu`` (src/gp2d.rs:516-724) over
``lcf_batch_gp2d_fit`` and ``lcf_batch_bb_temperature``.

The host part is the reference's bookkeeping: bands with a known effective
wavelength and at least 3 points, sources with at least 5 points in 2 bands,
the union of band wavelengths, the global time grid and hyper-parameter
candidates; the GP fits, predictions and the blackbody grid fits run on the
device.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Mapping, Optional, Sequence

import numpy as np

from .gpu import Gp2dInput, GpuContext
from .thermal import BAND_WAVELENGTH


@dataclass
class Gp2dResult:
    """src/gp2d.rs:228-244."""

    amp: float
    ls_time: float
    ls_wave: float
    train_rms: float
    n_train: int
    n_bands: int
    bands: list = field(default_factory=list)
    band_log_wavelengths: list = field(default_factory=list)


@dataclass
class Gp2dThermalResult:
    """src/gp2d.rs:247-255."""

    times: list
    log_temps: list
    chi2s: list


def fit_gp_2d_batch_gpu(gpu: GpuContext, sources: Sequence[Mapping], n_time_steps: int):
    """fit_gp_2d_batch_gpu (src/gp2d.rs:516-724): per source Some((Gp2dResult,
    Gp2dThermalResult)) or None."""
    if not sources:
        return []
    preps = []
    for idx, bands in enumerate(sources):
        t, w, m, nv, names, lws, wavs = [], [], [], [], [], [], []
        for name, bd in bands.items():
            wav = BAND_WAVELENGTH.get(name)
            if wav is None or len(bd.times) < 3:
                continue
            lw = float(np.log10(wav))
            if name not in names:
                names.append(name)
                lws.append(lw)
                wavs.append(wav)
            for i in range(len(bd.times)):
                t.append(bd.times[i])
                w.append(lw)
                m.append(bd.values[i])
                nv.append(bd.errors[i] * bd.errors[i])
        if len(t) < 5 or len(names) < 2:
            continue
        preps.append((idx, Gp2dInput(t, w, m, nv), names, lws, wavs, min(t), max(t)))
    results: list = [None] * len(sources)
    if not preps:
        return results
    all_w: list = []
    for p in preps:
        for lw in p[3]:
            if not any(abs(x - lw) < 1e-6 for x in all_w):
                all_w.append(lw)
    all_w.sort()
    nw = len(all_w)
    t0 = min(p[5] for p in preps)
    t1 = max(p[6] for p in preps)
    dur = max(t1 - t0, 1.0)
    eval_t = [t0 + dur * i / max(n_time_steps - 1, 1) for i in range(n_time_steps)]
    qt = np.repeat(eval_t, nw)
    qw = np.tile(all_w, n_time_steps)
    amps = [0.3, 1.0, 3.0, 10.0]
    wr = max(all_w[-1] - all_w[0], 0.01)
    lst = [dur / 4.0, dur / 8.0, dur / 16.0, dur / 32.0]
    lsw = [wr * 2.0, wr, wr / 2.0, wr / 4.0]
    gout = gpu.batch_gp_2d([p[1] for p in preps], qt, qw, amps, lst, lsw, 40)
    # one SED per (source, time step): the source's bands at that time
    off, lam, mag, err, where = [0], [], [], [], []
    for gi, p in enumerate(preps):
        g = gout[gi]
        if not g.success:
            continue
        pred, std = np.asarray(g.pred_grid), np.asarray(g.std_grid)
        for ti in range(n_time_steps):
            for bi, blw in enumerate(p[3]):
                wi = next((k for k, x in enumerate(all_w) if abs(x - blw) < 1e-6), None)
                if wi is None:
                    continue
                v, s = pred[ti * nw + wi], std[ti * nw + wi]
                if np.isfinite(v) and np.isfinite(s):
                    lam.append(p[4][bi])
                    mag.append(v)
                    err.append(s)
            off.append(len(lam))
            where.append((gi, ti))
    bb = gpu.bb_temperature(off, lam, mag, err) if len(off) > 1 else None
    per = {gi: ([], [], []) for gi in range(len(preps))}
    for k, (gi, ti) in enumerate(where):
        if bb["ok"][k]:
            per[gi][0].append(eval_t[ti])
            per[gi][1].append(float(bb["log_t"][k]))
            per[gi][2].append(float(bb["chi2"][k]))
    for gi, p in enumerate(preps):
        g = gout[gi]
        if not g.success:
            continue
        res = Gp2dResult(g.amp, g.ls_time, g.ls_wave, g.train_rms, g.n_train, len(p[2]), list(p[2]), list(p[3]))
        results[p[0]] = (res, Gp2dThermalResult(*per[gi]))
    return results


GP2D_MAX_TRAIN = 200  # fit_gp_2d's training cap (src/gp2d.rs:371)


def _concat_bands(bands: Mapping):
    """Training arrays of fit_gp_2d (src/gp2d.rs:340-368): bands with a known
    wavelength and >= 3 points, concatenated in the mapping's order; None
    below 5 points or 2 bands."""
    t, w, m, nv, names, lws = [], [], [], [], [], []
    for name, bd in bands.items():
        wav = BAND_WAVELENGTH.get(name)
        if wav is None or len(bd.times) < 3:
            continue
        lw = float(np.log10(wav))
        names.append(name)
        lws.append(lw)
        t.extend(float(x) for x in bd.times)
        w.extend([lw] * len(bd.times))
        m.extend(float(x) for x in bd.values)
        nv.extend(float(e) * float(e) for e in bd.errors)
    if len(t) < 5 or len(names) < 2:
        return None
    return np.asarray(t), np.asarray(w), np.asarray(m), np.asarray(nv), names, lws


def _subsample_range(t: np.ndarray) -> tuple[float, float]:
    n = t.size
    if n <= GP2D_MAX_TRAIN:
        sub = t
    else:
        step = n / GP2D_MAX_TRAIN
        sub = t[[int(i * step) for i in range(GP2D_MAX_TRAIN)]]
    return float(np.min(sub)), float(np.max(sub))


def fit_gp_2d_with_thermal_batch(gpu: GpuContext, sources: Sequence[Mapping], n_time_steps: int,
                                 thermal: bool = True) -> list:
    """fit_gp_2d (src/gp2d.rs:336-439) and fit_gp_2d_with_thermal (:479-500)
    for a list of sources (magnitude bands per source), on the device: per
    source Some((Gp2dResult, Gp2dThermalResult or None)) or None.  The
    thermal step evaluates the GP at n_time_steps times spanning the training
    subsample and every band wavelength, then fits a blackbody per time
    (extract_thermal_from_gp2d, :445-476)."""
    preps = [(i, p) for i, p in ((i, _concat_bands(b)) for i, b in enumerate(sources)) if p is not None]
    results: list = [None] * len(sources)
    if not preps:
        return results
    off = np.zeros(len(preps) + 1, dtype=np.int64)
    np.cumsum([p[0].size for _, p in preps], out=off[1:])
    csr = {"offsets": off, "times": np.concatenate([p[0] for _, p in preps]),
           "waves": np.concatenate([p[1] for _, p in preps]), "mags": np.concatenate([p[2] for _, p in preps]),
           "noise_var": np.concatenate([p[3] for _, p in preps])}
    nb_max = max(len(p[4]) for _, p in preps)
    nq = max(n_time_steps, 1) * nb_max if thermal else 1
    qt, qw = np.zeros((len(preps), nq)), np.zeros((len(preps), nq))
    eval_times = []
    for k, (_, p) in enumerate(preps):
        t0, t1 = _subsample_range(p[0])
        et = [t0 + (t1 - t0) * i / max(n_time_steps - 1, 1) for i in range(n_time_steps)]
        eval_times.append(et)
        if thermal and n_time_steps > 0:
            nb = len(p[4])
            qt[k, :n_time_steps * nb] = np.repeat(et, nb)
            qw[k, :n_time_steps * nb] = np.tile(p[5], n_time_steps)
    g = gpu.fit_gp_2d_arrays(csr, qt, qw)
    bb_off, lam, mag, err, where = [0], [], [], [], []
    if thermal:
        for k, (_, p) in enumerate(preps):
            if g["n_train"][k] == 0:
                continue
            nb = len(p[4])
            for ti in range(n_time_steps):
                sl = slice(ti * nb, (ti + 1) * nb)
                lam.extend(BAND_WAVELENGTH[nm] for nm in p[4])
                mag.extend(g["pred"][k, sl])
                err.extend(g["std"][k, sl])
                bb_off.append(len(lam))
                where.append((k, ti))
    bb = gpu.bb_temperature(bb_off, lam, mag, err) if len(bb_off) > 1 else None
    per = {k: ([], [], []) for k in range(len(preps))}
    for j, (k, ti) in enumerate(where):
        if bb["ok"][j]:
            per[k][0].append(eval_times[k][ti])
            per[k][1].append(float(bb["log_t"][j]))
            per[k][2].append(float(bb["chi2"][j]))
    for k, (i, p) in enumerate(preps):
        if g["n_train"][k] == 0:
            continue
        h = g["hyp"][k]
        res = Gp2dResult(float(h[0]), float(h[1]), float(h[2]), float(g["train_rms"][k]), int(g["n_train"][k]),
                         len(p[4]), list(p[4]), list(p[5]))
        results[i] = (res, Gp2dThermalResult(*per[k]) if thermal else None)
    return results


def fit_gp_2d_batch(gpu: GpuContext, sources: Sequence[Mapping]) -> list:
    """fit_gp_2d (src/gp2d.rs:336-439) per source: Gp2dResult or None."""
    return [r[0] if r is not None else None for r in fit_gp_2d_with_thermal_batch(gpu, sources, 0, thermal=False)]
