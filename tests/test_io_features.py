"""Readers, the per-source JSON record and the feature map (no GPU):

  * the feature map has exactly the reference's keys and model codes
    (tests/golden/feature_keys.json, read from src/features.rs) when every
    result is present, and the reference's None fill when some are absent;
  * derived features on hand-computed cases (delta magnitudes, decay
    deceleration, g - r colours and their evolution, per-model delta chi2,
    MultiBazin BIC gain, mean posterior width, Bazin timescales) and the 2D-GP
    cooling rate against numpy's least-squares slope;
  * NPZ / GOPREAUX CSV / ZTF CSV readers on files written here, row by row
    against the reference's rules (src/bin/extract_features.rs:91-316),
    including dropped rows, filter aliases and upper limits;
  * the JSON record: serde's field names, enum names, skipped empty fields,
    non-finite numbers as null.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "feature_keys.json").read_text())


@pytest.fixture(scope="module")
def mods(pkg):
    from lightcurve_fitting_amd import features, io, gp2d, nonparametric, parametric, thermal, gpu
    return features, io, gp2d, nonparametric, parametric, thermal, gpu


def _np_band(nonparametric, band, **kw):
    r = nonparametric.NonparametricBandResult(band=band)
    for k, v in kw.items():
        setattr(r, k, v)
    return r


def _results(mods):
    features, io, gp2d, nonparametric, parametric, thermal, gpu = mods
    M = gpu.GpuModelName
    npr = [_np_band(nonparametric, "g", peak_mag=18.2, t0=10.0, mag_at_30d=18.9, mag_at_60d=19.8, mag_at_90d=20.5,
                    von_neumann_ratio=1.5, rise_amplitude_over_noise=4.0, post_peak_monotonicity=0.9, n_obs=20),
           _np_band(nonparametric, "r", peak_mag=18.0, t0=12.0, mag_at_30d=18.5, mag_at_60d=19.3, mag_at_90d=19.9,
                    von_neumann_ratio=0.5, rise_amplitude_over_noise=7.0, post_peak_monotonicity=0.7, n_obs=25,
                    n_upper_limits=3, rise_time=float("nan"), fwhm=30.0)]
    mb = gpu.MultiBazinResult(2, [0.1] * 10, 1.2, 50.0, [60.0, 50.0, float("nan"), float("nan")],
                              [60.0, 50.0, float("nan"), float("nan")])
    par = [parametric.ParametricBandResult(band="r", model=M.Bazin, pso_params=[0.1, 0.0, 5.0, math.log(2.0),
                                                                                 math.log(20.0), -3.0],
                                           pso_chi2=1.5, svi_mu=[0.1, 0.2, 0.3], svi_log_sigma=[math.log(0.1),
                                                                                                math.log(0.3)],
                                           svi_elbo=-12.0, n_obs=25, mag_chi2=2.0,
                                           per_model_chi2={M.Bazin: 1.5, M.Villar: 1.9, M.Tde: None},
                                           per_model_params={}, multi_bazin=mb),
           parametric.ParametricBandResult(band="g", model=M.Villar, pso_params=[0.0] * 7, pso_chi2=1.1, svi_mu=[],
                                           svi_log_sigma=[], svi_elbo=None, n_obs=20, mag_chi2=None,
                                           per_model_chi2={}, per_model_params={})]
    th = thermal.ThermalResult(4.1, 4.0, -0.01, 0.05, 0.002, 1.3, 12, 3, "r")
    g2 = gp2d.Gp2dResult(1.0, 20.0, 0.3, 0.05, 80, 3, ["g", "r", "i"], [3.67, 3.79, 3.88])
    g2t = gp2d.Gp2dThermalResult([0.0, 10.0, 20.0, 35.0], [4.2, 4.1, 4.05, 3.9], [2.0, 1.0, 1.5, 3.0])
    return npr, par, th, g2, g2t


def test_keys_and_model_codes_match_reference(mods):
    features, _, _, _, _, _, gpu = mods
    npr, par, th, g2, g2t = _results(mods)
    f = features.extract_features_from_results(npr, par, th, g2, g2t, "r")
    assert sorted(f) == GOLD["keys"]
    for name, code in GOLD["model_codes"].items():
        assert features.MODEL_CODE[gpu.GpuModelName[name]] == code
    assert [m.name for m in features.FEATURE_MODELS] == GOLD["model_names"]
    # absent thermal / 2D GP: their keys present as None; no results at all: no np / param keys
    f2 = features.extract_features_from_results(npr, par, None, None, None, "r")
    assert sorted(f2) == GOLD["keys"]
    assert all(f2[k] is None for k in f2 if k.startswith(("thermal_", "gp2d_")))
    f3 = features.extract_features_from_results([], [], None, None, None, "r")
    assert all(k.startswith(("thermal_", "gp2d_")) for k in f3)


def test_derived_features(mods):
    features = mods[0]
    npr, par, th, g2, g2t = _results(mods)
    f = features.extract_features_from_results(npr, par, th, g2, g2t, "r")
    close = lambda a, b: a is not None and abs(a - b) < 1e-12
    assert f["np_rise_time"] is None and f["np_fwhm"] == 30.0          # NaN -> None
    assert f["np_n_det"] == 25.0 and f["np_n_upper_limits"] == 3.0
    assert close(f["np_delta_mag_30d"], 0.5) and close(f["np_delta_mag_60d"], 1.3) and close(f["np_delta_mag_90d"], 1.9)
    assert close(f["np_decay_deceleration"], 0.75)                     # 0.6 / 0.8
    assert close(f["np_peak_color_g_r"], 0.2) and close(f["np_peak_time_offset_g_r"], -2.0)
    assert close(f["np_color_g_r_30d"], 0.4) and close(f["np_color_g_r_60d"], 0.5) and close(f["np_color_g_r_90d"], 0.6)
    assert close(f["np_color_evolution_rate"], 0.1 / 30.0)
    assert close(f["np_von_neumann_mean"], 1.0) and f["np_rise_significance_max"] == 7.0
    assert close(f["np_monotonicity_mean"], 0.8) and f["np_n_bands_fit"] == 2.0
    assert f["param_best_model"] == 0.0 and f["param_pso_chi2"] == 1.5 and f["param_n_obs"] == 25.0
    assert [f[f"param_svi_mu_{i}"] for i in range(4)] == [0.1, 0.2, 0.3, None]
    assert close(f["param_svi_sigma_mean"], 0.2)
    assert close(f["param_bazin_rise_tau"], 2.0) and close(f["param_bazin_decay_tau"], 20.0)
    assert close(f["param_delta_chi2"], 0.4)                           # 1.5 - 1.1
    assert f["param_chi2_Bazin"] == 1.5 and close(f["param_dchi2_Villar"], 0.4)
    assert f["param_chi2_Tde"] is None and f["param_dchi2_Tde"] is None and f["param_chi2_Afterglow"] is None
    assert f["mb_best_k"] == 2.0 and f["mb_bic_improvement_k2"] == 10.0
    assert f["thermal_n_color_obs"] == 12.0 and f["thermal_chi2"] == 1.3
    assert f["gp2d_log_temp_peak"] == 4.1 and f["gp2d_log_temp_latest"] == 3.9
    assert close(f["gp2d_temp_range"], 0.3) and f["gp2d_thermal_chi2_min"] == 1.0
    assert close(f["gp2d_thermal_chi2_mean"], 1.875)
    slope = np.polyfit(g2t.times, g2t.log_temps, 1)[0]
    assert abs(f["gp2d_cooling_rate"] - slope) < 1e-12
    # reference band fallback: the first band when the named one is absent
    f4 = features.extract_features_from_results(npr, par, th, g2, g2t, "z")
    assert f4["np_peak_mag"] == 18.2 and f4["param_best_model"] == 0.0


def test_npz_reader(mods, tmp_path):
    io = mods[1]
    rows = np.array([[1.0, 0, 0, -0.4, 0.01],      # g: mag 24.9, err 0.025
                     [2.0, 0, 1, -0.2, 0.02],      # r
                     [3.0, 0, 2, 0.0, 0.03],       # i
                     [4.0, 0, 1, np.nan, 0.02],    # non-finite flux: dropped
                     [5.0, 0, 1, -0.2, 0.0],       # err <= 0: dropped
                     [6.0, 0, 5, -0.2, 0.02],      # unknown band: dropped
                     [7.0, 0, 1.7, -0.3, 0.02],    # band id truncates to 1 (r)
                     [8.0, 0, np.nan, -0.1, 0.02]], dtype=np.float32)  # NaN id -> 0 (g)
    p = tmp_path / "s.npz"
    np.savez(p, data=rows)
    t, m, e, b = io.read_photometry_npz(p)
    assert t == [1.0, 2.0, 3.0, 7.0, 8.0] and b == ["g", "r", "i", "r", "g"]
    assert np.allclose(m, -2.5 * rows[[0, 1, 2, 6, 7], 3].astype(np.float64) + 23.9)
    assert np.allclose(e, 2.5 * rows[[0, 1, 2, 6, 7], 4].astype(np.float64))
    np.savez(tmp_path / "w.npz", data=rows.astype(np.float64))      # not f32: unreadable, as Array2<f32>
    assert io.read_photometry_npz(tmp_path / "w.npz") is None
    assert io.read_photometry_npz(tmp_path / "missing.npz") is None
    np.savez(tmp_path / "n.npz", data=rows[3:6])                    # no usable row
    assert io.read_photometry_npz(tmp_path / "n.npz") is None


def test_datacube_reader(mods, tmp_path):
    io = mods[1]
    d = tmp_path / "cube" / "sub"
    d.mkdir(parents=True)
    (d / "SN1_datacube_mangled.csv").write_text(
        "Phase,Filter,Mag,Magerr,Nondetection,Other\n"
        "-5.0,ztfg,18.5,0.05,False,x\n"
        "-4.0, r ,18.1,0.04,False\n"        # filter trimmed
        "-3.0,ps1::z,17.9,0.06,,\n"
        "-2.0,g,18.0,0.05, true\n"          # non-detection (trimmed)
        "-1.0,g,18.0,0.05,1\n"
        "0.0,g,inf,0.05,False\n"            # non-finite mag
        "1.0,g,18.0,0.0,False\n"            # err <= 0
        "2.0,H,18.0,0.05,False\n"           # unknown filter
        " 3.0,g,18.0,0.05,False\n"          # phase does not parse (whitespace)
        "4.0,c,17.5,0.1\n"                  # ATLAS cyan, short row
        "5.0\n")                            # missing fields
    (tmp_path / "cube" / "other.csv").write_text("x\n")
    idx = io.build_datacube_index(tmp_path / "cube")
    assert list(idx) == ["SN1"]
    t, m, e, b = io.read_photometry_datacube(idx, "SN1")
    assert t == [-5.0, -4.0, -3.0, 4.0] and b == ["g", "r", "z", "c"]
    assert m == [18.5, 18.1, 17.9, 17.5] and e == [0.05, 0.04, 0.06, 0.1]
    assert io.read_photometry_datacube(idx, "nope") is None


def test_ztf_reader_and_discovery(mods, tmp_path):
    io = mods[1]
    for name in ("ZTF21b", "ZTF20a", ".hidden"):
        (tmp_path / name).mkdir()
    (tmp_path / "afile.txt").write_text("")
    (tmp_path / "ZTF20a" / "photometry.csv").write_text(
        "jd,fid,magpsf,sigmapsf,diffmaglim\n"
        "2459000.5,1,18.2,0.05,20.0\n"      # g detection
        "2459001.5,2,,,20.5\n"              # r upper limit at diffmaglim, err 1
        "2459002.5,3, 18.9 ,0.07,20.1\n"    # magpsf is trimmed before parsing
        "2459003.5,2,18.0,0,20.0\n"         # detection with invalid error: dropped
        "2459004.5,2,nan,0.05,20.0\n"       # non-finite magnitude: dropped
        "0,1,18.0,0.05,20.0\n"              # jd <= 0
        "2459005.5,4,18.0,0.05,20.0\n"      # unknown fid
        "2459006.5,1,,,-1\n")               # limit <= 0: dropped
    t, m, e, b = io.read_photometry_ztf(tmp_path / "ZTF20a")
    assert t == [2459000.5, 2459001.5, 2459002.5] and b == ["g", "r", "i"]
    assert m == [18.2, 20.5, 18.9] and e == [0.05, 1.0, 0.07]
    assert [s.id for s in io.discover_ztf_sources(tmp_path)] == ["ZTF20a", "ZTF21b"]
    assert io.read_photometry_ztf(tmp_path / "ZTF21b") is None


def test_source_lists(mods, tmp_path):
    io = mods[1]
    (tmp_path / "l.csv").write_text("split,obj_id\ntrain,A1\nval,\ntest,A3\n")
    assert [(s.id, s.split) for s in io.read_npz_source_list(tmp_path / "l.csv")] == [("A1", "train"), ("A3", "test")]
    (tmp_path / "c.csv").write_text("Name,Type,Subtype,Z\nSN1,SN,Ia,0.1\nSN2,TDE,,0.2\n")
    got = [(s.id, s.split, s.subtype) for s in io.read_datacube_source_list(tmp_path / "c.csv")]
    assert got == [("SN1", "SN", "Ia"), ("SN2", "TDE", "")]


def test_json_record(mods, tmp_path):
    features, io = mods[0], mods[1]
    npr, par, th, g2, g2t = _results(mods)
    f = features.extract_features_from_results(npr, par, th, g2, g2t, "r")
    p = io.write_source_json(tmp_path, "SRC1", npr, par, th, g2, g2t, f)
    rec = json.loads(p.read_text())
    assert list(rec) == ["source_id", "nonparametric", "parametric", "thermal", "gp2d", "gp2d_thermal", "features"]
    assert rec["source_id"] == "SRC1" and p.name == "SRC1.json"
    assert rec["nonparametric"][1]["band"] == "r" and rec["nonparametric"][1]["rise_time"] is None
    assert rec["parametric"][0]["model"] == "Bazin" and rec["parametric"][0]["uncertainty_method"] == "Svi"
    assert rec["parametric"][0]["per_model_chi2"] == {"Bazin": 1.5, "Villar": 1.9, "Tde": None}
    assert "per_model_params" not in rec["parametric"][0]          # empty map skipped
    assert "multi_bazin" not in rec["parametric"][1]               # absent skipped
    assert rec["parametric"][0]["multi_bazin"]["per_k_bic"] == [60.0, 50.0, None, None]
    assert rec["thermal"]["ref_band"] == "r" and rec["gp2d"]["bands"] == ["g", "r", "i"]
    assert sorted(rec["features"]) == GOLD["keys"]
    rec2 = io.source_record("S", [], [], None, None, None, {})
    assert rec2["thermal"] is None and rec2["gp2d_thermal"] is None and rec2["parametric"] == []
