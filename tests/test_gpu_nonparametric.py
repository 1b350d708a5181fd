"""GPU parity of the nonparametric band pipeline (needs an MI355X) against the
oracle's process_band, through lcf_batch_nonparametric:

  * status (no result / dense / FITC) identical for every band;
  * GP mean, derivative and std on the grid within 1e-8 relative (std 1e-7)
    for dense bands, 1e-6 for FITC bands (ill-conditioned inducing-point
    system: last-bit differences of the exponential are amplified);
  * features within 1e-8 relative (t0, peak_mag, FWHM, crossing times, chi2,
    predictive features ...), except the least-squares slopes of the fine
    grid (rise / decay rates), compared to 1e-9 mag/day absolute because a
    flat GP mean makes them pure rounding noise (|rate| ~ 1e-15);
  * both branches: dense up to 200 detections on a 200-point grid (the
    nonparametric workload's settings) and FITC above the reference's
    100-detection threshold; edge cases (empty / short bands, all upper
    limits, zero time range);
  * the reference API (BandData maps -> NonparametricBandResult + DenseGP map),
    including the pooled "combined" band of sources whose bands are all short.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SLOPES = {"rise_rate", "decay_rate", "near_peak_rise_rate", "near_peak_decay_rate"}


def feat_close(orc, g, o, rel=1e-8):
    bad = []
    for j, name in enumerate(orc.NP_FEATURES):
        a, b = g[:, j], o[:, j]
        if not np.array_equal(np.isfinite(a), np.isfinite(b)):
            bad.append((name, "finite pattern", np.nonzero(np.isfinite(a) != np.isfinite(b))[0][:5]))
            continue
        f = np.isfinite(b)
        if name in SLOPES:
            ok = np.abs(a[f] - b[f]) <= 1e-9 + rel * np.abs(b[f])
        else:
            ok = np.abs(a[f] - b[f]) <= rel * np.maximum(np.abs(b[f]), 1.0)
        if not ok.all():
            k = np.nonzero(f)[0][~ok][0]
            bad.append((name, k, a[k], b[k]))
    return bad


def grid_close(a, b, rel):
    """Finite patterns equal and |a - b| <= rel (|b| + 1) element-wise (rel
    per row); returns (ok, message naming the worst element)."""
    fin = np.isfinite(b)
    if not np.array_equal(fin, np.isfinite(a)):
        return False, "finite pattern differs at %s" % str(np.argwhere(fin != np.isfinite(a))[:5].tolist())
    d = np.where(fin, np.abs(a - b) / (np.abs(np.where(fin, b, 0.0)) + 1.0), 0.0) / rel[:, None]
    r, c = np.unravel_index(np.argmax(d), d.shape)
    return bool(d[r, c] <= 1.0), f"worst [{r}, {c}]: {a[r, c]!r} vs {b[r, c]!r}, {d[r, c]:.3g} x tolerance"


@pytest.fixture(scope="module")
def ctx64(pkg):
    c = pkg.GpuContext(0)
    yield c
    c.close()


@pytest.mark.parametrize("ng,dense_max,nmin,nmax,seed", [(200, 200, 10, 200, 1), (50, 100, 10, 200, 2),
                                                         (50, 100, 120, 400, 3)])
def test_matches_oracle(pkg, orc, ctx64, ng, dense_max, nmin, nmax, seed):
    c = pkg.synthetic.bazin_mag_batch(40, seed=seed, n_min=nmin, n_max=nmax)
    g = ctx64.nonparametric_arrays(c, c["t_lo"], c["t_hi"], n_grid=ng, dense_max=dense_max)
    o = orc.batch_nonparametric(c, c["t_lo"], c["t_hi"], n_grid=ng, dense_max=dense_max)
    assert np.array_equal(g["status"], o["status"])
    if nmin > dense_max:
        assert np.sum(o["status"] == 2) > 10
    # FITC bands: B = K_mm + K_mn L^-1 K_nm with a 1e-6 jitter is ill-conditioned and
    # amplifies last-bit differences of the device exponential; dense bands 1e-8
    fitc = o["status"] == 2
    for k, rel in (("pred", 1e-8), ("dpred", 1e-8), ("std", 1e-7), ("state", 1e-8)):
        ok, msg = grid_close(g[k], o[k], np.where(fitc, 1e-6, rel))
        assert ok, f"{k}: {msg}"
    dense = ~fitc
    assert feat_close(orc, g["features"][dense], o["features"][dense]) == []
    assert feat_close(orc, g["features"][fitc], o["features"][fitc], 1e-6) == []


def test_edge_cases(pkg, orc, ctx64):
    rng = np.random.default_rng(4)
    bands = []
    bands.append((np.zeros(0), np.zeros(0), np.zeros(0)))                      # empty
    bands.append((np.arange(4.0), np.full(4, 18.0), np.full(4, 0.05)))        # 4 points
    t = np.sort(rng.uniform(0, 100, 30))
    bands.append((t, 18 + 0.1 * rng.normal(size=30), np.full(30, 0.5)))      # all upper limits (SNR < 3)
    bands.append((np.full(12, 5.0), 18 + 0.1 * rng.normal(size=12), np.full(12, 0.05)))  # zero range
    bands.append((t, 18 + np.sin(t / 9.0), np.full(30, 0.05)))                # ordinary
    off = np.zeros(len(bands) + 1, dtype=np.int64)
    off[1:] = np.cumsum([b[0].size for b in bands])
    c = {"offsets": off, "times": np.concatenate([b[0] for b in bands]),
         "mags": np.concatenate([b[1] for b in bands]), "mag_errs": np.concatenate([b[2] for b in bands])}
    lo = np.array([0.0, 0.0, 0.0, 5.0, 0.0])
    hi = np.array([100.0, 3.0, 100.0, 5.0, 100.0])
    g = ctx64.nonparametric_arrays(c, lo, hi)
    o = orc.batch_nonparametric(c, lo, hi)
    assert list(g["status"]) == [0, 0, 0, 0, 1] and np.array_equal(g["status"], o["status"])
    assert g["features"][2, 44] == 30 and g["features"][1, 10] == 4
    assert feat_close(orc, g["features"], o["features"]) == []
    with pytest.raises(pkg.GpuError):
        ctx64.nonparametric_arrays(c, lo, hi, dense_max=300)
    with pytest.raises(pkg.GpuError):
        ctx64.nonparametric_arrays(c, lo, hi, n_grid=1)


def test_reference_api(pkg, orc, ctx64):
    np_ = pkg.nonparametric
    c = pkg.synthetic.bazin_mag_batch(3, seed=8, n_min=20, n_max=120)
    off = c["offsets"]
    sources = []
    for s in range(3):
        sources.append({nm: np_.BandData(c["times"][off[3 * s + k]:off[3 * s + k + 1]],
                                         c["mags"][off[3 * s + k]:off[3 * s + k + 1]],
                                         c["mag_errs"][off[3 * s + k]:off[3 * s + k + 1]])
                        for k, nm in enumerate("gri")})
    # a source of three 3-point bands: only the pooled band can be fitted
    tt = np.array([1.0, 20.0, 40.0])
    sources.append({nm: np_.BandData(tt + k, 18.0 + 0.3 * k + np.array([0.5, 0.0, 0.4]), np.full(3, 0.05))
                    for k, nm in enumerate("gri")})
    out = np_.fit_nonparametric_batch_gpu(ctx64, sources)
    o = orc.batch_nonparametric(c, c["t_lo"], c["t_hi"])
    for s in range(3):
        res, gps = out[s]
        assert [r.band for r in res] == [nm for k, nm in enumerate("gri") if o["status"][3 * s + k] > 0]
        for r in res:
            b = 3 * s + "gri".index(r.band)
            assert r.n_obs == o["features"][b, 10] and r.n_upper_limits == o["features"][b, 44]
            assert r.t0 == pytest.approx(o["features"][b, 6], rel=1e-12)
            assert r.peak_mag == pytest.approx(o["features"][b, 7], rel=1e-10)
            assert r.band in gps and gps[r.band].kernel_amp() == o["features"][b, 45]
    res, gps = out[3]
    assert len(res) == 1 and res[0].band == "combined" and gps == {}
    pooled_t = np.concatenate([tt + k for k in range(3)])
    vals = [18.0 + 0.3 * k + np.array([0.5, 0.0, 0.4]) for k in range(3)]
    pooled_m = np.concatenate([v - v.sum() / v.size for v in vals])
    od = np.argsort(pooled_t, kind="stable")
    grand = np.concatenate(vals).sum() / 9
    cc = {"offsets": np.array([0, 9]), "times": pooled_t[od], "mags": pooled_m[od], "mag_errs": np.full(9, 0.05)}
    oc = orc.batch_nonparametric(cc, [1.0], [42.0])
    assert oc["status"][0] == 1
    assert res[0].peak_mag == pytest.approx(oc["features"][0, 7] + grand, rel=1e-10)
    assert res[0].t0 == pytest.approx(oc["features"][0, 6], rel=1e-12)
