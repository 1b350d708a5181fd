"""GPU parity of the batched Laplace approximation (FP64) against the oracle
(needs an MI355X): Hessian, log sigma and elbo per (band, model) job, for
both Hessian modes, every model (MetzgerKN in the 4-point mode), ragged bands
with upper limits and an empty band."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBS = ("times", "flux", "obs_var", "is_upper", "upper_flux")


@pytest.fixture(scope="module")
def lap_jobs(pkg, orc):
    csr = pkg.synthetic.ragged_selection_batch(6, seed=17, n_min=8, n_max=120)
    nb = csr["offsets"].size - 1
    band_idx, mids = [], []
    for b in range(nb):
        for m in range(8):
            band_idx.append(b)
            mids.append(m)
    jcsr = pkg.synthetic.subset(csr, np.asarray(band_idx))
    mids = np.asarray(mids, dtype=np.int32)
    maps = np.zeros((mids.size, 7))
    for m in range(8):
        sel = np.nonzero(mids == m)[0]
        sub = pkg.synthetic.subset(jcsr, sel)
        prm, _, _ = orc.batch_pso(m, sub, 24, 20 if m != 7 else 3, 20, 42)
        maps[sel, :prm.shape[1]] = prm
    return jcsr, mids, maps


EPS = 2.220446049250313e-16


def log_sigma_from(H):
    """log sigma from a Hessian through the oracle's Jacobi eigensolver
    (log_sigma_from_hessian, src/parametric.rs:2590-2613)."""
    lam, V = orc_eigen(H)
    sm = (V * V / np.maximum(lam, 1e-6)[None, :]).sum(axis=1)
    return 0.5 * np.log(np.maximum(sm, 1e-30))


def orc_eigen(H):
    import oracle
    return oracle.symmetric_eigen(H)


def f_noise(orc, m, p, t, flux, var, up, uf, se2):
    """First-order rounding bound of one objective evaluation: eps per term of
    the sum, plus each model value's rounding (relative 50 eps: exp/pow/log
    of arguments up to ~50 in magnitude) propagated through d term / d pred
    (|r| / v for detections, (|z| + 1) / sqrt(v) for upper limits)."""
    pred = orc.eval_batch(m, p, t)[0] * (10.0 if m == 7 else 1.0)
    v = var + se2
    y = np.where(up != 0, uf, flux)
    r = np.abs(pred - y)
    sens = np.where(up != 0, (r / np.sqrt(v) + 1.0) / np.sqrt(v), r / v)
    terms = 0.5 * (r * r / v + np.abs(np.log(2 * np.pi * v)))
    return EPS * (4.0 * (terms.sum() + t.size) + 50.0 * (sens * (np.abs(pred) + 1e-300)).sum())


def hessian_tol(H, noise, h, mode):
    """Rounding bound of the difference Hessian: the 4-point stencil divides
    four evaluations' noise by 4 h_i h_j; the gradient stencil divides two
    analytic-gradient evaluations (taken to carry the objective's noise per
    unit of parameter) by 2 h_i.  16x slack, plus 1e-6 relative to the
    diagonal scale."""
    d = np.sqrt(np.outer(np.abs(np.diag(H)), np.abs(np.diag(H))))
    if mode == 0:
        return 1e-6 * d + 16 * noise / np.outer(h, h)
    return 1e-6 * d + 16 * noise / np.minimum.outer(h, h)


@pytest.mark.parametrize("mode", [0, 1])
def test_laplace_matches_oracle(pkg, orc, lap_jobs, mode):
    """Per job: elbo = -f(MAP) to 1e-12; the Hessian within its rounding bound
    (hessian_tol); log sigma equal to the oracle's eigen-decomposition of the
    device's own Hessian (so FD noise in flat directions, which the 1e-6
    eigenvalue clamp amplifies, is not mistaken for a kernel error)."""
    jcsr, mids, maps = lap_jobs
    ctx = pkg.GpuContext(0, precision=pkg.FP64)
    sel = np.arange(mids.size) if mode == 0 else np.nonzero(mids != 7)[0]
    sub = pkg.synthetic.subset(jcsr, sel)
    data = pkg.GpuBatchData(ctx, csr=sub)
    out = ctx.laplace_fit(data, mids[sel], maps[sel], mode=mode)
    off = sub["offsets"]
    for k, j in enumerate(sel):
        m = int(mids[j])
        np_ = orc.n_params(m)
        a, b = off[k], off[k + 1]
        ls, elbo, H = orc.laplace_fit(m, maps[j, :np_], *(sub[x][a:b] for x in OBS), mode=mode)
        Hg = out["hessian"][k, :(np_ - 1) ** 2].reshape(np_ - 1, np_ - 1)
        h = 1e-4 * np.maximum(np.abs(maps[j, :np_ - 1]), 1.0)
        obs = [sub[x][a:b] for x in OBS]
        noise = f_noise(orc, m, maps[j, :np_], *obs, np.exp(2 * maps[j, np_ - 1]))
        dH = np.abs(Hg - H)
        assert np.all(dH <= hessian_tol(H, noise, h, mode)), (j, m)
        assert out["elbo"][k] == pytest.approx(elbo, rel=1e-12, abs=1e-12), (j, m)
        assert np.allclose(out["log_sigma"][k, :np_ - 1], log_sigma_from(Hg), rtol=0, atol=1e-9), (j, m)
        assert out["log_sigma"][k, np_ - 1] == ls[np_ - 1], (j, m)
        # first-order perturbation of (H^-1)_ii: |d log sigma| <= ||dH||_2 / (2 lambda_min)
        lam = np.linalg.eigvalsh(H).min()
        dn = np.linalg.norm(Hg - H)
        if lam > 1e-6 and dn < lam / 4:
            assert np.allclose(out["log_sigma"][k, :np_ - 1], ls[:np_ - 1], rtol=0,
                               atol=dn / lam + 1e-9), (j, m)


def test_laplace_empty_band_and_errors(pkg, ctx):
    e = {"offsets": np.array([0, 0], dtype=np.int64), "times": np.zeros(0), "flux": np.zeros(0),
         "obs_var": np.zeros(0), "is_upper": np.zeros(0, np.int32), "upper_flux": np.zeros(0)}
    data = pkg.GpuBatchData(ctx, csr=e)
    out = ctx.laplace_fit(data, [0], [[0.0, 0.0, 10.0, 1.0, 2.0, -3.0]])
    assert np.all(np.isfinite(out["log_sigma"][0, :6]))
    with pytest.raises(pkg.GpuError):
        ctx.laplace_fit(data, [7], [[-2.0, -1.0, 0.5, 0.0, -3.0]], mode=1)
