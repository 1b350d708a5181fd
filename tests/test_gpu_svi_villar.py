"""The Villar SVI workload (3 bands x 60 points per source, sigma = 0.05 peak
+ 2 uJy, 16 reparameterised samples, Adam lr 1e-2) on the device against the
oracle (needs an MI355X).

A third of this workload's points are upper limits.  With the reference's
upper-limit adjoint (phi / max(Phi, 1e-300), Phi from erf_approx,
cuda/svi.cu:200-204) the gradient at the initial guide reaches 1e284, Adam's
second moment overflows and the fit never leaves its start (the oracle shows
the same); the workload therefore runs with LCF_SVI_EXACT_TAIL, the exact
derivative of the log_normal_cdf that the ELBO sums.

  * the FP32 adjoint gradient of one step's ELBO per job against FP64 central
    finite differences (h = 1e-6 relative) of the oracle's ELBO at the same
    noise: within 1e-4 of the job's gradient scale, at the initial guide and
    at the fitted posterior;
  * FP32 posterior means and standard deviations within 2 % of the oracle's
    posterior width over 5 seeds (the reference's GPU stopping rule), and at
    the full 2000 steps without early stop;
  * the CPU fit's semantics (LCF_SVI_CPU_STOPPING | LCF_SVI_INFLATE_SIGMA:
    min 200 / stall 50 and + ln 4, src/parametric.rs:2262-2263, 2441-2446):
    FP64 kernel vs oracle to 1e-7 with either tail, FP32 within 2 % with the
    exact tail.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OBS = ("times", "flux", "obs_var", "is_upper", "upper_flux")
CPU_STOPPING, INFLATE, FIXED, EXACT = 1, 2, 8, 16


@pytest.fixture(scope="module")
def villar(pkg, orc):
    csr = pkg.synthetic.villar_svi_batch(24, seed=77)
    nj = csr["offsets"].size - 1
    prm, _, _ = orc.batch_pso(1, csr, 64, 200, 200, 42)
    pso = np.zeros((nj, 7))
    pso[:, :7] = prm
    pc, pw = np.zeros((nj, 7)), np.zeros((nj, 7))
    for j in range(nj):
        c, w = orc.svi_prior_for_model(1, pso[j])
        pc[j], pw[j] = c, w
    return csr, np.full(nj, 1, dtype=np.int32), pso, pc, pw


def _agreement(orc, out, mu, ls, mids):
    dmu, dsig = [], []
    for j in range(len(mids)):
        sig = np.exp(ls[j, :7])
        dmu.extend(np.abs(out["mu"][j, :7] - mu[j, :7]) / sig)
        dsig.extend(np.abs(np.exp(out["log_sigma"][j, :7]) - sig) / sig)
    return np.quantile(dmu, 0.95), np.quantile(dsig, 0.95)


def _fd_check(orc, csr, mids, mu, ls, pc, pw, g32, S, seed, step):
    off = csr["offsets"]
    gmu32, gls32 = g32
    worst = 0.0
    for j in range(len(mids)):
        eps = np.stack([orc.svi_normals(seed, j, step, s, 7) for s in range(S)])
        obs = [csr[k][off[j]:off[j + 1]] for k in OBS]
        f = lambda m_, l_: orc.svi_elbo_grad(1, m_, l_, pc[j], pw[j], eps, *obs, grad_mode=2)[0]
        fd_mu, fd_ls = np.zeros(7), np.zeros(7)
        for k in range(7):
            h = 1e-6 * max(abs(mu[j, k]), 1.0)
            p, m = mu[j].copy(), mu[j].copy()
            p[k] += h
            m[k] -= h
            fd_mu[k] = (f(p, ls[j]) - f(m, ls[j])) / (2 * h)
            h = 1e-6 * max(abs(ls[j, k]), 1.0)
            p, m = ls[j].copy(), ls[j].copy()
            p[k] += h
            m[k] -= h
            fd_ls[k] = (f(mu[j], p) - f(mu[j], m)) / (2 * h)
        scale = max(np.abs(fd_mu).max(), np.abs(fd_ls).max(), 1.0)
        worst = max(worst, np.abs(gmu32[j, :7] - fd_mu).max() / scale, np.abs(gls32[j, :7] - fd_ls).max() / scale)
    return worst


def test_fp32_adjoint_vs_fp64_central_fd(ctx, orc, pkg, villar):
    csr, mids, pso, pc, pw = villar
    data = pkg.GpuBatchData(ctx, csr=csr)
    S, seed, step = 16, 3, 0
    # at the initial guide (log sigma = -1) ...
    ls0 = np.full_like(pso, -1.0)
    _, gmu, gls = ctx.svi_elbo_grad(data, mids, pso, ls0, pc, pw, S, seed=seed, step=step, flags=EXACT)
    w0 = _fd_check(orc, csr, mids, pso, ls0, pc, pw, (gmu, gls), S, seed, step)
    # ... and at the fitted posterior
    mu, ls, _, _ = orc.batch_svi(csr, mids, pso, pc, pw, 2000, 16, 0.01, seed=42, flags=EXACT)
    _, gmu, gls = ctx.svi_elbo_grad(data, mids, mu, ls, pc, pw, S, seed=seed, step=7, flags=EXACT)
    w1 = _fd_check(orc, csr, mids, mu, ls, pc, pw, (gmu, gls), S, seed, 7)
    assert w0 <= 1e-4 and w1 <= 1e-4, (w0, w1)


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_fp32_posterior_within_2pct_over_seeds(ctx, orc, pkg, villar, seed):
    csr, mids, pso, pc, pw = villar
    data = pkg.GpuBatchData(ctx, csr=csr)
    out = ctx.svi_fit(data, mids, pso, pc, pw, 2000, 16, 0.01, seed=seed, flags=EXACT)
    mu, ls, _, steps = orc.batch_svi(csr, mids, pso, pc, pw, 2000, 16, 0.01, seed=seed, flags=EXACT)
    assert np.array_equal(out["steps"], steps)
    qmu, qsig = _agreement(orc, out, mu, ls, mids)
    assert qmu <= 0.02 and qsig <= 0.02, (qmu, qsig)


def test_fp32_posterior_at_2000_fixed_steps(ctx, orc, pkg, villar):
    csr, mids, pso, pc, pw = villar
    sub = pkg.synthetic.subset(csr, np.arange(12))
    data = pkg.GpuBatchData(ctx, csr=sub)
    out = ctx.svi_fit(data, mids[:12], pso[:12], pc[:12], pw[:12], 2000, 16, 0.01, seed=42, flags=FIXED | EXACT)
    assert np.all(out["steps"] == 2000)
    mu, ls, _, steps = orc.batch_svi(sub, mids[:12], pso[:12], pc[:12], pw[:12], 2000, 16, 0.01, seed=42,
                                     flags=FIXED | EXACT)
    assert np.all(steps == 2000)
    qmu, qsig = _agreement(orc, out, mu, ls, mids[:12])
    assert qmu <= 0.02 and qsig <= 0.02, (qmu, qsig)


@pytest.mark.parametrize("tail", [0, EXACT])
def test_cpu_semantics_flags_match_oracle(pkg, orc, villar, tail):
    csr, mids, pso, pc, pw = villar
    flags = CPU_STOPPING | INFLATE | tail
    mu, ls, elbo, steps = orc.batch_svi(csr, mids, pso, pc, pw, 1000, 4, 0.01, seed=42, flags=flags)
    assert np.all(steps >= 201)
    c64 = pkg.GpuContext(0, precision=pkg.FP64)
    d64 = pkg.GpuBatchData(c64, csr=csr)
    o64 = c64.svi_fit(d64, mids, pso, pc, pw, 1000, 4, 0.01, seed=42, flags=flags)
    assert np.array_equal(o64["steps"], steps)
    assert np.allclose(o64["mu"], mu, rtol=1e-7, atol=1e-8)
    assert np.allclose(o64["log_sigma"], ls, rtol=1e-7, atol=1e-8)
    # the ln 4 inflation is in the returned log sigma (against the plain run)
    plain = c64.svi_fit(d64, mids, pso, pc, pw, 1000, 4, 0.01, seed=42, flags=flags & ~INFLATE)
    assert np.allclose(o64["log_sigma"] - plain["log_sigma"], np.log(4.0), atol=1e-12)
    if not tail:
        # the reference's tail: FP64 Adam moments overflow and freeze the
        # guide at its start, which FP32 (skipping non-finite steps) cannot
        # reproduce; FP64 parity above is the check
        assert np.median(np.abs(mu - pso).max(axis=1)) == 0.0
        return
    assert np.median(np.abs(mu - pso).max(axis=1)) > 1e-3  # the exact tail moves the guide
    c32 = pkg.GpuContext(0, precision=pkg.FP32)
    d32 = pkg.GpuBatchData(c32, csr=csr)
    o32 = c32.svi_fit(d32, mids, pso, pc, pw, 1000, 4, 0.01, seed=42, flags=flags)
    qmu, qsig = _agreement(orc, o32, mu, ls, mids)
    assert qmu <= 0.02 and qsig <= 0.02, (qmu, qsig)
