// gp_kernels.hip -- batched dense Gaussian-process fits in FP64 for gfx950,
// 1D (time) and 2D (time x log10 wavelength, anisotropic squared exponential).
//
// One 256-thread workgroup owns one band (1D) or source (2D) at a time
// (persistent, pulled from a device queue in decreasing training-set size).
// For every hyper-parameter candidate it builds K + diag(noise) as a packed
// lower triangle in LDS (m <= 128: 66 KB), factorises it in place with a
// right-looking Cholesky whose column scaling and trailing updates are spread
// over all 256 threads, and solves for alpha on wavefront 0 with the right-hand
// side held in registers (lane l owns rows l and l+64), so the triangular
// solves need no barriers.  Candidates are scored by the exact negative log
// marginal likelihood or by the train RMS; the winner is refitted and
// predicted (mean, time derivative of the mean, standard deviation) at the
// query grid and at every observation.
//
// Arithmetic follows DenseGP (src/sparse_gp.rs:16-176), batch_gp_fit
// (cuda/gp.cu:23-367) and batch_gp2d_fit_predict (cuda/gp2d.cu:21-359): the
// right-looking updates subtract l_ik l_jk in the same k order as the
// reference's Crout loops, so factors agree to rounding.
#include "gp_kernels.hpp"
#include "lcf_device.hpp"

namespace lcf {

constexpr int GP_THREADS = 256;

__device__ __forceinline__ int tri(int i, int j) { return (i * (i + 1) >> 1) + j; }

struct GpShared {
    double* L;      // packed lower triangle, cap*(cap+1)/2
    double* x;      // training times [cap]
    double* w;      // training log wavelengths [cap] (2D)
    double* y;      // training values [cap]
    double* nv;     // training noise variances [cap]
    double* alpha;  // [cap]
    double* misc;   // scalars
};

__device__ __forceinline__ GpShared gp_carve(double* base, int cap) {
    GpShared s;
    s.L = base;
    s.x = s.L + cap * (cap + 1) / 2;
    s.w = s.x + cap;
    s.y = s.w + cap;
    s.nv = s.y + cap;
    s.alpha = s.nv + cap;
    s.misc = s.alpha + cap;
    return s;
}

struct Hyp {
    double amp, it, iw;  // amplitude, 1/(2 l_t^2), 1/(2 l_w^2) (0 in 1D)
    __device__ __forceinline__ double k(double t1, double w1, double t2, double w2) const {
        double dt = t1 - t2, dw = w1 - w2;
        return amp * ::exp(-dt * dt * it - dw * dw * iw);
    }
};

// Build K + diag(max(nv, 1e-10)) and factorise in place.  Returns false
// (uniformly) when a pivot is not positive.
__device__ bool gp_factor(const GpShared& s, int m, const Hyp& h) {
    const int tid = threadIdx.x;
    const int mm = m * m;
    for (int e = tid; e < mm; e += GP_THREADS) {
        int i = e / m, j = e - i * m;
        if (j <= i) {
            double v = h.k(s.x[i], s.w[i], s.x[j], s.w[j]);
            if (i == j) v += fmax(s.nv[i], 1e-10);
            s.L[tri(i, j)] = v;
        }
    }
    __syncthreads();
    for (int j = 0; j < m; j++) {
        double piv = s.L[tri(j, j)];
        if (piv <= 0.0) return false;  // uniform: every thread read the same pivot
        double ljj = ::sqrt(piv);
        __syncthreads();  // all have read the pivot before it is overwritten
        if (tid == 0) s.L[tri(j, j)] = ljj;
        for (int i = j + 1 + tid; i < m; i += GP_THREADS) s.L[tri(i, j)] /= ljj;
        __syncthreads();
        const int r = m - j - 1;
        const int rr = r * r;
        for (int e = tid; e < rr; e += GP_THREADS) {
            int a = e / r, b = e - a * r;
            if (b <= a) {
                int i = j + 1 + a, k = j + 1 + b;
                s.L[tri(i, k)] -= s.L[tri(i, j)] * s.L[tri(k, j)];
            }
        }
        __syncthreads();
    }
    return true;
}

// Wavefront 0: alpha = L^-T L^-1 (y - ybar); returns 0.5 yc.alpha + sum ln L_ii
// (wave-uniform).  Lane l owns rows l and l + 64 in registers.
__device__ double gp_solve_alpha(const GpShared& s, int m, double ybar) {
    const int lane = lane_id();
    double b0 = lane < m ? s.y[lane] - ybar : 0.0;
    double b1 = lane + 64 < m ? s.y[lane + 64] - ybar : 0.0;
    const double yc0 = b0, yc1 = b1;
    for (int j = 0; j < m; j++) {  // forward: L z = yc
        double bj = j < 64 ? bcast(b0, j) : bcast(b1, j - 64);
        double zj = bj / s.L[tri(j, j)];
        if (lane == (j & 63)) { if (j < 64) b0 = zj; else b1 = zj; }
        if (lane > j && lane < m) b0 -= s.L[tri(lane, j)] * zj;
        if (lane + 64 > j && lane + 64 < m) b1 -= s.L[tri(lane + 64, j)] * zj;
    }
    for (int j = m - 1; j >= 0; j--) {  // backward: L^T a = z
        double bj = j < 64 ? bcast(b0, j) : bcast(b1, j - 64);
        double aj = bj / s.L[tri(j, j)];
        if (lane == (j & 63)) { if (j < 64) b0 = aj; else b1 = aj; }
        if (lane < j) b0 -= s.L[tri(j, lane)] * aj;
        if (lane + 64 < j) b1 -= s.L[tri(j, lane + 64)] * aj;
    }
    if (lane < m) s.alpha[lane] = b0;
    if (lane + 64 < m) s.alpha[lane + 64] = b1;
    double part = (lane < m ? yc0 * b0 + ::log(s.L[tri(lane, lane)]) * 2.0 : 0.0) +
                  (lane + 64 < m ? yc1 * b1 + ::log(s.L[tri(lane + 64, lane + 64)]) * 2.0 : 0.0);
    return 0.5 * wave_sum(part);
}

// Wavefront 0: RMS of the predictive mean at the training points.
__device__ double gp_train_rms(const GpShared& s, int m, const Hyp& h, double ybar) {
    const int lane = lane_id();
    double rss = 0.0;
    for (int i = lane; i < m; i += 64) {
        double mu = 0.0;
        for (int k = 0; k < m; k++) mu += h.k(s.x[i], s.w[i], s.x[k], s.w[k]) * s.alpha[k];
        double r = mu + ybar - s.y[i];
        rss += r * r;
    }
    return ::sqrt(wave_sum(rss) / (double)m);
}

__device__ __forceinline__ Hyp candidate(const GpArgs& a, int c, bool& valid) {
    const int nw = a.n_lsw > 0 ? a.n_lsw : 1;
    const int ia = c / (a.n_ls * nw), il = (c / nw) % a.n_ls, iw = c % nw;
    Hyp h;
    h.amp = a.amps[ia];
    const double ls = a.lss[il];
    valid = ls >= 0.1;
    h.it = 0.5 / (ls * ls);
    h.iw = 0.0;
    if (a.w) {
        const double lw = a.lsws[iw];
        valid = valid && lw >= 0.001;
        h.iw = 0.5 / (lw * lw);
    }
    return h;
}

__device__ void gp_band(const GpArgs& a, double* lds, int band) {
    const int tid = threadIdx.x;
    const int cap = a.cap;
    const bool two_d = a.w != nullptr;
    GpShared s = gp_carve(lds, cap);
    const int64_t o0 = a.offsets[band];
    const int n = (int)(a.offsets[band + 1] - o0);
    double* st = a.state + (int64_t)band * a.state_stride;
    const double* q = a.query + (a.query_per_band ? (int64_t)band * a.n_query : 0);
    const double* qw = a.query_w ? a.query_w + (a.query_per_band ? (int64_t)band * a.n_query : 0) : nullptr;
    for (int i = tid; i < a.state_stride; i += GP_THREADS) st[i] = 0.0;
    for (int k = tid; k < a.n_query; k += GP_THREADS) {
        a.pred[(int64_t)band * a.n_query + k] = NAN;
        a.stdv[(int64_t)band * a.n_query + k] = NAN;
        if (a.dpred) a.dpred[(int64_t)band * a.n_query + k] = NAN;
    }
    for (int i = tid; i < n; i += GP_THREADS) a.pred_obs[o0 + i] = NAN;
    if (n < 3) return;
    const int m = n < cap ? n : cap;
    for (int i = tid; i < m; i += GP_THREADS) {
        int idx;
        if (n <= m) idx = i;
        else if (a.subsample_rule == 0) {
            double step = (double)(n - 1) / (double)(m - 1);
            idx = (int)(i * step + 0.5);
        } else {
            idx = (int)(((double)i + 0.5) * n / m);
        }
        idx = idx < n - 1 ? idx : n - 1;
        s.x[i] = a.t[o0 + idx];
        s.w[i] = two_d ? a.w[o0 + idx] : 0.0;
        s.y[i] = a.y[o0 + idx];
        s.nv[i] = a.nv[o0 + idx];
    }
    __syncthreads();
    double ybar = 0.0;  // all threads, same summation order
    for (int i = 0; i < m; i++) ybar += s.y[i];
    ybar /= (double)m;

    double best = 1e99;
    int best_c = 0;
    bool any = false;
    const int nc = a.n_amp * a.n_ls * (a.n_lsw > 0 ? a.n_lsw : 1);
    for (int c = 0; c < nc; c++) {
        bool valid;
        const Hyp h = candidate(a, c, valid);
        double score = 1e99;
        if (valid) {
            if (gp_factor(s, m, h)) {
                if (tid < 64) {
                    double half_fit_logdet = gp_solve_alpha(s, m, ybar);
                    double sc;
                    if (a.criterion == 0) sc = half_fit_logdet + 0.5 * (double)m * 1.83787706640934548356;
                    else sc = gp_train_rms(s, m, h, ybar);
                    if (tid == 0) s.misc[0] = __builtin_isfinite(sc) ? sc : 1e99;
                }
                __syncthreads();
                score = s.misc[0];
            }
            __syncthreads();
        }
        if (c == 0) any = valid;
        if (score < best) { best = score; best_c = c; any = true; }
    }
    if (!any) return;
    bool valid;
    const Hyp h = candidate(a, best_c, valid);
    if (!gp_factor(s, m, h)) return;
    if (tid < 64) gp_solve_alpha(s, m, ybar);
    __syncthreads();
    // state 1D: alpha[cap], x[cap], amp, 1/(2 lt^2), ybar, m
    //       2D: alpha[cap], x[cap], w[cap], amp, 1/(2 lt^2), 1/(2 lw^2), ybar, m
    for (int i = tid; i < m; i += GP_THREADS) {
        st[i] = s.alpha[i];
        st[cap + i] = s.x[i];
        if (two_d) st[2 * cap + i] = s.w[i];
    }
    if (tid == 0) {
        int o = two_d ? 3 * cap : 2 * cap;
        st[o++] = h.amp;
        st[o++] = h.it;
        if (two_d) st[o++] = h.iw;
        st[o++] = ybar;
        st[o] = (double)m;
    }
    // predictive mean, time derivative and std at the query grid
    for (int k = tid; k < a.n_query; k += GP_THREADS) {
        const double t = q[k], wq = qw ? qw[k] : 0.0;
        double mu = 0.0, dmu = 0.0;
        for (int i = 0; i < m; i++) {
            double kv = h.k(t, wq, s.x[i], s.w[i]);
            mu += kv * s.alpha[i];
            dmu += -2.0 * h.it * (t - s.x[i]) * kv * s.alpha[i];
        }
        double vtv = 0.0;  // v = L^-1 k*, var = amp - |v|^2 (floored at 1e-10)
        double vloc[GP_LDS_CAP];
        for (int i = 0; i < m; i++) {
            double si = h.k(t, wq, s.x[i], s.w[i]);
            for (int j = 0; j < i; j++) si -= s.L[tri(i, j)] * vloc[j];
            double vi = si / s.L[tri(i, i)];
            vloc[i] = vi;
            vtv += vi * vi;
        }
        const int64_t o = (int64_t)band * a.n_query + k;
        a.pred[o] = mu + ybar;
        a.stdv[o] = ::sqrt(fmax(h.amp - vtv, 1e-10));
        if (a.dpred) a.dpred[o] = dmu;
    }
    for (int i = tid; i < n; i += GP_THREADS) {  // predictive mean at every observation
        const double t = a.t[o0 + i], wi = two_d ? a.w[o0 + i] : 0.0;
        double mu = 0.0;
        for (int k = 0; k < m; k++) mu += h.k(t, wi, s.x[k], s.w[k]) * s.alpha[k];
        a.pred_obs[o0 + i] = mu + ybar;
    }
}

__global__ __launch_bounds__(GP_THREADS) void gp_fit_kernel(GpArgs a) {
    extern __shared__ double gp_lds[];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
        __syncthreads();
        const int t = next;
        __syncthreads();
        if (t >= a.n_tasks) break;
        gp_band(a, gp_lds, a.order[t]);
        __syncthreads();
    }
}

size_t gp_lds_bytes(int cap) {
    return sizeof(double) * ((size_t)cap * (cap + 1) / 2 + 5 * (size_t)cap + 8);
}

hipError_t launch_gp_fit(const GpArgs& a, int grid, hipStream_t st) {
    if (a.cap < 3 || a.cap > GP_LDS_CAP || grid <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gp_fit_kernel, dim3(grid), dim3(GP_THREADS), gp_lds_bytes(a.cap), st, a);
    return hipGetLastError();
}

hipError_t gp_occupancy(int cap, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)gp_fit_kernel,
                                                         GP_THREADS, gp_lds_bytes(cap));
}

}  // namespace lcf
