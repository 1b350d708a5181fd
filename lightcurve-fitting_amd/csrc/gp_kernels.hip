// gp_kernels.hip -- batched dense 1D Gaussian-process fits in FP64 for gfx950.
//
// One 256-thread workgroup owns one band at a time (persistent, bands pulled
// from a device queue in decreasing training-set size).  For every (amplitude,
// length-scale) candidate it builds the kernel matrix K + diag(noise) as a
// packed lower triangle in LDS (m <= 128: 66 KB), factorises it in place with
// a right-looking Cholesky whose column scaling and trailing updates are
// spread over all 256 threads, and solves for alpha on wavefront 0 with the
// right-hand side held in registers (lane l owns rows l and l+64), so the two
// triangular solves need no barriers.  The candidate is scored by the exact
// negative log marginal likelihood or by the train RMS; the winner is refitted
// and predicted (mean, derivative, standard deviation) at the query grid and
// at every observation of the band.
//
// Arithmetic follows DenseGP (src/sparse_gp.rs:16-176) and batch_gp_fit
// (cuda/gp.cu:23-367): the right-looking updates subtract l_ik l_jk in the
// same k order as the reference's Crout loops, so factors agree to rounding.
#include "gp_kernels.hpp"
#include "lcf_device.hpp"

namespace lcf {

constexpr int GP_THREADS = 256;

__device__ __forceinline__ int tri(int i, int j) { return (i * (i + 1) >> 1) + j; }

struct GpShared {
    double* L;      // packed lower triangle, cap*(cap+1)/2
    double* x;      // training times [cap]
    double* y;      // training values [cap]
    double* nv;     // training noise variances [cap]
    double* alpha;  // [cap]
    double* misc;   // scalars
};

__device__ __forceinline__ GpShared gp_carve(double* base, int cap) {
    GpShared s;
    s.L = base;
    s.x = s.L + cap * (cap + 1) / 2;
    s.y = s.x + cap;
    s.nv = s.y + cap;
    s.alpha = s.nv + cap;
    s.misc = s.alpha + cap;
    return s;
}

// Build K + diag(max(nv, 1e-10)) and factorise in place.  Returns false
// (uniformly) when a pivot is not positive.
__device__ bool gp_factor(const GpShared& s, int m, double amp, double i2) {
    const int tid = threadIdx.x;
    const int mm = m * m;
    for (int e = tid; e < mm; e += GP_THREADS) {
        int i = e / m, j = e - i * m;
        if (j <= i) {
            double d = s.x[i] - s.x[j];
            double v = amp * ::exp(-d * d * i2);
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
    // forward: L z = yc
    for (int j = 0; j < m; j++) {
        double bj = j < 64 ? bcast(b0, j) : bcast(b1, j - 64);
        double zj = bj / s.L[tri(j, j)];
        if (lane == (j & 63)) { if (j < 64) b0 = zj; else b1 = zj; }
        if (lane > j && lane < m) b0 -= s.L[tri(lane, j)] * zj;
        if (lane + 64 > j && lane + 64 < m) b1 -= s.L[tri(lane + 64, j)] * zj;
    }
    // backward: L^T a = z
    for (int j = m - 1; j >= 0; j--) {
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
__device__ double gp_train_rms(const GpShared& s, int m, double amp, double i2, double ybar) {
    const int lane = lane_id();
    double rss = 0.0;
    for (int i = lane; i < m; i += 64) {
        double mu = 0.0;
        for (int k = 0; k < m; k++) {
            double d = s.x[i] - s.x[k];
            mu += amp * ::exp(-d * d * i2) * s.alpha[k];
        }
        double r = mu + ybar - s.y[i];
        rss += r * r;
    }
    return ::sqrt(wave_sum(rss) / (double)m);
}

__device__ void gp_band(const GpArgs& a, double* lds, int band) {
    const int tid = threadIdx.x;
    const int cap = a.cap;
    GpShared s = gp_carve(lds, cap);
    const int64_t o0 = a.offsets[band];
    const int n = (int)(a.offsets[band + 1] - o0);
    double* st = a.state + (int64_t)band * a.state_stride;
    const double* q = a.query + (a.query_per_band ? (int64_t)band * a.n_query : 0);
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
        s.y[i] = a.y[o0 + idx];
        s.nv[i] = a.nv[o0 + idx];
    }
    __syncthreads();
    // y mean over the training set (all threads, same summation order)
    double ybar = 0.0;
    for (int i = 0; i < m; i++) ybar += s.y[i];
    ybar /= (double)m;

    double best = 1e99;
    int best_c = 0;
    bool any = false;
    const int nc = a.n_amp * a.n_ls;
    for (int c = 0; c < nc; c++) {
        const double amp = a.amps[c / a.n_ls], ls = a.lss[c % a.n_ls];
        double score = 1e99;
        if (ls >= 0.1) {
            const double i2 = 0.5 / (ls * ls);
            if (gp_factor(s, m, amp, i2)) {
                if (tid < 64) {
                    double half_fit_logdet = gp_solve_alpha(s, m, ybar);
                    double sc;
                    if (a.criterion == 0) sc = half_fit_logdet + 0.5 * (double)m * 1.83787706640934548356;
                    else sc = gp_train_rms(s, m, amp, i2, ybar);
                    if (tid == 0) s.misc[0] = __builtin_isfinite(sc) ? sc : 1e99;
                }
                __syncthreads();
                score = s.misc[0];
            }
            __syncthreads();
        }
        if (c == 0) any = ls >= 0.1;
        if (score < best) { best = score; best_c = c; any = true; }
    }
    if (!any) return;
    const double amp = a.amps[best_c / a.n_ls], ls = a.lss[best_c % a.n_ls];
    const double i2 = 0.5 / (ls * ls);
    bool ok = gp_factor(s, m, amp, i2);
    if (!ok) return;
    if (tid < 64) gp_solve_alpha(s, m, ybar);
    __syncthreads();
    // state: alpha[cap], x[cap], amp, 1/(2 l^2), ybar, m
    for (int i = tid; i < m; i += GP_THREADS) {
        st[i] = s.alpha[i];
        st[cap + i] = s.x[i];
    }
    if (tid == 0) {
        st[2 * cap] = amp;
        st[2 * cap + 1] = i2;
        st[2 * cap + 2] = ybar;
        st[2 * cap + 3] = (double)m;
    }
    // predictive mean, derivative and std at the query grid (one query per thread)
    for (int k = tid; k < a.n_query; k += GP_THREADS) {
        const double t = q[k];
        double mu = 0.0, dmu = 0.0;
        for (int i = 0; i < m; i++) {
            double d = t - s.x[i];
            double kv = amp * ::exp(-d * d * i2);
            mu += kv * s.alpha[i];
            dmu += -2.0 * i2 * d * kv * s.alpha[i];
        }
        // v = L^-1 k*, var = amp - |v|^2 (floored at 1e-10)
        double vtv = 0.0;
        double vloc[128];
        for (int i = 0; i < m; i++) {
            double d = t - s.x[i];
            double si = amp * ::exp(-d * d * i2);
            for (int j = 0; j < i; j++) si -= s.L[tri(i, j)] * vloc[j];
            double vi = si / s.L[tri(i, i)];
            vloc[i] = vi;
            vtv += vi * vi;
        }
        const int64_t o = (int64_t)band * a.n_query + k;
        a.pred[o] = mu + ybar;
        a.stdv[o] = ::sqrt(fmax(amp - vtv, 1e-10));
        if (a.dpred) a.dpred[o] = dmu;
    }
    // predictive mean at every observation of the band
    for (int i = tid; i < n; i += GP_THREADS) {
        const double t = a.t[o0 + i];
        double mu = 0.0;
        for (int k = 0; k < m; k++) {
            double d = t - s.x[k];
            mu += amp * ::exp(-d * d * i2) * s.alpha[k];
        }
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
    return sizeof(double) * ((size_t)cap * (cap + 1) / 2 + 4 * (size_t)cap + 8);
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
