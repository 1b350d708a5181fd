// gp_kernels.hip -- batched dense Gaussian-process fits in FP64 for gfx950,
// 1D (time) and 2D (time x log10 wavelength, anisotropic squared exponential).
//
// One 256-thread workgroup owns one band (1D) or source (2D) at a time
// (persistent, pulled from a device queue in decreasing training-set size).
// For every hyper-parameter candidate it builds K + diag(noise) as a packed
// lower triangle in LDS (m <= 200: 157 KB), factorises it in place with a
// right-looking Cholesky whose column scaling and trailing updates are spread
// over all 256 threads, and solves for alpha on wavefront 0 with the right-hand
// side held in registers (lane l owns rows l, l+64, l+128, l+192), so the
// triangular solves need no barriers.  The training set is addressed through
// an index list in LDS into the band's observations (L1/L2 resident), which
// leaves the whole LDS budget to the factor.  Candidates are scored by the
// exact negative log marginal likelihood or by the train RMS; the winner is
// refitted and predicted at the query grid (mean and time derivative per
// thread; standard deviation by a wave-level triangular solve per query) and
// at every observation.
//
// Arithmetic follows DenseGP (src/sparse_gp.rs:16-176), batch_gp_fit
// (cuda/gp.cu:23-367) and batch_gp2d_fit_predict (cuda/gp2d.cu:21-359): the
// right-looking updates subtract l_ik l_jk in the same k order as the
// reference's Crout loops, so factors agree to rounding.
#include "gp_kernels.hpp"
#include "gp_device.hpp"

namespace lcf {

struct GpShared {
    double* L;      // packed lower triangle, cap*(cap+1)/2
    double* alpha;  // [cap]
    double* misc;   // scalars
    int* idx;       // [cap] training-set indices into the band
};

__device__ __forceinline__ GpShared gp_carve(double* base, int cap) {
    GpShared s;
    s.L = base;
    s.alpha = s.L + cap * (cap + 1) / 2;
    s.misc = s.alpha + cap;
    s.idx = reinterpret_cast<int*>(s.misc + 8);
    return s;
}

// The band's training set, read through the subsample indices.
struct Train {
    const double *t, *w, *y, *nv;
    const int* idx;
    __device__ __forceinline__ double x(int i) const { return t[idx[i]]; }
    __device__ __forceinline__ double wl(int i) const { return w ? w[idx[i]] : 0.0; }
    __device__ __forceinline__ double yy(int i) const { return y[idx[i]]; }
    __device__ __forceinline__ double v(int i) const { return nv[idx[i]]; }
};

struct Hyp {
    double amp, it, iw;  // amplitude, 1/(2 l_t^2), 1/(2 l_w^2) (0 in 1D)
    __device__ __forceinline__ double k(double t1, double w1, double t2, double w2) const {
        double dt = t1 - t2, dw = w1 - w2;
        return amp * ::exp(-dt * dt * it - dw * dw * iw);
    }
};

// Build K + diag(max(nv, 1e-10)) and factorise in place.  Returns false
// (uniformly) when a pivot is not positive.
__device__ bool gp_factor(const GpShared& s, const Train& tr, int m, const Hyp& h) {
    const int tid = threadIdx.x;
    const int mm = m * m;
    for (int e = tid; e < mm; e += GP_THREADS) {
        int i = e / m, j = e - i * m;
        if (j <= i) {
            double v = h.k(tr.x(i), tr.wl(i), tr.x(j), tr.wl(j));
            if (i == j) v += fmax(tr.v(i), 1e-10);
            s.L[tri(i, j)] = v;
        }
    }
    __syncthreads();
    return chol_packed(s.L, m);
}

// Wavefront 0: alpha = L^-T L^-1 (y - ybar); returns 0.5 yc.alpha + sum ln L_ii
// (wave-uniform).  Lane l owns rows l + 64 r in registers.
__device__ double gp_solve_alpha(const GpShared& s, const Train& tr, int m, double ybar) {
    const int lane = lane_id();
    double b[GP_ROWS], yc[GP_ROWS];
#pragma unroll
    for (int r = 0; r < GP_ROWS; r++) {
        int i = lane + WAVE * r;
        b[r] = i < m ? tr.yy(i) - ybar : 0.0;
        yc[r] = b[r];
    }
    for (int j = 0; j < m; j++) {  // forward: L z = yc
        double zj = row_of(b, j) / s.L[tri(j, j)];
#pragma unroll
        for (int r = 0; r < GP_ROWS; r++) {
            int i = lane + WAVE * r;
            if (i == j) b[r] = zj;
            else if (i > j && i < m) b[r] -= s.L[tri(i, j)] * zj;
        }
    }
    for (int j = m - 1; j >= 0; j--) {  // backward: L^T a = z
        double aj = row_of(b, j) / s.L[tri(j, j)];
#pragma unroll
        for (int r = 0; r < GP_ROWS; r++) {
            int i = lane + WAVE * r;
            if (i == j) b[r] = aj;
            else if (i < j) b[r] -= s.L[tri(j, i)] * aj;
        }
    }
    double part = 0.0;
#pragma unroll
    for (int r = 0; r < GP_ROWS; r++) {
        int i = lane + WAVE * r;
        if (i < m) {
            s.alpha[i] = b[r];
            part += yc[r] * b[r] + ::log(s.L[tri(i, i)]) * 2.0;
        }
    }
    return 0.5 * wave_sum(part);
}

// Wavefront 0: RMS of the predictive mean at the training points.
__device__ double gp_train_rms(const GpShared& s, const Train& tr, int m, const Hyp& h, double ybar) {
    const int lane = lane_id();
    double rss = 0.0;
    for (int i = lane; i < m; i += WAVE) {
        const double xi = tr.x(i), wi = tr.wl(i);
        double mu = 0.0;
        for (int k = 0; k < m; k++) mu += h.k(xi, wi, tr.x(k), tr.wl(k)) * s.alpha[k];
        double r = mu + ybar - tr.yy(i);
        rss += r * r;
    }
    return ::sqrt(wave_sum(rss) / (double)m);
}

__device__ __forceinline__ Hyp candidate(const GpArgs& a, int band, int c, bool& valid) {
    const int nw = a.n_lsw > 0 ? a.n_lsw : 1;
    const int ia = c / (a.n_ls * nw), il = (c / nw) % a.n_ls, iw = c % nw;
    Hyp h;
    h.amp = a.amps[ia];
    const double ls = a.lss_band ? a.lss_band[(int64_t)band * a.n_ls + il] : a.lss[il];
    valid = ls >= a.ls_floor && (!a.min_ls || ls >= a.min_ls[band]);
    h.it = 0.5 / (ls * ls);
    h.iw = 0.0;
    if (a.w) {
        const double lw = a.lsws_band ? a.lsws_band[(int64_t)band * a.n_lsw + iw] : a.lsws[iw];
        valid = valid && lw >= a.lw_floor;
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
        } else if (a.subsample_rule == 1) {
            idx = (int)(((double)i + 0.5) * ((double)n / (double)m));
        } else {  // fit_gp_2d (src/gp2d.rs:373-385)
            idx = (int)((double)i * ((double)n / (double)m));
        }
        s.idx[i] = idx < n - 1 ? idx : n - 1;
    }
    __syncthreads();
    Train tr{a.t + o0, two_d ? a.w + o0 : nullptr, a.y + o0, a.nv + o0, s.idx};
    double ybar = 0.0;  // all threads, same summation order
    for (int i = 0; i < m; i++) ybar += tr.yy(i);
    ybar /= (double)m;

    double best = 1e99;
    int best_c = 0;
    bool any = false;
    const int nc = a.n_amp * a.n_ls * (a.n_lsw > 0 ? a.n_lsw : 1);
    for (int c = 0; c < nc; c++) {
        bool valid;
        const Hyp h = candidate(a, band, c, valid);
        double score = 1e99;
        if (valid) {
            if (gp_factor(s, tr, m, h)) {
                if (tid < WAVE) {
                    double half_fit_logdet = gp_solve_alpha(s, tr, m, ybar);
                    double sc;
                    if (a.criterion == 0) sc = half_fit_logdet + 0.5 * (double)m * 1.83787706640934548356;
                    else sc = gp_train_rms(s, tr, m, h, ybar);
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
    // the CPU dense path keeps no fit when no candidate has a finite score
    if (a.criterion == 1 && !(best < 1e99)) return;
    bool valid;
    const Hyp h = candidate(a, band, best_c, valid);
    if (!gp_factor(s, tr, m, h)) return;
    if (tid < WAVE) gp_solve_alpha(s, tr, m, ybar);
    __syncthreads();
    // state 1D: alpha[cap], x[cap], amp, 1/(2 lt^2), ybar, m
    //       2D: alpha[cap], x[cap], w[cap], amp, 1/(2 lt^2), 1/(2 lw^2), ybar, m
    for (int i = tid; i < m; i += GP_THREADS) {
        st[i] = s.alpha[i];
        st[cap + i] = tr.x(i);
        if (two_d) st[2 * cap + i] = tr.wl(i);
    }
    if (tid == 0) {
        int o = two_d ? 3 * cap : 2 * cap;
        st[o++] = h.amp;
        st[o++] = h.it;
        if (two_d) st[o++] = h.iw;
        st[o++] = ybar;
        st[o] = (double)m;
    }
    // predictive mean and its time derivative at the query grid (one query per thread)
    for (int k = tid; k < a.n_query; k += GP_THREADS) {
        const double t = q[k], wq = qw ? qw[k] : 0.0;
        double mu = 0.0, dmu = 0.0;
        for (int i = 0; i < m; i++) {
            const double xi = tr.x(i);
            double kv = h.k(t, wq, xi, tr.wl(i));
            mu += kv * s.alpha[i];
            dmu += -2.0 * h.it * (t - xi) * kv * s.alpha[i];
        }
        const int64_t o = (int64_t)band * a.n_query + k;
        a.pred[o] = mu + ybar;
        if (a.dpred) a.dpred[o] = dmu;
    }
    // predictive std: v = L^-1 k*, var = amp - |v|^2 floored at 1e-10; one
    // query per wavefront, lanes holding the rows of v
    {
        const int wave = tid / WAVE, lane = tid - wave * WAVE;
        for (int k = wave; k < a.n_query; k += GP_WAVES) {
            const double t = q[k], wq = qw ? qw[k] : 0.0;
            double sv[GP_ROWS];
#pragma unroll
            for (int r = 0; r < GP_ROWS; r++) {
                int i = lane + WAVE * r;
                sv[r] = i < m ? h.k(t, wq, tr.x(i), tr.wl(i)) : 0.0;
            }
            double vtv = 0.0;
            for (int j = 0; j < m; j++) {
                double vj = row_of(sv, j) / s.L[tri(j, j)];
                vtv += vj * vj;
#pragma unroll
                for (int r = 0; r < GP_ROWS; r++) {
                    int i = lane + WAVE * r;
                    if (i > j && i < m) sv[r] -= s.L[tri(i, j)] * vj;
                }
            }
            if (lane == 0) a.stdv[(int64_t)band * a.n_query + k] = ::sqrt(fmax(h.amp - vtv, 1e-10));
        }
    }
    for (int i = tid; i < n; i += GP_THREADS) {  // predictive mean at every observation
        const double t = a.t[o0 + i], wi = two_d ? a.w[o0 + i] : 0.0;
        double mu = 0.0;
        for (int k = 0; k < m; k++) mu += h.k(t, wi, tr.x(k), tr.wl(k)) * s.alpha[k];
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
    return sizeof(double) * ((size_t)cap * (cap + 1) / 2 + (size_t)cap + 8) + sizeof(int) * (size_t)cap;
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
