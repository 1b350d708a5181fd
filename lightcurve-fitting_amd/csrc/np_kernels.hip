// np_kernels.hip -- batched nonparametric band pipeline for gfx950, FP64.
//
// One 256-thread workgroup owns one band at a time (persistent, pulled from a
// device queue in decreasing size) and runs the whole of process_band
// (src/nonparametric.rs:1228-1525) on it:
//   1. detections (SNR >= 3) compacted in order with wave ballots; the minimum
//      length scale from the median positive gap of the sorted detection
//      times (rank sort, rank selection);
//   2. every (amp, l) candidate: dense GP on all detections (packed Cholesky
//      in LDS, blocked over the workgroup; alpha by wave-level solves)
//      scored by train RMS (from the weights: residual -D_i alpha_i), or
//      above the dense threshold the FITC
//      approximation (inducing-point factor in LDS; the rank-1 updates of B
//      accumulated per chunk of 64 observations, each entry in observation
//      order) scored by approx_nlml; for a FITC band also the subsampled
//      dense GP kept for thermal reuse;
//   3. predictive mean / derivative per query thread and predictive std
//      (dense: 16 queries per wavefront, blocked forward substitution on the
//      FP64 matrix cores) on the grid, the mean at every detection (residual RMS,
//      chi2, baseline chi2), the std at (a 200-point subsample of) the
//      detection times and the std rescale;
//   4. the peak on the grid, the 1-day fine grid (<= 2000 points, LDS) and
//      its crossing / slope metrics by workgroup min / max / sum reductions,
//      the predictive features, the power-law decay fit, magnitudes at fixed
//      offsets and the variability features of the raw band.
// Arithmetic follows the reference's expressions (this unit is built without
// multiply-add contraction); sums over observations or grid points are
// workgroup reductions, so they round in a different order than the
// reference's sequential loops.
#include "np_kernels.hpp"
#include "gp_device.hpp"

namespace lcf {

namespace {

constexpr double MAG_FACTOR = 1.0857362047581294;   // 2.5 / ln 10
constexpr double OFF_EFOLD = 1.0857362047581294;    // 2.5 log10(e)
constexpr double OFF_HALF = 0.7525749891599529;     // 2.5 log10(2)
constexpr int FITC_TRI = NP_MAX_IND * (NP_MAX_IND + 1) / 2;
constexpr int FITC_DOUBLES = 2 * FITC_TRI + NP_CHUNK * NP_MAX_IND + 2 * NP_CHUNK + NP_MAX_IND;
constexpr int MISC_DOUBLES = 64;

__host__ __device__ constexpr int tri_n(int c) { return c * (c + 1) / 2; }

__host__ __device__ inline int alpha_doubles(int cap) { return cap > 2 * NP_MAX_IND ? cap : 2 * NP_MAX_IND; }

// r0 holds, in turn: the packed factor (tri_n(cap)), the FITC work area, the
// fine grid (NP_FINE_MAX); its last alpha_doubles(cap) doubles are the weight
// slot (alpha, or FITC inducing points + weights), written once the factor is
// no longer read.
__host__ __device__ inline int r0_doubles(int cap, int fitc) {
    const int ac = alpha_doubles(cap);
    int r = tri_n(cap);
    if (r < NP_FINE_MAX + ac) r = NP_FINE_MAX + ac;
    if (fitc && r < FITC_DOUBLES + ac) r = FITC_DOUBLES + ac;
    return r;
}

// misc slots
enum {
    M_RED = 0,      // [4] reduction partials
    M_RED2 = 4,     // [4]
    M_SCAL = 8,     // scalar broadcast slots
    M_Q = 16,       // [9] predictive queries
    M_INT = 32,     // int view starts here (16 doubles = 32 ints)
    M_RINV = 48,    // [CHOL_NB] pivot reciprocals of the blocked Cholesky
};

struct Lds {
    double* r0;
    double* al;     // weight slot at the end of r0
    double* xs;     // [cap] training times of a dense band
    double* misc;
    int* imisc;
};

__device__ __forceinline__ Lds carve(double* base, int cap, int fitc) {
    Lds s;
    const int r0n = r0_doubles(cap, fitc);
    s.r0 = base;
    s.al = base + r0n - alpha_doubles(cap);
    s.xs = base + r0n;
    s.misc = s.xs + cap;
    s.imisc = reinterpret_cast<int*>(s.misc + M_INT);
    return s;
}

// rbf (src/sparse_gp.rs:17-20)
__device__ __forceinline__ double rbf(double x1, double x2, double amp, double it) {
    double d = x1 - x2;
    return amp * ::exp(-d * d * it);
}

// detection_mask through mag2flux (src/nonparametric.rs:46-57, src/common.rs:32-36)
__device__ __forceinline__ bool is_detection(double m, double e) {
    double flux = ::pow(10.0, -0.4 * (m - 23.9));
    double ferr = e / MAG_FACTOR * flux;
    return flux > 0.0 && ferr > 0.0 && (flux / ferr) >= 3.0;
}

// Detections of the band: times (LDS when staged, else the global compacted
// copy), magnitudes and errors (global compacted copies).
struct Det {
    const double *xs, *gx, *y, *e;  // gx: the times in global memory
    __device__ __forceinline__ double x(int i) const { return xs[i]; }
    __device__ __forceinline__ double xg(int i) const { return gx[i]; }
    __device__ __forceinline__ double yy(int i) const { return y[i]; }
    __device__ __forceinline__ double ee(int i) const { return e[i]; }
    __device__ __forceinline__ double nv(int i) const { return fmax(e[i] * e[i], 1e-6); }
};

// The subsample of src/gp.rs:5-27 of the detections: k(i) = floor((i + 1/2)
// nd / ms) (identity when nd <= ms); times staged in LDS in subsample order.
struct DetSub {
    const double *xs, *gx, *y, *e;
    int nd, sub;
    double step;
    __device__ __forceinline__ int k(int i) const {
        if (!sub) return i;
        int q = (int)::floor(((double)i + 0.5) * step);
        return q < nd - 1 ? q : nd - 1;
    }
    __device__ __forceinline__ double x(int i) const { return xs[i]; }
    __device__ __forceinline__ double xg(int i) const { return gx[k(i)]; }
    __device__ __forceinline__ double yy(int i) const { return y[k(i)]; }
    __device__ __forceinline__ double nv(int i) const {
        double v = e[k(i)];
        return fmax(v * v, 1e-6);
    }
};

// The fitted predictor: dense over x = detections (times in LDS) or FITC over
// inducing points z (LDS).  Every pointer refers to LDS and none is null: a
// null merged with an LDS address makes the compiler test the LDS aperture,
// which it cannot encode here.
template <int ROWS>
struct Pred {
    int sparse, m;
    double amp, it, ybar;
    Det d;               // dense training set
    const double* z;     // FITC inducing points
    const double* alpha; // weights
    const double* L;     // dense factor / FITC K_mm factor
    const double* Lb;    // FITC B factor
    __device__ __forceinline__ double xk(int k) const { return sparse ? z[k] : d.x(k); }
    __device__ __forceinline__ double mean(double t) const {
        double dot = 0.0;
        for (int k = 0; k < m; k++) dot += rbf(t, xk(k), amp, it) * alpha[k];
        return dot + ybar;
    }
    __device__ __forceinline__ double dmean(double t) const {
        double s = 0.0;
        for (int k = 0; k < m; k++) {
            double xk_ = xk(k);
            s += -2.0 * it * (t - xk_) * rbf(t, xk_, amp, it) * alpha[k];
        }
        return s;
    }
    // One wavefront, four wave-uniform query times at once (independent
    // chains hide the latency of the lane reads); the std of each in every
    // lane.  var = amp - |L^-1 k*|^2 (+ |L_B^-1 k*|^2 for FITC); the solve
    // multiplies by 1 / l_jj instead of dividing (within an ulp per step; the
    // reference divides).
    __device__ __forceinline__ void wave_std4(const double (&tq)[4], double (&sd)[4]) const {
        double q[4];
        wave_sqnorm4(L, tq, q);
        if (!sparse) {
#pragma unroll
            for (int u = 0; u < 4; u++) sd[u] = ::sqrt(fmax(amp - q[u], 1e-10));
            return;
        }
        double qb[4];
        wave_sqnorm4(Lb, tq, qb);
#pragma unroll
        for (int u = 0; u < 4; u++) sd[u] = ::sqrt(fmax(amp - q[u] + qb[u], 1e-10));
    }
    // Dense predictor, one wavefront: the predictive std of 16 queries --
    // lane l serves query l & 15, whose time tq is -- in every lane of the
    // query's column.  v = L^-1 k* by 16-row blocks I: the finished blocks'
    // contribution sum_{J<I} L_IJ v_J as v_mfma_f64_16x16x4f64 tile products
    // (A: the factor's tile, read from LDS; B: v_J, kept in registers in the
    // product's own output layout, whose element e is the B operand of the
    // k-step e), then the block's forward substitution, each solved row
    // multiplied by its pivot's reciprocal and sent from its lane group to
    // the other three.  var = amp - |v|^2, floored at 1e-10 (the reference's
    // predict_with_std, src/sparse_gp.rs:131-158, solves column by column;
    // the sums here round in tile order).  Layouts: A lane l -> A[l % 16][l / 16],
    // B lane l -> B[l / 16][l % 16], D lane l element e -> D[4 e + l / 16][l % 16].
    // NB: the block capacity (m <= 16 NB).
    template <int NB>
    __device__ __forceinline__ double dense_std16(double tq) const {
        const int lane = lane_id();
        const int g = lane >> 4, c = lane & 15;
        const int nb = (m + 15) >> 4;
        double v[NB][4];
        double q = 0.0;
#pragma unroll
        for (int I = 0; I < NB; I++) {
            if (I < nb) {
                const int i0 = 16 * I;
                double rhs[4];
                int rr[4];  // start of this lane's rows at column i0 (row 0 for rows past m)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int i = i0 + 4 * e + g;
                    rhs[e] = i < m ? rbf(tq, d.x(i), amp, it) : 0.0;
                    rr[e] = rs(i < m ? i : 0) + i0;
                }
                if (I > 0) {
                    const int ia = i0 + c;
                    const bool a_ok = ia < m;
                    const int ra = rs(a_ok ? ia : 0) + g;
                    v4d acc = {0.0, 0.0, 0.0, 0.0};
                    // constant trip count, so that v[J] is indexed at compile time
#pragma unroll
                    for (int J = 0; J < NB - 1; J++) {
                        if (J >= I) continue;
#pragma unroll
                        for (int s4 = 0; s4 < 4; s4++) {
                            const double av = a_ok ? L[ra + 16 * J + 4 * s4] : 0.0;
                            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, v[J][s4], acc, 0, 0, 0);
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 4; e++) rhs[e] -= acc[e];
                }
                // lanes 0..15: the reciprocal of pivot i0 + lane (0 past m, so
                // rows past m solve to 0)
                const int jd = i0 + c;
                const double rdl = jd < m ? 1.0 / L[rs(jd) + jd] : 0.0;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const int gr = r & 3, er = r >> 2;
                    const double x = rhs[er] * bcast(rdl, r);  // row r, in lane group gr
                    const double xb = __shfl(x, (gr << 4) | c, WAVE);
                    if (g == gr) rhs[er] = x;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        if (4 * e + 3 <= r) continue;  // rows already solved in every group
                        const double lv = L[rr[e] + r];
                        if (4 * e + g > r) rhs[e] = __builtin_fma(-lv, xb, rhs[e]);
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    v[I][e] = rhs[e];
                    q = __builtin_fma(rhs[e], rhs[e], q);
                }
            }
        }
        q += __shfl_xor(q, 16, WAVE);
        q += __shfl_xor(q, 32, WAVE);
        return ::sqrt(fmax(amp - q, 1e-10));
    }
    __device__ __forceinline__ void wave_sqnorm4(const double* F, const double (&tq)[4], double (&q)[4]) const {
        const int lane = lane_id();
        double b[4][ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
            const int i = lane + WAVE * r;
            const double xi = i < m ? xk(i) : 0.0;
#pragma unroll
            for (int u = 0; u < 4; u++) b[u][r] = i < m ? rbf(tq[u], xi, amp, it) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = 0.0;
        double rd[ROWS];  // pivot reciprocals of this lane's rows, formed once
        int rsi[ROWS];
        wave_rdiag(F, m, rd);
#pragma unroll
        for (int r = 0; r < ROWS; r++) rsi[r] = rs(lane + WAVE * r);
        sqnorm4_seg<0>(F, b, rd, rsi, q);
        if constexpr (ROWS > 1) if (m > 64) sqnorm4_seg<1>(F, b, rd, rsi, q);
        if constexpr (ROWS > 2) if (m > 128) sqnorm4_seg<2>(F, b, rd, rsi, q);
        if constexpr (ROWS > 3) if (m > 192) sqnorm4_seg<3>(F, b, rd, rsi, q);
    }
    // Forward substitution of the four right-hand sides over the rows j of
    // register R (j in [64 R, 64 R + 64)); the multiply-adds are fused (an
    // ulp from the reference's separate rounding, far inside the std's
    // tolerance).
    template <int R>
    __device__ __forceinline__ void sqnorm4_seg(const double* F, double (&b)[4][ROWS], const double (&rd)[ROWS],
                                                const int (&rsi)[ROWS], double (&q)[4]) const {
        const int lane = lane_id();
        const int jhi = m < 64 * R + 64 ? m : 64 * R + 64;
        for (int j = 64 * R; j < jhi; j++) {
            const int src = j & 63;
            const double rljj = bcast(rd[R], src);
            double z[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                z[u] = bcast(b[u][R], src) * rljj;
                q[u] = __builtin_fma(z[u], z[u], q[u]);
            }
            if (lane > src && lane + 64 * R < m) {
                const double lij = F[rsi[R] + j];
#pragma unroll
                for (int u = 0; u < 4; u++) b[u][R] = __builtin_fma(-lij, z[u], b[u][R]);
            }
#pragma unroll
            for (int r = R + 1; r < ROWS; r++)
                if (lane + 64 * r < m) {
                    const double lij = F[rsi[r] + j];
#pragma unroll
                    for (int u = 0; u < 4; u++) b[u][r] = __builtin_fma(-lij, z[u], b[u][r]);
                }
        }
    }
};

// ---- dense GP (src/sparse_gp.rs:83-116, 168-175) ----------------------------

// K + diag(max(nv, 1e-10)) with y - ybar appended as row m, factorised:
// the factor in rows 0..m-1, L^-1 (y - ybar) left in row m.  Row m may cover
// the staged training times (a band of cap detections): restage them before
// reading x again.
template <typename D>
__device__ __forceinline__ bool dense_factor(double* L, const D& d, int m, double amp, double it, double ybar,
                                             int* flag, double* rinv) {
    const int ne = rs(m);
    for (int e = threadIdx.x; e < ne; e += GP_THREADS) {
        int i, j;
        tri_ij(e, i, j);
        double v = rbf(d.x(i), d.x(j), amp, it);
        if (i == j) v += fmax(d.nv(i), 1e-10);
        L[e] = v;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += GP_THREADS) L[ne + j] = d.yy(j) - ybar;
    __syncthreads();
    return chol_blocked(L, m, m + 1, flag, rinv);
}

// Wavefront 0: alpha = L^-T z with z = L^-1 (y - ybar) from row m, written
// to alpha after the solve (alpha may overlap the tail of L).
template <int ROWS>
__device__ __forceinline__ void dense_alpha(const double* L, int m, double* alpha) {
    const int lane = lane_id();
    const int rz = rs(m);
    double b[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        int i = lane + WAVE * r;
        b[r] = i < m ? L[rz + i] : 0.0;
    }
    wave_solve_lt(L, m, b);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        int i = lane + WAVE * r;
        if (i < m) alpha[i] = b[r];
    }
}

// All threads: a barrier, the training times back into LDS where the
// augmented row m of the factor (packed at [rs(m), rs(m) + m) of r0, which
// xs follows) reached them, then a barrier.
template <typename D>
__device__ __forceinline__ void restage(const double* r0, double* xs, const D& d, int m) {
    __syncthreads();
    if (rs(m) + m <= (int)(xs - r0)) return;
    for (int i = threadIdx.x; i < m; i += GP_THREADS) xs[i] = d.xg(i);
    __syncthreads();
}

// Train RMS (DenseGP::train_rms, src/sparse_gp.rs:169-175) from the weights
// alone.  With K alpha = y - ybar and K = A + D (A the noise-free kernel
// matrix at the training times, D = diag(max(nv, 1e-10))), the prediction at
// training point i, (A alpha)_i + ybar, equals y_i - D_i alpha_i: the
// residual is -D_i alpha_i.  m products instead of the m^2 kernel evaluations
// of the explicit predictions, whose sum also cancels down to that residual
// (relative rounding there grows with amp / D_i; here it does not).
__device__ __forceinline__ double dense_train_rms(const Det& d, int m, const double* alpha, double* red) {
    double rss = 0.0;
    for (int i = threadIdx.x; i < m; i += GP_THREADS) {
        const double r = fmax(d.nv(i), 1e-10) * alpha[i];
        rss += r * r;
    }
    rss = block_sum(rss, red);
    return ::sqrt(rss / (double)m);
}

// ---- FITC (src/sparse_gp.rs:195-271, 315-327) --------------------------------

struct Fitc {
    double *lmm, *lb, *kc, *li, *lv, *w;  // in r0
    double *z, *az;                        // in the weight slot
    int* flag;
    double* rinv;                          // pivot reciprocals (misc)
};

__device__ __forceinline__ Fitc fitc_carve(const Lds& s) {
    Fitc f;
    f.lmm = s.r0;
    f.lb = f.lmm + FITC_TRI;
    f.kc = f.lb + FITC_TRI;
    f.li = f.kc + NP_CHUNK * NP_MAX_IND;
    f.lv = f.li + NP_CHUNK;
    f.w = f.lv + NP_CHUNK;
    f.z = s.al;
    f.az = s.al + NP_MAX_IND;
    f.flag = s.imisc + 20;
    f.rinv = s.misc + M_RINV;
    return f;
}

__device__ __forceinline__ bool fitc_fit(const Fitc& f, const Det& d, int nd, int m, double amp, double it, double ybar,
                         double tmin, double tmax) {
    const int tid = threadIdx.x;
    if (tid < m) f.z[tid] = tmin + (double)tid * (tmax - tmin) / (double)(m - 1 > 1 ? m - 1 : 1);
    __syncthreads();
    for (int e = tid; e < m * m; e += GP_THREADS) {
        int i = e / m, j = e - i * m;
        if (j <= i) {
            double v = rbf(f.z[i], f.z[j], amp, it);
            if (i == j) v += 1e-6;
            f.lmm[tri(i, j)] = v;
            f.lb[tri(i, j)] = v;
        }
    }
    if (tid < m) f.w[tid] = 0.0;
    __syncthreads();
    if (!chol_blocked(f.lmm, m, m, f.flag, f.rinv)) return false;
    for (int j0 = 0; j0 < nd; j0 += NP_CHUNK) {
        const int cn = nd - j0 < NP_CHUNK ? nd - j0 : NP_CHUNK;
        for (int e = tid; e < cn * m; e += GP_THREADS) {
            int jj = e / m, i = e - jj * m;
            f.kc[jj * NP_MAX_IND + i] = rbf(f.z[i], d.x(j0 + jj), amp, it);
        }
        __syncthreads();
        if (tid < cn) {
            // v = L_mm^-1 k_j in registers, Q_jj = |v|^2 in row order
            const double* kr = f.kc + tid * NP_MAX_IND;
            double v[NP_MAX_IND];
            double q = 0.0;
#pragma unroll
            for (int i = 0; i < NP_MAX_IND; i++) {
                if (i < m) {
                    double s = kr[i];
#pragma unroll
                    for (int k = 0; k < i; k++) s -= f.lmm[tri(i, k)] * v[k];
                    v[i] = s / f.lmm[tri(i, i)];
                } else {
                    v[i] = 0.0;
                }
            }
#pragma unroll
            for (int i = 0; i < NP_MAX_IND; i++)
                if (i < m) q += v[i] * v[i];
            const int j = j0 + tid;
            double lam = fmax(amp - q, 0.0) + d.nv(j);
            double li = 1.0 / fmax(lam, 1e-12);
            f.li[tid] = li;
            f.lv[tid] = li * (d.yy(j) - ybar);
        }
        __syncthreads();
        for (int e = tid; e < m * m; e += GP_THREADS) {  // B[r][c] += (li k_r) k_c, c >= r, stored at (c, r)
            int r = e / m, c = e - r * m;
            if (c >= r) {
                double acc = f.lb[tri(c, r)];
                for (int jj = 0; jj < cn; jj++) {
                    const double* kr = f.kc + jj * NP_MAX_IND;
                    acc += f.li[jj] * kr[r] * kr[c];
                }
                f.lb[tri(c, r)] = acc;
            }
        }
        if (tid < m) {
            double acc = f.w[tid];
            for (int jj = 0; jj < cn; jj++) acc += f.kc[jj * NP_MAX_IND + tid] * f.lv[jj];
            f.w[tid] = acc;
        }
        __syncthreads();
    }
    if (!chol_blocked(f.lb, m, m, f.flag, f.rinv)) return false;
    if (tid < WAVE) {
        double b[1];  // m <= NP_MAX_IND <= 64: one row per lane
        b[0] = tid < m ? f.w[tid] : 0.0;
        wave_solve_l(f.lb, m, b);
        wave_solve_lt(f.lb, m, b);
        if (tid < m) f.az[tid] = b[0];
    }
    __syncthreads();
    return true;
}

__device__ __forceinline__ double fitc_nlml(const Fitc& f, const Det& d, int nd, int m, double amp, double it, double ybar,
                            double* red) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nd; i += GP_THREADS) {
        const double ti = d.x(i);
        double dot = 0.0;
        for (int k = 0; k < m; k++) dot += rbf(ti, f.z[k], amp, it) * f.az[k];
        double r = d.yy(i) - (dot + ybar);
        double nv = d.nv(i);
        acc += r * r / nv + ::log(nv);
    }
    return block_sum(acc, red) / (double)nd;
}

// ---- curve helpers over an array in LDS (src/common.rs:168-460) -------------

struct Curve {
    const double* m;  // values
    int n;
    double t0, step;  // times t0 + i step
    __device__ __forceinline__ double t(int i) const { return t0 + (double)i * step; }
};

// First i in [a, b) with m[i] >= thr, or INT_MAX.
__device__ __forceinline__ int first_at_least(const Curve& c, int a, int b, double thr, int* ired) {
    int best = 0x7fffffff;
    for (int i = a + threadIdx.x; i < b; i += GP_THREADS)
        if (c.m[i] >= thr && i < best) best = i;
    return block_min_int(best, ired);
}

// Last i in [a, b) with m[i] >= thr, or -1.
__device__ __forceinline__ int last_at_least(const Curve& c, int a, int b, double thr, int* ired) {
    int best = -1;
    for (int i = a + threadIdx.x; i < b; i += GP_THREADS)
        if (c.m[i] >= thr && i > best) best = i;
    return block_max_int(best, ired);
}

// Least-squares slope over i in [a, b) (all threads).
__device__ __forceinline__ double slope_range(const Curve& c, int a, int b, double tiny, double* red) {
    const int n = b - a;
    if (n < 2) return NAN;
    double st = 0, sm = 0, stt = 0, stm = 0;
    for (int i = a + threadIdx.x; i < b; i += GP_THREADS) {
        double ti = c.t(i), mi = c.m[i];
        st += ti;
        sm += mi;
        stt += ti * ti;
        stm += ti * mi;
    }
    st = block_sum(st, red);
    sm = block_sum(sm, red);
    stt = block_sum(stt, red);
    stm = block_sum(stm, red);
    const double nf = (double)n;
    double den = nf * stt - st * st;
    if (fabs(den) < tiny) return NAN;
    return (nf * stm - st * sm) / den;
}

// compute_fine_grid_metrics' eleven metrics (FineGridMetrics order).
__device__ __forceinline__ void curve_metrics(const Curve& c, int pk, double* red, int* ired, double* o) {
    const int n = c.n;
    const double pm = c.m[pk];
    // decay crossings (find_decay_crossing)
    const double offs[2] = {OFF_EFOLD, OFF_HALF};
    for (int q = 0; q < 2; q++) {
        double r = NAN;
        if (pk < (n - 1 > 0 ? n - 1 : 0)) {
            const double thr = pm + offs[q];
            int i = first_at_least(c, pk + 1, n, thr, ired);
            if (i != 0x7fffffff) {
                double dm = c.m[i] - c.m[i - 1];
                if (fabs(dm) < 1e-15) r = c.t(i) - c.t(pk);
                else {
                    double fr = (thr - c.m[i - 1]) / dm;
                    double tc = c.t(i - 1) + fr * (c.t(i) - c.t(i - 1));
                    r = tc - c.t(pk);
                }
            }
        }
        o[q == 0 ? 0 : 2] = r;
    }
    // dm15: first grid time at or after t_peak + 15
    {
        double r = NAN;
        const double tt = c.t(pk) + 15.0;
        int best = 0x7fffffff;
        for (int i = pk + 1 + threadIdx.x; i < n; i += GP_THREADS)
            if (c.t(i) >= tt && i < best) best = i;
        int i = block_min_int(best, ired);
        if (i != 0x7fffffff) {
            double t0 = c.t(i - 1), t1 = c.t(i);
            double dt = t1 - t0;
            if (fabs(dt) < 1e-15) r = c.m[i - 1] - pm;
            else {
                double fr = (tt - t0) / dt;
                double mi = c.m[i - 1] + fr * (c.m[i] - c.m[i - 1]);
                r = mi - pm;
            }
        }
        o[1] = r;
    }
    // FWHM on grid points (compute_fwhm)
    {
        const double half = pm + 0.75;
        int ib = last_at_least(c, 0, pk, half, ired);
        int ia = first_at_least(c, pk + 1, n, half, ired);
        o[3] = (ib >= 0 && ia != 0x7fffffff) ? c.t(ia) - c.t(ib) : NAN;
    }
    // rise time scale (extract_rise_timescale)
    {
        double r = NAN;
        if (pk > 0) {
            const int k = pk < 3 ? pk : 3;
            double base = 0.0;
            for (int i = 0; i < k; i++) base += c.m[i];
            base = base / (double)k;
            const double target = base + (pm - base) * (1.0 - ::exp(-1.0));
            double bv = INFINITY;
            int bi = 0x7fffffff;
            for (int i = threadIdx.x; i < pk; i += GP_THREADS) {
                double dd = fabs(c.m[i] - target);
                if (dd < bv) {  // first strict minimum, NaN never taken
                    bv = dd;
                    bi = i;
                }
            }
            int i = block_argmin(bv, bi, red, ired);
            if (i != 0x7fffffff && fabs(c.m[i] - target) < INFINITY) r = c.t(pk) - c.t(i);
        }
        o[4] = r;
    }
    // rise crossings (find_rise_crossing)
    for (int q = 0; q < 2; q++) {
        double r = NAN;
        if (pk > 0) {
            const double thr = pm + (q == 0 ? OFF_HALF : OFF_EFOLD);
            int i = last_at_least(c, 0, pk, thr, ired);
            if (i >= 0) {
                double dm = c.m[i] - c.m[i + 1];
                if (fabs(dm) < 1e-15) r = c.t(pk) - c.t(i);
                else {
                    double fr = (thr - c.m[i + 1]) / dm;
                    double tc = c.t(i + 1) + fr * (c.t(i) - c.t(i + 1));
                    r = c.t(pk) - tc;
                }
            }
        }
        o[5 + q] = r;
    }
    // rise / decay rates over the first / last quarter
    if (n < 2) {
        o[7] = o[8] = NAN;
    } else {
        int k = (int)::ceil((double)n * 0.25);
        int ke = k < 2 ? 2 : k;
        ke = ke > n - 1 ? n - 1 : ke;
        o[7] = slope_range(c, 0, ke, 1e-10, red);
        int kd = k < 2 ? 2 : k;
        kd = kd > n ? n : kd;
        o[8] = slope_range(c, n - kd, n, 1e-10, red);
    }
    // near-peak rates over 30 days
    {
        double r = NAN;
        if (pk > 0 && pk < n) {
            const double ts = c.t(pk) - 30.0;
            int best = 0x7fffffff;
            for (int i = threadIdx.x; i <= pk; i += GP_THREADS)
                if (c.t(i) >= ts && i < best) best = i;
            int a0 = block_min_int(best, ired);
            r = a0 == 0x7fffffff ? NAN : slope_range(c, a0, pk + 1, 1e-15, red);
        }
        o[9] = r;
    }
    {
        double r = NAN;
        if (pk < (n - 1 > 0 ? n - 1 : 0)) {
            const double te = c.t(pk) + 30.0;
            int best = -1;
            for (int i = pk + threadIdx.x; i < n; i += GP_THREADS)
                if (c.t(i) <= te && i > best) best = i;
            int b0 = block_max_int(best, ired);
            r = b0 < 0 ? NAN : slope_range(c, pk, b0 + 1, 1e-15, red);
        }
        o[10] = r;
    }
}

// Stable rank of a[i] among a[0..n) (ties by index).
__device__ __forceinline__ int stable_rank(const double* a, int n, int i) {
    const double v = a[i];
    int r = 0;
    for (int j = 0; j < n; j++) {
        double u = a[j];
        r += (u < v) || (u == v && j < i);
    }
    return r;
}

__device__ __forceinline__ double finite_or_nan(double v) { return __builtin_isfinite(v) ? v : NAN; }

// Inclusive prefix sum across the workgroup of one value per thread; returns
// this thread's inclusive sum and sets *total.
__device__ __forceinline__ double block_scan(double v, double* red, double* total) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        double u = __shfl_up(v, off, WAVE);
        if (lane >= off) v += u;
    }
    __syncthreads();
    if (lane == WAVE - 1) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    const int w = threadIdx.x / WAVE;
    double before = 0.0, tot = 0.0;
#pragma unroll
    for (int k = 0; k < GP_WAVES; k++) {
        if (k < w) before += red[k];
        tot += red[k];
    }
    *total = tot;
    return before + v;
}

template <int ROWS>
__device__ __forceinline__ void np_band(const NpArgs& a, const Lds& s, int band) {
    const int tid = threadIdx.x;
    const int lane = lane_id(), wave = tid / WAVE;
    double* red = s.misc + M_RED;
    double* sc = s.misc + M_SCAL;
    int* ired = s.imisc;
    const int64_t o0 = a.offsets[band];
    const int n = (int)(a.offsets[band + 1] - o0);
    const int ng = a.n_grid;
    const int sstride = 2 * a.state_cap + 4;
    double* feat = a.feat + (int64_t)band * NP_NFEAT;
    double* gpred = a.pred + (int64_t)band * ng;
    double* gstd = a.stdv + (int64_t)band * ng;
    double* gdpred = a.dpred + (int64_t)band * ng;
    double* st = a.state + (int64_t)band * sstride;
    const double* bt = a.t + o0;
    const double* bm = a.mag + o0;
    const double* be = a.err + o0;
    double* dx = a.dx + o0;  // compacted detections (times, mags, errors)
    double* dy = a.dy + o0;
    double* de = a.de + o0;
    double* buf = a.buf + o0;
    int* flag = s.imisc + 20;

    for (int k = tid; k < NP_NFEAT; k += GP_THREADS) feat[k] = k == 10 ? (double)n : NAN;
    // Each feature goes out as soon as it is known (thread 0; the values are
    // workgroup-uniform), so no register holds it through the later phases.
    auto put = [&](int k, double v) {
        if (tid == 0) feat[k] = finite_or_nan(v);
    };
    for (int k = tid; k < ng; k += GP_THREADS) gpred[k] = gstd[k] = gdpred[k] = NAN;
    for (int k = tid; k < sstride; k += GP_THREADS) st[k] = 0.0;
    if (tid == 0) a.status[band] = 0;
    const double t_lo = a.t_lo[band];
    const double dur = a.t_hi[band] - t_lo;
    if (n < 5 || !(dur > 0.0)) return;

    // 1. detections, in order
    int nd = 0;
    for (int c0 = 0; c0 < n; c0 += GP_THREADS) {
        const int i = c0 + tid;
        const bool dd = i < n && is_detection(bm[i], be[i]);
        const unsigned long long bal = __ballot(dd);
        const int below = __popcll(bal & ((1ull << lane) - 1ull));
        __syncthreads();
        if (lane == 0) ired[wave] = __popcll(bal);
        __syncthreads();
        int base = nd, tot = 0;
        for (int k = 0; k < GP_WAVES; k++) {
            if (k < wave) base += ired[k];
            tot += ired[k];
        }
        if (dd) {
            dx[base + below] = bt[i];
            dy[base + below] = bm[i];
            de[base + below] = be[i];
        }
        nd += tot;
    }
    __syncthreads();
    if (tid == 0) feat[44] = (double)(n - nd);
    if (nd < 5) return;

    // minimum length scale: 2 x median positive gap of the sorted detection times
    for (int i = tid; i < nd; i += GP_THREADS) {
        const double v = dx[i];
        int r = 0;
        for (int j = 0; j < nd; j++) {
            double u = dx[j];
            r += (u < v) || (u == v && j < i);
        }
        buf[r] = v;
    }
    __syncthreads();
    int n_gap = 0;
    for (int c0 = 0; c0 < nd - 1; c0 += GP_THREADS) {
        const int k = c0 + tid;
        bool ok = false;
        if (k < nd - 1) {
            double g = buf[k + 1] - buf[k];
            ok = __builtin_isfinite(g) && g > 0.0;
        }
        n_gap += block_count(ok, ired);
    }
    double med = 1.0;
    if (n_gap > 0) {
        const int target = n_gap / 2;
        for (int k = tid; k < nd - 1; k += GP_THREADS) {
            double g = buf[k + 1] - buf[k];
            if (!(__builtin_isfinite(g) && g > 0.0)) continue;
            int r = 0;
            for (int j = 0; j < nd - 1; j++) {
                double u = buf[j + 1] - buf[j];
                if (__builtin_isfinite(u) && u > 0.0) r += (u < g) || (u == g && j < k);
            }
            if (r == target) sc[0] = g;
        }
        __syncthreads();
        med = sc[0];
        __syncthreads();
    }
    const double min_ls = fmax(med * 2.0, 0.1);

    const bool sparse = nd > a.dense_max;
    if ((sparse && !a.fitc) || (!sparse && nd > a.cap)) {  // band routed to a launch too small for it
        if (tid == 0) a.status[band] = -1;
        return;
    }
    // training times of a dense band into LDS
    if (!sparse)
        for (int i = tid; i < nd; i += GP_THREADS) s.xs[i] = dx[i];
    __syncthreads();
    const Det dgl{dx, dx, dy, de};    // all detections, times from global memory
    const Det dls{s.xs, dx, dy, de};  // a dense band's detections, times staged in LDS
    double ybar = 0.0;  // sequential, as the reference
    for (int i = 0; i < nd; i++) ybar += dy[i];
    ybar = ybar / (double)nd;
    double tmin = INFINITY, tmax = -INFINITY;
    if (sparse) {
        double lo = INFINITY, hi = -INFINITY;
        for (int i = tid; i < nd; i += GP_THREADS) {
            lo = fmin(lo, dx[i]);
            hi = fmax(hi, dx[i]);
        }
        tmin = block_min(lo, red);
        tmax = block_max(hi, red);
    }
    const Fitc fw = fitc_carve(s);
    const int m_ind = nd < a.n_ind ? nd : a.n_ind;

    // 2. candidates
    double best = INFINITY;
    int best_c = -1;
    const int nc = a.n_amp * a.n_ls;
    for (int c = 0; c < nc; c++) {
        const double amp = a.amps[c / a.n_ls];
        const double ls = fmax(dur / a.lsf[c % a.n_ls], 0.1);
        if (ls < min_ls) continue;
        const double it = 0.5 / (ls * ls);
        double score;
        if (!sparse) {
            const bool ok = dense_factor(s.r0, dls, nd, amp, it, ybar, flag, s.misc + M_RINV);
            if (ok && tid < WAVE) dense_alpha<ROWS>(s.r0, nd, s.al);
            restage(s.r0, s.xs, dls, nd);
            if (!ok) continue;
            score = dense_train_rms(dls, nd, s.al, red);
        } else {
            if (!fitc_fit(fw, dgl, nd, m_ind, amp, it, ybar, tmin, tmax)) continue;
            score = fitc_nlml(fw, dgl, nd, m_ind, amp, it, ybar, red);
        }
        if (__builtin_isfinite(score) && score < best) {
            best = score;
            best_c = c;
        }
        __syncthreads();
    }
    if (best_c < 0) return;
    const double famp = a.amps[best_c / a.n_ls];
    const double fls = fmax(dur / a.lsf[best_c % a.n_ls], 0.1);
    const double fit = 0.5 / (fls * fls);
    put(45, famp);
    put(46, fls);

    // refit the winner; a FITC band first fits its subsampled dense GP.  The
    // dense weights go to the state row first: the factor is still needed
    // for the standard deviations, and the weight slot overlaps its tail.
    Pred<ROWS> P;
    P.sparse = sparse;
    P.amp = famp;
    P.it = fit;
    P.ybar = ybar;
    P.d = dls;
    if (!sparse) {
        dense_factor(s.r0, dls, nd, famp, fit, ybar, flag, s.misc + M_RINV);
        if (tid < WAVE) dense_alpha<ROWS>(s.r0, nd, st);
        restage(s.r0, s.xs, dls, nd);
        for (int i = tid; i < nd; i += GP_THREADS) st[a.state_cap + i] = dx[i];
        if (tid == 0) {
            st[2 * a.state_cap] = famp;
            st[2 * a.state_cap + 1] = fit;
            st[2 * a.state_cap + 2] = ybar;
            st[2 * a.state_cap + 3] = (double)nd;
        }
        __syncthreads();
        P.m = nd;
        P.z = fw.z;  // unused by a dense predictor (no null LDS pointers: see Pred)
        P.alpha = s.al;
        P.L = s.r0;
        P.Lb = fw.lb;
    } else {
        const int ms = nd < a.max_sub ? nd : a.max_sub;
        const DetSub dsub{s.xs, dx, dy, de, nd, nd > a.max_sub, (double)nd / (double)a.max_sub};
        for (int i = tid; i < ms; i += GP_THREADS) s.xs[i] = dx[dsub.k(i)];
        __syncthreads();
        double ys = 0.0;
        for (int i = 0; i < ms; i++) ys += dsub.yy(i);
        ys = ys / (double)ms;
        for (int c = 0; c < nc; c++) {  // first candidate that factorises, no length-scale floor
            const double amp = a.amps[c / a.n_ls];
            const double ls = fmax(dur / a.lsf[c % a.n_ls], 0.1);
            const double it = 0.5 / (ls * ls);
            const bool ok = dense_factor(s.r0, dsub, ms, amp, it, ys, flag, s.misc + M_RINV);
            if (ok && tid < WAVE) dense_alpha<ROWS>(s.r0, ms, st);
            restage(s.r0, s.xs, dsub, ms);
            if (!ok) continue;
            for (int i = tid; i < ms; i += GP_THREADS) st[a.state_cap + i] = dsub.x(i);
            if (tid == 0) {
                st[2 * a.state_cap] = amp;
                st[2 * a.state_cap + 1] = it;
                st[2 * a.state_cap + 2] = ys;
                st[2 * a.state_cap + 3] = (double)ms;
            }
            __syncthreads();
            break;
        }
        if (!fitc_fit(fw, dgl, nd, m_ind, famp, fit, ybar, tmin, tmax)) return;
        P.m = m_ind;
        P.z = fw.z;
        P.alpha = fw.az;
        P.L = fw.lmm;
        P.Lb = fw.lb;
    }
    if (tid == 0) a.status[band] = sparse ? 2 : 1;

    // 3. predictive std (needs the factors) on the grid and at (a 200-point
    // subsample of) the detection times
    const double gstep_den = (double)(ng - 1);
    constexpr int NB = ((ROWS * 64 < NP_CAP_MAX ? ROWS * 64 : NP_CAP_MAX) + 15) / 16;
    double ssum = 0.0, scnt = 0.0;
    if (!sparse && nd <= 16 * NB) {
        // the grid queries, then the detection subsample: 16 per wavefront
        const int ns = nd > 200 ? 200 : nd;
        const int nq = ng + ns;
        for (int k0 = 16 * wave; k0 < nq; k0 += 16 * GP_WAVES) {
            const int kq = k0 + (lane & 15);
            const int kc = kq < nq ? kq : nq - 1;
            double tq;
            if (kc < ng) {
                tq = t_lo + (double)kc * dur / gstep_den;
            } else {
                const int k = kc - ng;
                tq = dx[nd > 200 ? (int)(((double)k + 0.5) * ((double)nd / 200.0)) : k];
            }
            const double sd = P.template dense_std16<NB>(tq);
            if (lane < 16 && kq < nq) {
                if (kq < ng) {
                    gstd[kq] = sd;
                } else if (__builtin_isfinite(sd)) {
                    ssum += sd;
                    scnt += 1.0;
                }
            }
        }
        ssum = block_sum(ssum, red);
        scnt = block_sum(scnt, red);
    } else {  // FITC: the wave-level solves against K_mm and B
        for (int k0 = 4 * wave; k0 < ng; k0 += 4 * GP_WAVES) {
            double tq[4], sd[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + u < ng ? k0 + u : ng - 1;
                tq[u] = t_lo + (double)k * dur / gstep_den;
            }
            P.wave_std4(tq, sd);
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (lane == u && k0 + u < ng) gstd[k0 + u] = sd[u];
        }
        {
            const int ns = nd > 200 ? 200 : nd;
            for (int k0 = 4 * wave; k0 < ns; k0 += 4 * GP_WAVES) {
                double tq[4], sd[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u < ns ? k0 + u : ns - 1;
                    const int i = nd > 200 ? (int)(((double)k + 0.5) * ((double)nd / 200.0)) : k;
                    tq[u] = dx[i];
                }
                P.wave_std4(tq, sd);
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (k0 + u < ns && __builtin_isfinite(sd[u])) {
                        ssum += sd[u];
                        scnt += 1.0;
                    }
            }
            // every lane of a wave holds the same partial: count it once per wave
            ssum = block_sum(lane == 0 ? ssum : 0.0, red);
            scnt = block_sum(lane == 0 ? scnt : 0.0, red);
        }
    }
    if (!sparse) {  // dense weights into the slot, over the factor's tail
        for (int i = tid; i < nd; i += GP_THREADS) s.al[i] = st[i];
        __syncthreads();
    }
    // predictive mean and derivative on the grid, residuals at the detections
    for (int k = tid; k < ng; k += GP_THREADS) {
        const double tq = t_lo + (double)k * dur / gstep_den;
        gpred[k] = P.mean(tq);
        gdpred[k] = P.dmean(tq);
    }
    double rss = 0.0, chi = 0.0, bv = 0.0, omin = INFINITY, snr = -INFINITY;
    for (int i = tid; i < nd; i += GP_THREADS) {
        const double y = dy[i], e = de[i];
        const double r = y - P.mean(dx[i]);
        const double e2 = e * e + 1e-10;
        rss += r * r;
        chi += r * r / e2;
        bv += (y - ybar) * (y - ybar) / e2;
        omin = fmin(omin, y);
        snr = fmax(snr, fabs(y) / e);
    }
    rss = block_sum(rss, red);
    chi = block_sum(chi, red);
    bv = block_sum(bv, red);
    omin = block_min(omin, red);
    snr = block_max(snr, red);
    const double rms = ::sqrt(rss / (double)nd);
    put(8, chi / (double)nd);
    put(9, bv / (double)nd);
    put(25, snr);
    {
        double scale = 1.0;
        if (scnt > 0.0) {
            const double mstd = ssum / scnt;
            if (mstd > 1e-12 && __builtin_isfinite(rms)) scale = fmin(fmax(rms / mstd, 0.05), 5.0) * 0.6;
        }
        __syncthreads();  // grid mean and std written
        if (scale != 1.0)
            for (int k = tid; k < ng; k += GP_THREADS) gstd[k] *= scale;
        __syncthreads();
    }

    // 4. peak on the grid
    int pk;
    {
        double bvv = INFINITY;
        int bi = 0x7fffffff;
        for (int k = tid; k < ng; k += GP_THREADS) {
            double v = gpred[k];
            if (__builtin_isfinite(v) && v < bvv) {
                bvv = v;
                bi = k;
            }
        }
        pk = block_argmin(bvv, bi, red, ired);
        if (pk == 0x7fffffff) pk = 0;
    }
    const double t0 = t_lo + (double)pk * dur / gstep_den;
    const double peak_mag = fmax(gpred[pk], omin - 2.0);
    put(6, t0);
    put(7, peak_mag);
    double t_last = -INFINITY;
    for (int i = tid; i < n; i += GP_THREADS) t_last = fmax(t_last, bt[i]);
    t_last = block_max(t_last, red);
    put(21, t0 - t_last);

    // predictive queries: t_last - 2 .. t_last + 3, t0 + 30 / 60 / 90
    if (tid < 9) {
        const double tq = tid < 6 ? t_last + (double)(tid - 2) : t0 + 30.0 * (double)(tid - 5);
        s.misc[M_Q + tid] = P.mean(tq);
    }

    // fine grid (1 day, <= 2000 points) in LDS; the factors are no longer needed
    double fine[11];
    {
        const double fa = t_lo + 0.0 * dur / gstep_den;
        const double fb = t_lo + (double)(ng - 1) * dur / gstep_den;
        int nf = (int)fmax(fb - fa, 1.0) + 1;
        const double fstep = nf > NP_FINE_MAX ? (fb - fa) / 1999.0 : 1.0;
        nf = nf < NP_FINE_MAX ? nf : NP_FINE_MAX;
        __syncthreads();
        double* fp = s.r0;
        for (int i = tid; i < nf; i += GP_THREADS) fp[i] = P.mean(fa + (double)i * fstep);
        __syncthreads();
        Curve cv{fp, nf, fa, fstep};
        double bvv = INFINITY;
        int bi = 0x7fffffff;
        for (int i = tid; i < nf; i += GP_THREADS) {
            double v = fp[i];
            if (__builtin_isfinite(v) && v < bvv) {
                bvv = v;
                bi = i;
            }
        }
        int pkf = block_argmin(bvv, bi, red, ired);
        if (pkf == 0x7fffffff) pkf = 0;
        curve_metrics(cv, pkf, red, ired, fine);
        put(0, fine[4]);
        put(1, fine[5]);
        put(2, fine[6]);
        put(3, fine[0]);
        put(4, fine[1]);
        put(5, fine[2]);
        put(11, fine[3]);
        put(12, fine[7]);
        put(13, fine[8]);
        put(14, fine[9]);
        put(15, fine[10]);
    }

    // grid statistics (compute_predictive_features, fit_decay_power_law, ...)
    const double dg = (t_lo + 1.0 * dur / gstep_den) - (t_lo + 0.0 * dur / gstep_den);
    double sum = 0.0, mx = -INFINITY, mn = INFINITY, dmx = -INFINITY, dmn = INFINITY, sv = 0.0, svx = -INFINITY;
    for (int k = tid; k < ng; k += GP_THREADS) {
        const double p = gpred[k];
        sum += p;
        mx = fmax(mx, p);
        mn = fmin(mn, p);
        if (k + 1 < ng) {
            double d = (gpred[k + 1] - p) / dg;
            dmx = fmax(dmx, d);
            dmn = fmin(dmn, d);
        }
        const double sd = gstd[k];
        sv += sd * sd;
        svx = fmax(svx, sd * sd);
    }
    sum = block_sum(sum, red);
    mx = block_max(mx, red);
    mn = block_min(mn, red);
    dmx = block_max(dmx, red);
    dmn = block_min(dmn, red);
    sv = block_sum(sv, red);
    svx = block_max(svx, red);
    const double ngf = (double)ng;
    const double gmean = sum / ngf;
    double var = 0.0;
    for (int k = tid; k < ng; k += GP_THREADS) var += (gpred[k] - gmean) * (gpred[k] - gmean);
    var = block_sum(var, red);
    const double sigma_f = ::sqrt(var / ngf);
    put(23, sigma_f);
    put(24, mx - mn);
    put(26, dmx);
    put(27, dmn);
    put(28, gpred[ng - 1] / mn);
    put(29, sv / ngf);
    put(30, svx);
    double skew = NAN, kurt = NAN;
    if (ng > 3) {
        const double sdev = ::sqrt(var / ngf);
        if (sdev > 1e-10) {
            double s3 = 0.0, s4 = 0.0;
            for (int k = tid; k < ng; k += GP_THREADS) {
                double z = (gpred[k] - gmean) / sdev;
                double z2 = z * z;
                s3 += z2 * z;
                s4 += z2 * z2;
            }
            s3 = block_sum(s3, red);
            s4 = block_sum(s4, red);
            skew = s3 / ngf;
            kurt = s4 / ngf - 3.0;
        }
    }
    int n_infl = 0;
    {
        int part = 0;
        for (int k = 1 + tid; k + 2 < ng; k += GP_THREADS) {  // pair (d2[k], d2[k+1]), d2 at 1..ng-2
            double da = (gpred[k + 1] - 2.0 * gpred[k] + gpred[k - 1]) / (dg * dg);
            double db = (gpred[k + 2] - 2.0 * gpred[k + 1] + gpred[k]) / (dg * dg);
            if (__builtin_isfinite(da) && __builtin_isfinite(db) && fabs(da) > 1e-6 && fabs(db) > 1e-6 &&
                da * db < 0.0)
                part++;
        }
        n_infl = (int)block_sum((double)part, red);
    }
    put(31, skew);
    put(32, kurt);
    put(33, (double)n_infl);
    // power law over t - t_peak >= 5 after the grid peak
    double pl_b = NAN, pl_chi = NAN;
    {
        const double tpk = t0;
        double cw = 0, sw = 0, swx = 0, swy = 0, swxx = 0, swxy = 0;
        for (int k = pk + 1 + tid; k < ng; k += GP_THREADS) {
            const double tk = t_lo + (double)k * dur / gstep_den;
            const double dt = tk - tpk;
            const double p = gpred[k];
            if (dt < 5.0 || !__builtin_isfinite(p)) continue;
            const double sd = gstd[k];
            const double w = (__builtin_isfinite(sd) && sd > 1e-10) ? 1.0 / (sd * sd) : 1.0;
            const double x = ::log10(dt + 1.0);
            cw += 1.0;
            sw += w;
            swx += w * x;
            swy += w * p;
            swxx += w * x * x;
            swxy += w * x * p;
        }
        cw = block_sum(cw, red);
        sw = block_sum(sw, red);
        swx = block_sum(swx, red);
        swy = block_sum(swy, red);
        swxx = block_sum(swxx, red);
        swxy = block_sum(swxy, red);
        if (cw >= 3.0) {
            const double den = sw * swxx - swx * swx;
            if (!(fabs(den) < 1e-20)) {
                const double b = (sw * swxy - swx * swy) / den;
                const double a0 = (swy - b * swx) / sw;
                double chi2 = 0.0;
                for (int k = pk + 1 + tid; k < ng; k += GP_THREADS) {
                    const double tk = t_lo + (double)k * dur / gstep_den;
                    const double dt = tk - tpk;
                    const double p = gpred[k];
                    if (dt < 5.0 || !__builtin_isfinite(p)) continue;
                    const double sd = gstd[k];
                    const double w = (__builtin_isfinite(sd) && sd > 1e-10) ? 1.0 / (sd * sd) : 1.0;
                    const double r = p - (a0 + b * ::log10(dt + 1.0));
                    chi2 += w * r * r;
                }
                chi2 = block_sum(chi2, red);
                pl_b = b;
                pl_chi = chi2 / (double)((cw > 3.0 ? (int)cw : 3) - 2);
            }
        }
    }
    put(34, pl_b);
    put(35, pl_chi);
    // post-peak monotonicity and local maxima
    double mono = NAN;
    {
        double part = 0.0;
        for (int k = pk + tid; k + 1 < ng; k += GP_THREADS) {
            double u = gpred[k], v = gpred[k + 1];
            if (__builtin_isfinite(v) && __builtin_isfinite(u) && v > u) part += 1.0;
        }
        part = block_sum(part, red);
        const int len = ng - pk;
        if (pk < (ng - 2 > 0 ? ng - 2 : 0) && len >= 3) mono = part / (double)(len - 1);
    }
    int n_max;
    {
        double part = 0.0;
        for (int k = 1 + tid; k + 1 < ng; k += GP_THREADS) {
            double u = gpred[k - 1], v = gpred[k], w = gpred[k + 1];
            if (!__builtin_isfinite(v) || !__builtin_isfinite(u) || !__builtin_isfinite(w)) continue;
            if (v < u - 0.1 && v < w - 0.1) part += 1.0;
        }
        int c = (int)block_sum(part, red);
        if (ng >= 3) {
            if (__builtin_isfinite(gpred[0]) && __builtin_isfinite(gpred[1]) && gpred[0] < gpred[1] - 0.1) c++;
            if (__builtin_isfinite(gpred[ng - 1]) && __builtin_isfinite(gpred[ng - 2]) &&
                gpred[ng - 1] < gpred[ng - 2] - 0.1)
                c++;
            n_max = c > 1 ? c : 1;
        } else {
            n_max = ng == 0 ? 0 : 1;
        }
    }
    put(42, mono);
    put(43, (double)n_max);

    // variability features of the raw band (all observations)
    double vn = NAN, stk = NAN, exv = NAN, cus = NAN, pre_rms = NAN, pre_sig = NAN;
    {
        const double nf = (double)n;
        // time-sorted magnitudes (stable rank of the times)
        for (int i = tid; i < n; i += GP_THREADS) {
            const double v = bt[i];
            int r = 0;
            for (int j = 0; j < n; j++) {
                double u = bt[j];
                r += (u < v) || (u == v && j < i);
            }
            buf[r] = bm[i];
        }
        double ms = 0.0;
        for (int i = tid; i < n; i += GP_THREADS) ms += bm[i];
        ms = block_sum(ms, red);
        const double mean = ms / nf;
        double v2 = 0.0, d2 = 0.0, ex = 0.0;
        for (int i = tid; i < n; i += GP_THREADS) {
            const double dm = bm[i] - mean;
            v2 += dm * dm;
            ex += dm * dm - be[i] * be[i];
        }
        v2 = block_sum(v2, red);
        ex = block_sum(ex, red);
        __syncthreads();  // sorted magnitudes written
        for (int k = tid; k + 1 < n; k += GP_THREADS) {
            const double dd = buf[k + 1] - buf[k];
            d2 += dd * dd;
        }
        d2 = block_sum(d2, red);
        const double var1 = v2 / (nf - 1.0);
        vn = var1 < 1e-20 ? NAN : d2 / ((nf - 1.0) * var1);
        exv = fabs(mean) < 1e-20 ? NAN : ex / (nf * mean * mean);
        // Stetson K
        double ws = 0.0, wm = 0.0;
        for (int i = tid; i < n; i += GP_THREADS) {
            const double w = 1.0 / fmax(be[i] * be[i], 1e-30);
            ws += w;
            wm += bm[i] * w;
        }
        ws = block_sum(ws, red);
        wm = block_sum(wm, red);
        const double wmean = wm / ws;
        double sa = 0.0, sq = 0.0;
        for (int i = tid; i < n; i += GP_THREADS) {
            const double dl = (bm[i] - wmean) / fmax(be[i], 1e-15);
            sa += fabs(dl);
            sq += dl * dl;
        }
        sa = block_sum(sa, red);
        sq = block_sum(sq, red);
        const double srms = ::sqrt(sq / nf);
        stk = srms < 1e-20 ? NAN : (sa / nf) / srms;
        // CUSUM range over the input order
        const double sd = ::sqrt(v2 / (nf - 1.0));
        if (!(sd < 1e-20)) {
            const double norm = nf * sd;
            double carry = 0.0, lo = INFINITY, hi = -INFINITY;
            for (int c0 = 0; c0 < n; c0 += GP_THREADS) {
                const int i = c0 + tid;
                double tot;
                double incl = block_scan(i < n ? bm[i] - mean : 0.0, red, &tot);
                if (i < n) {
                    double x = (carry + incl) / norm;
                    lo = fmin(lo, x);
                    hi = fmax(hi, x);
                }
                carry += tot;
            }
            lo = block_min(lo, red);
            hi = block_max(hi, red);
            cus = hi - lo;
        }
        // pre-peak statistics (observations before the grid peak)
        double pc = 0.0, psum = 0.0;
        for (int i = tid; i < n; i += GP_THREADS)
            if (bt[i] < t0) {
                pc += 1.0;
                psum += bm[i];
            }
        pc = block_sum(pc, red);
        psum = block_sum(psum, red);
        if (pc >= 3.0) {
            const double pmean = psum / pc;
            double pv = 0.0;
            for (int i = tid; i < n; i += GP_THREADS)
                if (bt[i] < t0) pv += (bm[i] - pmean) * (bm[i] - pmean);
            pv = block_sum(pv, red);
            pre_rms = ::sqrt(pv / (pc - 1.0));
            // median error of the pre-peak observations: stable rank selection
            const int target = (int)pc / 2;
            for (int i = tid; i < n; i += GP_THREADS) {
                if (!(bt[i] < t0)) continue;
                const double v = be[i];
                int r = 0;
                for (int j = 0; j < n; j++) {
                    if (!(bt[j] < t0)) continue;
                    double u = be[j];
                    r += (u < v) || (u == v && j < i);
                }
                if (r == target) sc[1] = v;
            }
            __syncthreads();
            const double medv = sc[1];
            __syncthreads();
            pre_sig = medv > 1e-10 ? (pmean - peak_mag) / medv : NAN;
        }
    }

    put(39, vn);
    put(40, pre_rms);
    put(41, pre_sig);
    put(47, stk);
    put(48, exv);
    put(49, cus);
    __syncthreads();  // predictive queries written
    if (tid == 0) {
        const double* yw = s.misc + M_Q;
        double sl = 0.0;
        int nsl = 0;
        for (int i = 0; i < 5; i++) {
            double d = yw[i + 1] - yw[i];
            if (__builtin_isfinite(d)) {
                sl += d;
                nsl++;
            }
        }
        const double f_m1 = yw[1], f_0 = yw[2], f_p1 = yw[3], f_p2 = yw[4], f_p3 = yw[5];
        put(16, nsl > 0 ? sl / (double)nsl : NAN);
        put(17, (f_p1 - f_0) / 1.0);
        put(18, (f_p1 - 2.0 * f_0 + f_m1) / (1.0 * 1.0));
        put(19, f_p1);
        put(20, f_p2);
        put(22, (f_p3 - f_p2) / 1.0);
        put(36, yw[6]);
        put(37, yw[7]);
        put(38, yw[8]);
    }
}

}  // namespace

// Rows per lane of the wave-level solves (ROWS) and workgroups per CU (WGS)
// follow the class capacity: a 64-detection class keeps one row per lane, so
// its register arrays are a quarter of the 200-detection class's and four
// workgroups fit a CU; a 96-detection class fits three (LDS 38 KB each).
template <int ROWS, int WGS>
__global__ __launch_bounds__(GP_THREADS, WGS) void np_band_kernel(NpArgs a) {
    extern __shared__ double np_lds[];
    const Lds s = carve(np_lds, a.cap, a.fitc);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s.imisc[16] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int t = s.imisc[16];
        __syncthreads();
        if (t >= a.n_tasks) break;
        np_band<ROWS>(a, s, a.order[t]);
    }
}

size_t np_lds_bytes(int cap, int fitc) {
    return sizeof(double) * ((size_t)r0_doubles(cap, fitc) + (size_t)cap + MISC_DOUBLES);
}

static const void* np_kernel_for(int cap) {
    if (cap <= 64) return (const void*)np_band_kernel<1, 4>;
    if (cap <= 96) return (const void*)np_band_kernel<2, 3>;
    if (cap <= 128) return (const void*)np_band_kernel<2, 2>;
    return (const void*)np_band_kernel<4, 2>;
}

hipError_t launch_np(const NpArgs& a, int grid, hipStream_t st) {
    if (a.cap < 1 || a.cap > NP_CAP_MAX || a.n_ind < 1 || a.n_ind > NP_MAX_IND || a.n_grid < 2 || grid <= 0)
        return hipErrorInvalidValue;
    void* args[] = {const_cast<NpArgs*>(&a)};
    return hipLaunchKernel(np_kernel_for(a.cap), dim3(grid), dim3(GP_THREADS), args, np_lds_bytes(a.cap, a.fitc), st);
}

hipError_t np_occupancy(int cap, int fitc, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, np_kernel_for(cap), GP_THREADS,
                                                         np_lds_bytes(cap, fitc));
}

}  // namespace lcf
