// gp2d_hyper.hip -- 2D (time x log10 wavelength) GP hyper-parameter fit, FP64.
//
// One 256-thread workgroup per source (persistent, sources taken from a
// size-sorted queue; three workgroups per CU: the large phases are separate
// functions so that each fits 168 registers), the source's matrices in a
// global slot of ld x ld doubles (ld a multiple of 64, L2 / MALL resident),
// all O(n^3) work as 64 x 64 tile products staged through LDS and run on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64, a 32 x 32 quadrant per wave):
//   K   = a exp(-dt^2 / (2 lt^2) - dw^2 / (2 lw^2)) + diag(nv + sn^2),
//         theta = ln (a, lt, lw, sn), r = y - mean(y);
//   f   NLML = 1/2 z^T z + sum ln L_ii + n/2 ln 2 pi, z = L^-1 r, from a
//       right-looking blocked Cholesky of [[K, 0], [r^T, 1]] (padded with
//       the identity): the augmented row ends as z^T, so the forward solve
//       comes out of the factorisation;
//       The 64 x 64 diagonal blocks are factored left-looking in LDS by
//       one wave, the panels solved thread per row 16 columns at a time;
//   g   dNLML/dtheta_k = 1/2 tr((K^-1 - alpha alpha^T) dK/dtheta_k): X = L^-1
//       by blocked inversion (diagonal blocks by one wave, off-diagonal
//       blocks as X_IJ = -X_II sum_K L_IK X_KJ), alpha = X^T z, and
//       K^-1 = X^T X formed tile by tile and contracted at once with the
//       recomputed kernel derivatives, so K^-1 is never stored;
//   L-BFGS (m = 6) with a projected backtracking Armijo search (given up
//       once a step moves theta by <= 1e-10) and box bounds, exactly as the
//       oracle's driver; thread 0 holds the 4-vector
//       algebra, every thread takes part in the evaluations.
// Predictions at query points use alpha and X of the returned point.
#include "gp2d_hyper.hpp"
#include "gp_device.hpp"

namespace lcf {

namespace {

constexpr int T = GPH_TILE;
constexpr int NT = GPH_THREADS;
// block_sum (gp_device.hpp) combines GP_WAVES per-wave partial sums
static_assert(NT == GP_THREADS, "gp2d workgroups must match the block reductions' wave count");
constexpr int LM = 6;  // L-BFGS memory
constexpr double LN2PI = 1.83787706640934548356;

// box on theta = ln (a, lt, lw, sn): a in [1e-6, 1e6], lt in [0.1, 1e4] d,
// lw in [1e-3, 10] dex, sn in [1e-6, 10]
__constant__ double c_th_lo[4] = {-13.815510557964274, -2.302585092994046, -6.907755278982137,
                                  -13.815510557964274};
__constant__ double c_th_hi[4] = {13.815510557964274, 9.210340371976184, 2.302585092994046,
                                  2.302585092994046};

// L-BFGS state in LDS (indices into Smem::st)
enum {
    S_TH = 0, S_G = 4, S_F = 8, S_TT = 9, S_FT = 13, S_D = 14, S_STEP = 18, S_DEC = 19, S_YBAR = 20,
    S_GN = 24, S_S = 28, S_Y = 52, S_RHO = 76, S_HYP = 82, S_END = 96
};

// Second operand slab: a 32-deep half of a 64 x 64 tile operand (2048
// doubles) or one packed lower-triangular 64 x 64 block (2080 doubles).  At
// 52 KB of LDS in all, three workgroups share a CU.
constexpr int SB = T * (T + 1) / 2;
constexpr int KH = T / 2;  // depth of a half slab

struct Smem {
    double a[T * T];
    double b[SB];
    double red[NT];
    double st[S_END];
    int flag[8];  // 0 pivot failure, 1 stop, 2 accepted, 3 nm, 4 head, 5 step too small,
                  // 6 direction is the free steepest descent
};

struct Src {
    const double* t;
    const double* w;
    const double* y;
    const double* nv;
    int n;
    int np;  // padded size: rows [n + 1, np) are identity
};

// s[kk * 64 + r] = M[(r0 + r) * ld + k0 + kk] for kk < KK, zero for
// k0 + kk >= kend (rows of a row-major matrix become the slab's columns).
template <int KK = T>
__device__ __forceinline__ void stage_rows(double* s, const double* M, int ld, int r0, int k0, int kend) {
    const int rs_ = threadIdx.x >> 3, ks = threadIdx.x & 7;
#pragma unroll
    for (int q = 0; q < KK / 4; q++) {
        const int r = rs_ + 32 * (q & 1), kk = ks + 8 * (q >> 1);
        const int k = k0 + kk;
        s[kk * T + r] = k < kend ? M[(size_t)(r0 + r) * ld + k] : 0.0;
    }
}

// s[kk * 64 + c] = M[(k0 + kk) * ld + c0 + c] for kk < KK, zero for
// k0 + kk >= kend.
template <int KK = T>
__device__ __forceinline__ void stage_kmaj(double* s, const double* M, int ld, int c0, int k0, int kend) {
#pragma unroll
    for (int q = 0; q < KK / 4; q++) {
        const int idx = threadIdx.x + NT * q;
        const int kk = idx >> 6, c = idx & 63;
        const int k = k0 + kk;
        s[kk * T + c] = k < kend ? M[(size_t)k * ld + c0 + c] : 0.0;
    }
}

// 64 x 64 tile products on the FP64 matrix cores: C += A B with
// A[r][kk] = sa[kk * 64 + r] and B[kk][c] = sb[kk * 64 + c].  Wave w owns the
// 32 x 32 quadrant (32 (w >> 1), 32 (w & 1)) as 2 x 2 blocks of 16 x 16, each
// an accumulator of v_mfma_f64_16x16x4f64 (four doubles per lane), fed with
// one double of A and one of B per lane per four-deep k step.  Operand and
// result lane layouts of the instruction (checked on gfx950): A lane l ->
// A[l % 16][l / 16], B lane l -> B[l / 16][l % 16], D lane l element i ->
// D[4 i + l / 16][l % 16].
typedef double v4d __attribute__((ext_vector_type(4)));
struct Tile {
    v4d d[2][2];
};

__device__ __forceinline__ void zero(Tile& t) {
#pragma unroll
    for (int rb = 0; rb < 2; rb++)
#pragma unroll
        for (int cb = 0; cb < 2; cb++) t.d[rb][cb] = v4d{0.0, 0.0, 0.0, 0.0};
}

// Row / column inside the 64 x 64 tile of element i of block (rb, cb).
__device__ __forceinline__ int tile_row(int rb, int i) {
    return 32 * (threadIdx.x >> 7) + 16 * rb + 4 * i + ((threadIdx.x >> 4) & 3);
}
__device__ __forceinline__ int tile_col(int cb) {
    return 32 * ((threadIdx.x >> 6) & 1) + 16 * cb + (threadIdx.x & 15);
}

// KK-deep product (KK = 64: a whole tile; 32: one half of the k range, the
// halves accumulated in increasing k as one 64-deep product would).
template <int KK = T>
__device__ __forceinline__ void mma64(Tile& t, const double* sa, const double* sb) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double* pa = sa + (l >> 4) * T + 32 * (w >> 1) + (l & 15);
    const double* pb = sb + (l >> 4) * T + 32 * (w & 1) + (l & 15);
#pragma unroll 4
    for (int k0 = 0; k0 < KK; k0 += 4) {
        const double a0 = pa[k0 * T], a1 = pa[k0 * T + 16];
        const double b0 = pb[k0 * T], b1 = pb[k0 * T + 16];
        t.d[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, t.d[0][0], 0, 0, 0);
        t.d[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, t.d[0][1], 0, 0, 0);
        t.d[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, t.d[1][0], 0, 0, 0);
        t.d[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, t.d[1][1], 0, 0, 0);
    }
}

// K (lower triangle incl. padding and the augmented row) into A.
__device__ __noinline__ void build_k(double* A, int ld, const Src& s, const double* th, double ybar) {
    const double a = ::exp(th[0]), lt = ::exp(th[1]), lw = ::exp(th[2]), sn = ::exp(th[3]);
    const double it = 1.0 / (2.0 * lt * lt), iw = 1.0 / (2.0 * lw * lw), sn2 = sn * sn;
    const int n = s.n, np = s.np;
    // one element of the lower triangle per thread and step (rows of every
    // length keep all threads busy)
    const int ne = rs(np);
    for (int e = threadIdx.x; e < ne; e += NT) {
        int i, j;
        tri_ij(e, i, j);
        double k;
        if (i < n) {
            const double dt = s.t[i] - s.t[j], dw = s.w[i] - s.w[j];
            k = a * ::exp(-(dt * dt * it + dw * dw * iw));
            if (j == i) k += s.nv[i] + sn2;
        } else if (i == n) {
            k = j < n ? s.y[j] - ybar : 1.0;
        } else {
            k = j == i ? 1.0 : 0.0;
        }
        A[(size_t)i * ld + j] = k;
    }
    __syncthreads();
}

// Right-looking blocked Cholesky of A's lower triangle in place; columns
// >= n take the identity pivot.  Returns false (uniformly) on a pivot <= 0.
__device__ __noinline__ bool cholesky(double* A, int ld, const Src& s, Smem& sm) {
    const int n = s.n, np = s.np, tid = threadIdx.x, lane = lane_id();
    const int rows = n + 1;  // rows below n are identity and never change
    for (int J0 = 0; J0 < rows; J0 += T) {
        // (1) diagonal block on wave 0, left-looking in LDS: the block's lower
        // triangle column-major in sm.a (sm.a[c * T + r] = L[r][c]: a lane
        // reads its own row coalesced, the pivot row as a broadcast); lane r
        // computes column j of row r as A[r][j] - sum_k<j L[r][k] L[j][k],
        // the sum in four interleaved partial sums (a quarter of the
        // dependent chain), and scales it by the pivot's reciprocal square
        // root (v_rsq_f64 and one Newton step: a few ulp).  The reciprocals
        // go to sm.b[0 .. 63] for the panel.
        if (tid < WAVE) {
            const double* rowp = A + (size_t)(J0 + lane) * ld + J0;
            for (int c = 0; c < T; c++) sm.a[c * T + lane] = c <= lane ? rowp[c] : 0.0;
            int bad = 0;
            double rl = 1.0;
            for (int j = 0; j < T; j++) {
                if (J0 + j >= n) {
                    if (lane == j) sm.a[j * T + j] = 1.0;
                    continue;
                }
                double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
                int k = 0;
                for (; k + 4 <= j; k += 4) {
                    p0 = fma(sm.a[k * T + lane], sm.a[k * T + j], p0);
                    p1 = fma(sm.a[(k + 1) * T + lane], sm.a[(k + 1) * T + j], p1);
                    p2 = fma(sm.a[(k + 2) * T + lane], sm.a[(k + 2) * T + j], p2);
                    p3 = fma(sm.a[(k + 3) * T + lane], sm.a[(k + 3) * T + j], p3);
                }
                for (; k < j; k++) p0 = fma(sm.a[k * T + lane], sm.a[k * T + j], p0);
                const double acc = sm.a[j * T + lane] - ((p0 + p1) + (p2 + p3));
                const double piv = bcast(acc, j);
                bad |= !(piv > 0.0);
                double r = __builtin_amdgcn_rsq(piv);
                r = r * fma(-0.5 * piv * r, r, 1.5);
                if (lane == j) rl = r;
                if (lane >= j) sm.a[j * T + lane] = lane == j ? piv * r : acc * r;
            }
            double* wr = A + (size_t)(J0 + lane) * ld + J0;
            for (int c = 0; c <= lane; c++) wr[c] = sm.a[c * T + lane];
            sm.b[lane] = rl;  // 1 for the identity pivots of the padding
            if (tid == 0) sm.flag[0] = bad;
        }
        __syncthreads();
        if (sm.flag[0]) {
            __syncthreads();
            return false;
        }
        // (2) panel rows below the block (real rows and the augmented row),
        // thread per row, 16 columns at a time: the row's solved columns of
        // the earlier chunks are read back in increasing k, so every element
        // loses its terms in the same order as a full-row substitution.
        for (int i = J0 + T + tid; i < rows; i += NT) {
            double* rp = A + (size_t)i * ld + J0;
            for (int cb = 0; cb < T; cb += 16) {
                double v[16];
#pragma unroll
                for (int c = 0; c < 16; c++) v[c] = rp[cb + c];
                for (int k = 0; k < cb; k++) {
                    const double lik = rp[k];
#pragma unroll
                    for (int j = 0; j < 16; j++) v[j] = fma(-lik, sm.a[k * T + cb + j], v[j]);
                }
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    double acc = v[j];
#pragma unroll
                    for (int k = 0; k < j; k++) acc = fma(-v[k], sm.a[(cb + k) * T + cb + j], acc);
                    v[j] = acc * sm.b[cb + j];  // the pivot's reciprocal
                }
#pragma unroll
                for (int c = 0; c < 16; c++) rp[cb + c] = v[c];
            }
        }
        __syncthreads();
        // (3) trailing update of the lower tiles with rows < n + 1: the
        // panel rows of tile row I0 are staged once for the whole row of
        // tiles, and the diagonal tile reads them as both operands.
        for (int I0 = J0 + T; I0 < rows; I0 += T) {
            __syncthreads();
            stage_rows(sm.a, A, ld, I0, J0, J0 + T);
            for (int C0 = J0 + T; C0 <= I0; C0 += T) {
                Tile acc;
                zero(acc);
                if (C0 < I0) {  // the other tile row's panel in two 32-deep halves
                    for (int h = 0; h < T; h += KH) {
                        __syncthreads();
                        stage_rows<KH>(sm.b, A, ld, C0, J0 + h, J0 + T);
                        __syncthreads();
                        mma64<KH>(acc, sm.a + h * T, sm.b);
                    }
                } else {
                    __syncthreads();
                    mma64(acc, sm.a, sm.a);
                }
#pragma unroll
                for (int rb = 0; rb < 2; rb++)
#pragma unroll
                    for (int cb = 0; cb < 2; cb++)
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            A[(size_t)(I0 + tile_row(rb, i)) * ld + C0 + tile_col(cb)] -= acc.d[rb][cb][i];
            }
        }
        __syncthreads();
    }
    (void)np;
    return true;
}

// NLML from the factor; stores z (augmented row) in zv.
__device__ __noinline__ double nlml_from_factor(const double* A, int ld, const Src& s, double* zv, Smem& sm) {
    const int n = s.n;
    double q = 0.0, l = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const double z = A[(size_t)n * ld + i];
        zv[i] = z;
        q += z * z;
        l += ::log(A[(size_t)i * ld + i]);
    }
    q = block_sum(q, sm.red);
    l = block_sum(l, sm.red);
    return 0.5 * q + l + 0.5 * (double)n * LN2PI;
}

// Gradient at the factor in A (z in zv): X = L^-1 into Xm, alpha into av.
__device__ __noinline__ void gradient(double* A, double* Xm, int ld, const Src& s, const double* th, const double* zv,
                         double* av, Smem& sm, double* g) {
    const int n = s.n, tid = threadIdx.x, lane = lane_id();
    const int nb = (n + T - 1) / T;
    // the augmented row leaves the factor: [[L, 0], [0, 1]]
    for (int j = tid; j < n; j += NT) A[(size_t)n * ld + j] = 0.0;
    __syncthreads();
    // diagonal blocks of X = L^-1, two at a time (sm.a, sm.b, each the
    // block's lower triangle packed by rows: L[r][c] at r (r + 1) / 2 + c):
    // wave w of the pair's block, lane c solves L_II x = e_c
    for (int I = 0; I < nb; I += 2) {
        __syncthreads();
        for (int e = tid; e < T * T; e += NT) {
            const int r = e >> 6, c = e & 63;
            if (c > r) continue;
            const int pk = (r * (r + 1) >> 1) + c;
            sm.a[pk] = A[(size_t)(I * T + r) * ld + I * T + c];
            if (I + 1 < nb) sm.b[pk] = A[(size_t)((I + 1) * T + r) * ld + (I + 1) * T + c];
        }
        __syncthreads();
        const int w = tid >> 6;
        if (w < 2 && I + w < nb) {
            const double* L = w ? sm.b : sm.a;
            const int I0 = (I + w) * T;
            // 16 rows at a time; x of the earlier chunks read back from Xm
            // in increasing q (the order of a full-column substitution)
            double* xc = Xm + (size_t)I0 * ld + I0 + lane;
            for (int rb = 0; rb < T; rb += 16) {
                double x[16];
#pragma unroll
                for (int j = 0; j < 16; j++) x[j] = rb + j == lane ? 1.0 : 0.0;
                for (int q = 0; q < rb; q++) {
                    const double xq = xc[(size_t)q * ld];
#pragma unroll
                    for (int j = 0; j < 16; j++) x[j] = fma(-L[((rb + j) * (rb + j + 1) >> 1) + q], xq, x[j]);
                }
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const double* Lr = L + ((rb + j) * (rb + j + 1) >> 1);
                    double acc = x[j];
#pragma unroll
                    for (int q = 0; q < j; q++) acc = fma(-Lr[rb + q], x[q], acc);
                    x[j] = acc / Lr[rb + j];
                }
#pragma unroll
                for (int j = 0; j < 16; j++) xc[(size_t)(rb + j) * ld] = x[j];
            }
        }
    }
    __syncthreads();
    // off-diagonal blocks, column block by column block, rows in order
    for (int J = 0; J < nb; J++) {
        const int J0 = J * T;
        for (int I = J + 1; I < nb; I++) {
            const int I0 = I * T;
            Tile acc;
            zero(acc);
            for (int k0 = J0; k0 < I0; k0 += T) {
                __syncthreads();
                stage_rows(sm.a, A, ld, I0, k0, I0);
                for (int h = 0; h < T; h += KH) {
                    if (h) __syncthreads();
                    stage_kmaj<KH>(sm.b, Xm, ld, J0, k0 + h, I0);
                    __syncthreads();
                    mma64<KH>(acc, sm.a + h * T, sm.b);
                }
            }
            // X_IJ = -X_II S: S (the sum above) becomes the second operand in
            // sm.a, X_II's rows the first, in two 32-deep halves
            __syncthreads();
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) sm.a[tile_row(rb, i) * T + tile_col(cb)] = acc.d[rb][cb][i];
            Tile x2;
            zero(x2);
            for (int h = 0; h < T; h += KH) {
                if (h) __syncthreads();
                stage_rows<KH>(sm.b, Xm, ld, I0, I0 + h, I0 + T);
                __syncthreads();
                mma64<KH>(x2, sm.b, sm.a + h * T);
            }
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
                for (int cb = 0; cb < 2; cb++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        Xm[(size_t)(I0 + tile_row(rb, i)) * ld + J0 + tile_col(cb)] = -x2.d[rb][cb][i];
        }
    }
    __syncthreads();
    // alpha = X^T z
    for (int j = tid; j < n; j += NT) {
        double acc = 0.0;
        for (int i = j; i < n; i++) acc = fma(Xm[(size_t)i * ld + j], zv[i], acc);
        av[j] = acc;
    }
    __syncthreads();
    // K^-1 = X^T X tile by tile, contracted with the kernel derivatives
    const double a = ::exp(th[0]), lt = ::exp(th[1]), lw = ::exp(th[2]), sn = ::exp(th[3]);
    const double it = 1.0 / (2.0 * lt * lt), iw = 1.0 / (2.0 * lw * lw), sn2 = sn * sn;
    double ga = 0.0, gt = 0.0, gw = 0.0, gs = 0.0;
    for (int I = 0; I < nb; I++) {
        const int I0 = I * T;
        for (int J = 0; J <= I; J++) {
            const int J0 = J * T;
            Tile acc;
            zero(acc);
            for (int k0 = I0; k0 < n; k0 += T) {
                __syncthreads();
                stage_kmaj(sm.a, Xm, ld, I0, k0, n);
                if (J < I) {  // the second operand in two 32-deep halves
                    for (int h = 0; h < T; h += KH) {
                        if (h) __syncthreads();
                        stage_kmaj<KH>(sm.b, Xm, ld, J0, k0 + h, n);
                        __syncthreads();
                        mma64<KH>(acc, sm.a + h * T, sm.b);
                    }
                } else {  // a diagonal tile reads sm.a twice
                    __syncthreads();
                    mma64(acc, sm.a, sm.a);
                }
            }
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = I0 + tile_row(rb, r);
                if (i >= n) continue;
                const double ti = s.t[i], wi = s.w[i], ai = av[i];
#pragma unroll
                for (int cb = 0; cb < 2; cb++) {
                    const int j = J0 + tile_col(cb);
                    if (j > i) continue;
                    const double W = acc.d[rb][cb][r] - ai * av[j];
                    const double dt = ti - s.t[j], dw = wi - s.w[j];
                    const double dt2 = dt * dt, dw2 = dw * dw;
                    const double ak = a * ::exp(-(dt2 * it + dw2 * iw));
                    const double m = j == i ? W : 2.0 * W;
                    ga = fma(m, ak, ga);
                    gt = fma(m * ak, dt2 * (2.0 * it), gt);
                    gw = fma(m * ak, dw2 * (2.0 * iw), gw);
                    if (j == i) gs = fma(W, 2.0 * sn2, gs);
                }
            }
        }
    }
    ga = block_sum(ga, sm.red);
    gt = block_sum(gt, sm.red);
    gw = block_sum(gw, sm.red);
    gs = block_sum(gs, sm.red);
    g[0] = 0.5 * ga;
    g[1] = 0.5 * gt;
    g[2] = 0.5 * gw;
    g[3] = 0.5 * gs;
}

// NLML at th (all threads); false when K is not positive definite.
__device__ bool eval_f(double* A, int ld, const Src& s, const double* th, double ybar, double* zv, Smem& sm,
                       double& f) {
    build_k(A, ld, s, th, ybar);
    if (!cholesky(A, ld, s, sm)) return false;
    f = nlml_from_factor(A, ld, s, zv, sm);
    return __builtin_isfinite(f);
}

// Components of the gradient at th that press against their bound (the
// projected gradient is zero there; they stay fixed for this iteration).
__device__ __forceinline__ bool pinned_at(const double* th, const double* g, int k) {
    return (th[k] <= c_th_lo[k] && g[k] > 0.0) || (th[k] >= c_th_hi[k] && g[k] < 0.0);
}

// Thread 0: steepest descent on the free components, scaled to at most unit
// length; the curvature memory is dropped.
__device__ void steepest_free(Smem& sm) {
    double* st = sm.st;
    const double* th = st + S_TH;
    const double* g = st + S_G;
    double* d = st + S_D;
    double gn = 0.0;
    for (int k = 0; k < 4; k++)
        if (!pinned_at(th, g, k)) gn += g[k] * g[k];
    const double sc = 1.0 / fmax(1.0, ::sqrt(gn));
    for (int k = 0; k < 4; k++) d[k] = pinned_at(th, g, k) ? 0.0 : -g[k] * sc;
    sm.flag[3] = 0;
    sm.flag[6] = 1;
    double dmax = 0.0;
    for (int k = 0; k < 4; k++) dmax = fmax(dmax, fabs(d[k]));
    st[S_STEP] = dmax > 2.0 ? 2.0 / dmax : 1.0;
}

// Central finite differences of the NLML at th (all threads), h = 1e-5
// max(|theta_k|, 1); one-sided against f0 where one side is not positive
// definite, 0 where neither is.  Eight factorisations per gradient: the
// plain form the adjoint replaces.  Leaves A holding another point's factor.
__device__ __noinline__ void fd_gradient(double* A, int ld, const Src& s, const double* th, double ybar,
                                         double* zv, Smem& sm, double f0, double* g) {
#pragma unroll 1
    for (int k = 0; k < 4; k++) {
        const double h = 1e-5 * fmax(fabs(th[k]), 1.0);
        double tp[4], tm[4], fp = NAN, fm = NAN;
        for (int j = 0; j < 4; j++) tp[j] = tm[j] = th[j];
        tp[k] += h;
        tm[k] -= h;
        const bool okp = eval_f(A, ld, s, tp, ybar, zv, sm, fp);
        const bool okm = eval_f(A, ld, s, tm, ybar, zv, sm, fm);
        g[k] = okp && okm ? (fp - fm) / (2.0 * h) : (okp ? (fp - f0) / h : (okm ? (f0 - fm) / h : 0.0));
    }
}

// Thread 0: L-BFGS direction (two-loop recursion over the free components:
// the pinned components of g are zeroed before the recursion and of the
// direction after it) and first step length; returns false when the
// projected gradient is within gtol.
__device__ bool lbfgs_direction(Smem& sm, double gtol) {
    double* st = sm.st;
    const double* th = st + S_TH;
    const double* g = st + S_G;
    double gi = 0.0;
    bool pin[4];
    for (int k = 0; k < 4; k++) {
        pin[k] = pinned_at(th, g, k);
        gi = fmax(gi, pin[k] ? 0.0 : fabs(g[k]));
    }
    if (gi <= gtol) return false;
    const int nm = sm.flag[3], head = sm.flag[4];
    double q[4], al[LM];
    for (int k = 0; k < 4; k++) q[k] = pin[k] ? 0.0 : g[k];
    for (int m = 0; m < nm; m++) {
        const int p = (head - 1 - m + LM) % LM;
        double s = 0.0;
        for (int k = 0; k < 4; k++) s += st[S_S + 4 * p + k] * q[k];
        al[m] = st[S_RHO + p] * s;
        for (int k = 0; k < 4; k++) q[k] -= al[m] * st[S_Y + 4 * p + k];
    }
    double gam;
    if (nm > 0) {
        const int p = (head - 1 + LM) % LM;
        double sy = 0.0, yy = 0.0;
        for (int k = 0; k < 4; k++) {
            sy += st[S_S + 4 * p + k] * st[S_Y + 4 * p + k];
            yy += st[S_Y + 4 * p + k] * st[S_Y + 4 * p + k];
        }
        gam = sy / yy;
    } else {
        double gn = 0.0;
        for (int k = 0; k < 4; k++)
            if (!pin[k]) gn += g[k] * g[k];
        gam = 1.0 / fmax(1.0, ::sqrt(gn));
    }
    for (int k = 0; k < 4; k++) q[k] *= gam;
    for (int m = nm - 1; m >= 0; m--) {
        const int p = (head - 1 - m + LM) % LM;
        double s = 0.0;
        for (int k = 0; k < 4; k++) s += st[S_Y + 4 * p + k] * q[k];
        const double b = st[S_RHO + p] * s;
        for (int k = 0; k < 4; k++) q[k] += st[S_S + 4 * p + k] * (al[m] - b);
    }
    double* d = st + S_D;
    double gd = 0.0;
    for (int k = 0; k < 4; k++) {
        d[k] = pin[k] ? 0.0 : -q[k];
        gd += g[k] * d[k];
    }
    sm.flag[6] = 0;
    if (!(gd < 0.0)) {
        steepest_free(sm);
        return true;
    }
    double dmax = 0.0;
    for (int k = 0; k < 4; k++) dmax = fmax(dmax, fabs(d[k]));
    st[S_STEP] = dmax > 2.0 ? 2.0 / dmax : 1.0;
    return true;
}

// Thread 0: trial point th + step d, projected on the box, and the Armijo
// decrease term g^T (tt - th).  When the projection leaves an L-BFGS step
// with no predicted decrease or no movement, the search restarts along the
// free steepest descent direction (which cannot lose its decrease to the
// projection).
__device__ void trial_point(Smem& sm) {
    double* st = sm.st;
    for (int pass = 0; pass < 2; pass++) {
        double dec = 0.0, mv = 0.0;
        for (int k = 0; k < 4; k++) {
            const double v = fmin(fmax(st[S_TH + k] + st[S_STEP] * st[S_D + k], c_th_lo[k]), c_th_hi[k]);
            st[S_TT + k] = v;
            dec += st[S_G + k] * (v - st[S_TH + k]);
            mv = fmax(mv, fabs(v - st[S_TH + k]));
        }
        st[S_DEC] = dec;
        sm.flag[5] = mv <= 1e-10;  // the step no longer moves theta
        if (sm.flag[6] || (dec < 0.0 && !sm.flag[5])) break;
        steepest_free(sm);
    }
}

// Thread 0: store the pair (s, y) and move to the trial point.
__device__ void accept_step(Smem& sm) {
    double* st = sm.st;
    const int head = sm.flag[4];
    double sy = 0.0, sk[4], yk[4];
    for (int k = 0; k < 4; k++) {
        sk[k] = st[S_TT + k] - st[S_TH + k];
        yk[k] = st[S_GN + k] - st[S_G + k];
        sy += sk[k] * yk[k];
    }
    if (sy > 1e-12) {  // curvature pair kept only when s^T y > 0
        for (int k = 0; k < 4; k++) {
            st[S_S + 4 * head + k] = sk[k];
            st[S_Y + 4 * head + k] = yk[k];
        }
        st[S_RHO + head] = 1.0 / sy;
        sm.flag[4] = (head + 1) % LM;
        sm.flag[3] = sm.flag[3] + 1 < LM ? sm.flag[3] + 1 : LM;
    }
    for (int k = 0; k < 4; k++) {
        st[S_TH + k] = st[S_TT + k];
        st[S_G + k] = st[S_GN + k];
    }
    st[S_F] = st[S_FT];
}

__device__ __noinline__ void predict(const Gp2dHyperArgs& a, int src, double* Kq, const double* Xm, int ld, const Src& s,
                        const double* th, double ybar, const double* av, Smem& sm) {
    const int n = s.n, nq = a.n_query, tid = threadIdx.x;
    const double amp = ::exp(th[0]), lt = ::exp(th[1]), lw = ::exp(th[2]);
    const double it = 1.0 / (2.0 * lt * lt), iw = 1.0 / (2.0 * lw * lw);
    const double* qt = a.qt + (a.q_per_src ? (int64_t)src * nq : 0);
    const double* qw = a.qw + (a.q_per_src ? (int64_t)src * nq : 0);
    for (int q0 = 0; q0 < nq; q0 += T) {
        const int nqc = nq - q0 < T ? nq - q0 : T;
        __syncthreads();
        for (int e = tid; e < nqc * n; e += NT) {
            const int qq = e / n, i = e - qq * n;
            const double dt = qt[q0 + qq] - s.t[i], dw = qw[q0 + qq] - s.w[i];
            Kq[(size_t)qq * ld + i] = amp * ::exp(-(dt * dt * it + dw * dw * iw));
        }
        __syncthreads();
        const int qq = tid & 63, part = tid >> 6;
        double v2 = 0.0, mu = 0.0;
        if (qq < nqc) {
            const double* kr = Kq + (size_t)qq * ld;
            for (int r = part; r < n; r += 4) {
                const double* xr = Xm + (size_t)r * ld;
                double v = 0.0;
                for (int i = 0; i <= r; i++) v = fma(xr[i], kr[i], v);
                v2 = fma(v, v, v2);
            }
            if (part == 0)
                for (int i = 0; i < n; i++) mu = fma(kr[i], av[i], mu);
        }
        sm.red[tid] = v2;
        __syncthreads();
        if (part == 0 && qq < nqc) {
            const double tot = sm.red[qq] + sm.red[qq + 64] + sm.red[qq + 128] + sm.red[qq + 192];
            const int64_t o = (int64_t)src * nq + q0 + qq;
            a.pred[o] = mu + ybar;
            a.stdv[o] = ::sqrt(fmax(amp - tot, 1e-10));
        }
    }
    __syncthreads();
}

__device__ void fit_source(const Gp2dHyperArgs& a, int src, double* slot, Smem& sm) {
    const int tid = threadIdx.x;
    const int64_t o = a.off[src];
    Src s;
    s.t = a.t + o;
    s.w = a.w + o;
    s.y = a.y + o;
    s.nv = a.nv + o;
    s.n = (int)(a.off[src + 1] - o);
    s.np = ((s.n + 1 + T - 1) / T) * T;
    const int ld = a.ld;
    double* A = slot;
    double* Xm = slot + (size_t)ld * ld;
    double* zv = Xm + (size_t)ld * ld;
    double* av = zv + ld;
    double* st = sm.st;
    if (s.n == 0) {
        if (tid == 0) {
            for (int k = 0; k < 4; k++) {
                a.theta[src * 4 + k] = a.theta0[src * 4 + k];
                a.grad[src * 4 + k] = NAN;
            }
            a.nlml[src] = NAN;
            a.iters[src] = 0;
            a.status[src] = 3;
            if (a.evals) {
                a.evals[2 * src] = 0;
                a.evals[2 * src + 1] = 0;
            }
        }
        for (int q = tid; q < a.n_query && a.pred; q += NT) {
            a.pred[(int64_t)src * a.n_query + q] = NAN;
            a.stdv[(int64_t)src * a.n_query + q] = NAN;
        }
        return;
    }
    double ys = 0.0;
    for (int i = tid; i < s.n; i += NT) ys += s.y[i];
    ys = block_sum(ys, sm.red);
    const double ybar = ys / (double)s.n;
    const bool train = a.max_iters >= 0;
    if (tid == 0) {
        for (int k = 0; k < 4; k++) {
            const double v = a.theta0[src * 4 + k];
            st[S_TH + k] = train ? fmin(fmax(v, c_th_lo[k]), c_th_hi[k]) : v;
        }
        sm.flag[3] = 0;
        sm.flag[4] = 0;
    }
    __syncthreads();
    double th[4], g[4], f = NAN;
#pragma unroll
    for (int k = 0; k < 4; k++) th[k] = st[S_TH + k];
    int status = 1, it = 0, nf = 1, ng = 0;
    if (!eval_f(A, a.ld, s, th, ybar, zv, sm, f)) {
        status = 3;
    } else {
        if (a.fd_grad && train) {
            fd_gradient(A, ld, s, th, ybar, zv, sm, f, g);
            nf += 8;
        } else {
            gradient(A, Xm, ld, s, th, zv, av, sm, g);
        }
        ng++;
        if (tid == 0) {
            for (int k = 0; k < 4; k++) st[S_G + k] = g[k];
            st[S_F] = f;
        }
        __syncthreads();
        if (!train) {
            status = 0;
        } else {
            for (; it < a.max_iters; it++) {
                if (tid == 0) sm.flag[1] = lbfgs_direction(sm, a.gtol) ? 0 : 1;
                __syncthreads();
                if (sm.flag[1]) {
                    status = 0;
                    break;
                }
                bool ok = false;
                for (int ls = 0; ls < 30; ls++) {
                    if (tid == 0) trial_point(sm);
                    __syncthreads();
                    if (sm.flag[5]) break;
                    double tt[4], ft = NAN;
#pragma unroll
                    for (int k = 0; k < 4; k++) tt[k] = st[S_TT + k];
                    const bool pd = eval_f(A, ld, s, tt, ybar, zv, sm, ft);
                    nf++;
                    if (tid == 0) {
                        sm.flag[2] = pd && ft <= st[S_F] + 1e-4 * st[S_DEC];
                        st[S_FT] = ft;
                        if (!sm.flag[2]) st[S_STEP] *= 0.5;
                    }
                    __syncthreads();
                    if (sm.flag[2]) {
                        ok = true;
                        break;
                    }
                }
                if (!ok) {
                    status = 2;
                    break;
                }
                double tt[4], gn[4];
#pragma unroll
                for (int k = 0; k < 4; k++) tt[k] = st[S_TT + k];
                if (a.fd_grad) {
                    fd_gradient(A, ld, s, tt, ybar, zv, sm, st[S_FT], gn);
                    nf += 8;
                } else {
                    gradient(A, Xm, ld, s, tt, zv, av, sm, gn);
                }
                ng++;
                if (tid == 0) {
                    for (int k = 0; k < 4; k++) st[S_GN + k] = gn[k];
                    accept_step(sm);
                }
                __syncthreads();
            }
            // final state at the returned point
#pragma unroll
            for (int k = 0; k < 4; k++) th[k] = st[S_TH + k];
            eval_f(A, ld, s, th, ybar, zv, sm, f);
            gradient(A, Xm, ld, s, th, zv, av, sm, g);
            nf++;
            ng++;
        }
    }
    if (tid == 0) {
        for (int k = 0; k < 4; k++) {
            a.theta[src * 4 + k] = th[k];
            a.grad[src * 4 + k] = status == 3 ? NAN : g[k];
        }
        a.nlml[src] = status == 3 ? NAN : f;
        a.iters[src] = it;
        a.status[src] = status;
        if (a.evals) {
            a.evals[2 * src] = nf;
            a.evals[2 * src + 1] = ng;
        }
    }
    if (a.pred && a.n_query > 0) {
        if (status == 3) {
            for (int q = tid; q < a.n_query; q += NT) {
                a.pred[(int64_t)src * a.n_query + q] = NAN;
                a.stdv[(int64_t)src * a.n_query + q] = NAN;
            }
        } else {
            predict(a, src, A, Xm, ld, s, th, ybar, av, sm);
        }
    }
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(NT, 3) void gp2d_hyper_kernel(Gp2dHyperArgs a) {
    __shared__ Smem sm;
    __shared__ int ticket;
    double* slot = a.work + (size_t)blockIdx.x * gp2d_hyper_slot_doubles(a.ld);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) ticket = atomicAdd(a.queue, 1);
        __syncthreads();
        const int t = ticket;
        if (t >= a.n_tasks) break;
        fit_source(a, a.order[t], slot, sm);
    }
}

hipError_t launch_gp2d_hyper(const Gp2dHyperArgs& a, int grid, hipStream_t st) {
    if (grid <= 0 || a.ld % T != 0) return hipErrorInvalidValue;
    void* args[] = {const_cast<Gp2dHyperArgs*>(&a)};
    return hipLaunchKernel((const void*)gp2d_hyper_kernel, dim3(grid), dim3(NT), args, 0, st);
}

hipError_t gp2d_hyper_occupancy(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void*)gp2d_hyper_kernel, NT, 0);
}

}  // namespace lcf
