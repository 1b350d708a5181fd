// gp_device.hpp -- device building blocks of the FP64 Gaussian-process
// kernels: packed-triangle Cholesky over a 256-thread workgroup, wave-level
// triangular solves with the right-hand side in registers, and workgroup
// reductions.
#pragma once
#include "lcf_device.hpp"

namespace lcf {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GP_THREADS = 256;
constexpr int GP_WAVES = GP_THREADS / WAVE;
constexpr int GP_ROWS = 4;  // rows per lane in the wave-level solves (m <= 256); callers may use fewer

// Packed lower triangle: element (i, j), j <= i.
__device__ __forceinline__ int tri(int i, int j) { return (i * (i + 1) >> 1) + j; }

// Row k's value from rows held as (lane + 64 r), r < ROWS: wave-uniform k.
template <int ROWS>
__device__ __forceinline__ double row_of(const double (&v)[ROWS], int k) {
    const int r = k >> 6;
    double c = v[0];
#pragma unroll
    for (int q = 1; q < ROWS; q++) c = r == q ? v[q] : c;
    return bcast(c, k & (WAVE - 1));
}

// In-place Cholesky of the packed m x m matrix in L (all GP_THREADS threads;
// the matrix must be complete and visible, i.e. written before a barrier).
// Right-looking: column j is scaled by its pivot, then the trailing triangle
// loses l_ij l_kj, so every element is reduced in the same k order as a
// row-major Cholesky-Crout.  Returns false (uniformly) on a pivot <= 0,
// after a barrier, so the caller may overwrite L straight away.
__device__ __forceinline__ bool chol_packed(double* L, int m) {
    const int tid = threadIdx.x;
    for (int j = 0; j < m; j++) {
        double piv = L[tri(j, j)];
        if (piv <= 0.0) {  // uniform: every thread read the same pivot
            __syncthreads();
            return false;
        }
        double ljj = ::sqrt(piv);
        __syncthreads();  // all have read the pivot before it is overwritten
        if (tid == 0) L[tri(j, j)] = ljj;
        for (int i = j + 1 + tid; i < m; i += GP_THREADS) L[tri(i, j)] /= ljj;
        __syncthreads();
        const int r = m - j - 1;
        const int rr = r * r;
        for (int e = tid; e < rr; e += GP_THREADS) {
            int a = e / r, b = e - a * r;
            if (b <= a) {
                int i = j + 1 + a, k = j + 1 + b;
                L[tri(i, k)] -= L[tri(i, j)] * L[tri(k, j)];
            }
        }
        __syncthreads();
    }
    return true;
}

// Row start of row i in the packed lower triangle.
__device__ __forceinline__ int rs(int i) { return (i * (i + 1)) >> 1; }

// (i, j) of packed index e (j <= i), from a single-precision square root
// corrected to the exact integer row.
__device__ __forceinline__ void tri_ij(int e, int& i, int& j) {
    int r = (int)((__builtin_amdgcn_sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
    r += rs(r + 1) <= e;
    r -= rs(r) > e;
    i = r;
    j = e - rs(r);
}

// DPP row broadcast (row_newbcast): every lane gets lane N of its 16-lane
// row.  N folds to a constant in the callers' fully unrolled loops; unlike a
// readlane it leaves the value in a vector register (no scalar round trip).
template <int N>
__device__ __forceinline__ double row_bcast_c(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), 0x150 + N, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x150 + N, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double row_bcast(double v, int n) {
    switch (n) {
        case 0: return row_bcast_c<0>(v);   case 1: return row_bcast_c<1>(v);
        case 2: return row_bcast_c<2>(v);   case 3: return row_bcast_c<3>(v);
        case 4: return row_bcast_c<4>(v);   case 5: return row_bcast_c<5>(v);
        case 6: return row_bcast_c<6>(v);   case 7: return row_bcast_c<7>(v);
        case 8: return row_bcast_c<8>(v);   case 9: return row_bcast_c<9>(v);
        case 10: return row_bcast_c<10>(v); case 11: return row_bcast_c<11>(v);
        case 12: return row_bcast_c<12>(v); case 13: return row_bcast_c<13>(v);
        case 14: return row_bcast_c<14>(v); default: return row_bcast_c<15>(v);
    }
}

constexpr int CHOL_NB = 16;  // block-column width of the blocked Cholesky

// In-place Cholesky of the packed m x m matrix in L over the workgroup,
// left-looking by block columns of CHOL_NB: (A) the panel below and including
// the diagonal block loses sum_{k < J0} l_ik l_jk, as 16 x 16 tile products
// on the FP64 matrix cores; (B) wavefront 0 factors the diagonal block with
// its rows in registers (lane r holds row J0 + r; column j is scaled, then
// the columns to its right lose l_ij l_cj); (C) every row below solves
// against the diagonal block.  (A) and (B) share one phase: wavefront 0
// takes the diagonal tile's panel update and goes straight on to factor it,
// while the other three update the rest of the panel.  Within a block the
// elements are reduced term by term in increasing k (fused multiply-adds),
// as the row-major Cholesky-Crout of src/sparse_gp.rs:27-49 orders them; the
// panel sums of (A) round as tile products, and the column scalings multiply
// by the pivot's reciprocal: the factor agrees with the reference's to
// within its backward error, not bit for bit.
// rows = m + 1 appends row m (its first m entries, e.g. y - ybar): it is
// reduced like any row below the factor and ends as L^-1 times itself, the
// forward substitution of solve_l (src/sparse_gp.rs:52-58).
// Two barriers per block column; returns false (uniformly, after a
// barrier) on a pivot <= 0.  The matrix must be complete and visible.
// rinv (LDS, CHOL_NB doubles) carries the reciprocals of the block's pivots
// from (B) to (C), whose rows multiply by them instead of dividing (within an
// ulp per element of the reference's division; the division sat on each
// row's dependent chain).
__device__ __forceinline__ bool chol_blocked(double* L, int m, int rows, int* flag, double* rinv) {
    const int tid = threadIdx.x;
    const int lane = lane_id();
    for (int J0 = 0; J0 < m; J0 += CHOL_NB) {
        const int jb = m - J0 < CHOL_NB ? m - J0 : CHOL_NB;
        // (A) panel update, k < J0, on the FP64 matrix cores: wave 0 takes
        // the diagonal tile (rows J0 .. J0 + 15), waves 1-3 the 16-row tiles
        // 1, 2, ... of rows J0 .. rows - 1 in turn, against the block's 16
        // columns; D[r][c] = sum_k L[i_r][k] L[J0 + c][k] over
        // four-deep k steps (two accumulators, alternate steps), subtracted
        // once at the end (the reference subtracts term by term: the sum
        // rounds differently, within the factor's backward error).
        // v_mfma_f64_16x16x4f64 lanes: A lane l -> A[l % 16][l / 16],
        // B lane l -> B[l / 16][l % 16], D lane l element e -> D[4 e + l / 16][l % 16].
        if (J0 > 0) {
            const int ntile = (rows - J0 + 15) >> 4;
            const int kk = lane >> 4, c16 = lane & 15;
            const int jb_row = J0 + c16;
            const bool b_ok = c16 < jb;
            const int rb = rs(b_ok ? jb_row : J0) + kk;
            const int wave = tid >> 6;
            for (int T = wave; T < ntile; T += wave == 0 ? ntile : GP_WAVES - 1) {
                const int i_a = J0 + 16 * T + c16;
                const bool a_ok = i_a < rows;
                const int ra = rs(a_ok ? i_a : J0) + kk;
                v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
                // Rows past the matrix (a) and columns past the block (b)
                // read row J0 instead: they only reach product rows /
                // columns that are never written, so the loads need no
                // select and are issued one k step ahead.
                double a0 = L[ra], b0 = L[rb], a1 = L[ra + 4], b1 = L[rb + 4];
                for (int k0 = 0; k0 < J0; k0 += 8) {  // J0 is a multiple of CHOL_NB = 16
                    const int kn = k0 + 8 < J0 ? k0 + 8 : k0;
                    const double na0 = L[ra + kn], nb0 = L[rb + kn], na1 = L[ra + kn + 4], nb1 = L[rb + kn + 4];
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc1, 0, 0, 0);
                    a0 = na0; b0 = nb0; a1 = na1; b1 = nb1;
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int i = J0 + 16 * T + 4 * e + kk, j = J0 + c16;
                    if (i < rows && c16 < jb && j <= i) L[rs(i) + j] -= acc0[e] + acc1[e];
                }
            }
        }
        // (B) diagonal block on wavefront 0, straight after its own panel
        // tile (its LDS writes are in order; no barrier): lane r holds row J0 + r
        if (tid < WAVE) {
            double v[CHOL_NB];
            const bool own = lane < jb;
            const int rr = own ? rs(J0 + lane) + J0 : 0;
#pragma unroll
            for (int c = 0; c < CHOL_NB; c++) v[c] = (own && c <= lane) ? L[rr + c] : 0.0;
            int bad = 0;
            // Entries above a lane's diagonal (v[c], c > lane) and the rows of
            // lanes >= jb are never read or stored, so they are updated
            // unconditionally (no selects on the chain).
            double rl = 0.0;  // lane j: 1 / l_jj
#pragma unroll
            for (int j = 0; j < CHOL_NB; j++) {
                if (j < jb) {
                    const double piv = row_bcast(v[j], j);  // lanes 0..15 hold the block's rows
                    bad |= piv <= 0.0;
                    // r = 1 / sqrt(piv): v_rsq_f64 and one Newton step (a few
                    // ulp); l_jj = piv r and the column scaled by r
                    double r = __builtin_amdgcn_rsq(piv);
                    r = r * __builtin_fma(-0.5 * piv * r, r, 1.5);
                    if (lane == j) rl = r;
                    v[j] = lane == j ? piv * r : v[j] * r;
                    // columns c > j of the rows below c lose l_ij l_cj
#pragma unroll
                    for (int c = 1; c < CHOL_NB; c++)
                        if (c > j) v[c] = __builtin_fma(-v[j], row_bcast(v[j], c), v[c]);
                }
            }
            if (own) {
#pragma unroll
                for (int c = 0; c < CHOL_NB; c++)
                    if (c <= lane) L[rr + c] = v[c];
                rinv[lane] = rl;
            }
            if (tid == 0) *flag = bad;
        }
        __syncthreads();
        if (*flag) {
            __syncthreads();
            return false;
        }
        // (C) rows below the diagonal block: forward solve against it
        for (int i = J0 + jb + tid; i < rows; i += GP_THREADS) {
            const int ri0 = rs(i) + J0;
            double v[CHOL_NB];
#pragma unroll
            for (int c = 0; c < CHOL_NB; c++) v[c] = c < jb ? L[ri0 + c] : 0.0;
#pragma unroll
            for (int j = 0; j < CHOL_NB; j++) {
                if (j < jb) {
                    const int rj0 = rs(J0 + j) + J0;
                    double acc = v[j];
#pragma unroll
                    for (int k = 0; k < j; k++) acc = __builtin_fma(-v[k], L[rj0 + k], acc);
                    v[j] = acc * rinv[j];
                }
            }
#pragma unroll
            for (int c = 0; c < CHOL_NB; c++)
                if (c < jb) L[ri0 + c] = v[c];
        }
        __syncthreads();
    }
    return true;
}

// The pivots' reciprocals are formed up front, one per lane and row, so no
// division sits on the dependent chain of the substitution (each step
// multiplies: within an ulp of the reference's division).
template <int ROWS>
__device__ __forceinline__ void wave_rdiag(const double* L, int m, double (&rd)[ROWS]) {
    const int lane = lane_id();
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
        const int i = lane + WAVE * r;
        rd[r] = i < m ? 1.0 / L[tri(i, i)] : 0.0;
    }
}

// The substitutions below run j in four segments of 64 (register R = j / 64
// holds row j), so the row's register is chosen at compile time and the
// registers of rows that are already final are left alone.

// b <- L^-1 b over rows [64 R, min(64 R + 64, m)); adds z_j^2 to nrm.
template <int R, int ROWS>
__device__ __forceinline__ void solve_l_seg(const double* L, int m, double (&b)[ROWS], const double (&rd)[ROWS],
                                            const int (&rsi)[ROWS], double& nrm) {
    const int lane = lane_id();
    const int jhi = m < 64 * R + 64 ? m : 64 * R + 64;
    for (int j = 64 * R; j < jhi; j++) {
        const int src = j & 63;
        const double zj = bcast(b[R], src) * bcast(rd[R], src);
        nrm += zj * zj;
        if (lane == src) b[R] = zj;
        else if (lane > src && lane + 64 * R < m) b[R] -= L[rsi[R] + j] * zj;
#pragma unroll
        for (int r = R + 1; r < ROWS; r++)
            if (lane + 64 * r < m) b[r] -= L[rsi[r] + j] * zj;
    }
}

// One wavefront: b <- L^-1 b (rows lane + 64 r in registers, m <= 64 ROWS);
// returns the squared norm of the solution (wave-uniform order j = 0..m-1).
template <int ROWS>
__device__ __forceinline__ double wave_solve_l(const double* L, int m, double (&b)[ROWS]) {
    const int lane = lane_id();
    double rd[ROWS];
    int rsi[ROWS];
    wave_rdiag(L, m, rd);
#pragma unroll
    for (int r = 0; r < ROWS; r++) rsi[r] = rs(lane + 64 * r);
    double nrm = 0.0;
    solve_l_seg<0>(L, m, b, rd, rsi, nrm);
    if constexpr (ROWS > 1) if (m > 64) solve_l_seg<1>(L, m, b, rd, rsi, nrm);
    if constexpr (ROWS > 2) if (m > 128) solve_l_seg<2>(L, m, b, rd, rsi, nrm);
    if constexpr (ROWS > 3) if (m > 192) solve_l_seg<3>(L, m, b, rd, rsi, nrm);
    return nrm;
}

// b <- L^-T b over rows j in [64 R, min(64 R + 64, m)), descending.  Row
// j's entries are loaded one step ahead, so the LDS latency stays off the
// dependent chain (lane reads, multiply, subtract).  The loads of register R
// run past row j's end in the lanes that do not use them (the next rows, or
// the training times that follow the packed factor: inside the allocation).
template <int R, int ROWS>
__device__ __forceinline__ void solve_lt_seg(const double* L, int m, double (&b)[ROWS], const double (&rd)[ROWS]) {
    const int lane = lane_id();
    const int jhi = m < 64 * R + 64 ? m : 64 * R + 64;
    if (jhi <= 64 * R) return;
    double lc[R + 1];
    {
        const int rj = rs(jhi - 1);
#pragma unroll
        for (int r = 0; r <= R; r++) lc[r] = L[rj + lane + 64 * r];
    }
    for (int j = jhi - 1; j >= 64 * R; j--) {
        double ln[R + 1];
        const int rn = rs(j > 64 * R ? j - 1 : j);
#pragma unroll
        for (int r = 0; r <= R; r++) ln[r] = L[rn + lane + 64 * r];
        const int src = j & 63;
        const double aj = bcast(b[R], src) * bcast(rd[R], src);
        const double upd = b[R] - lc[R] * aj;  // selects, not branches, on the chain
        b[R] = lane == src ? aj : (lane < src ? upd : b[R]);
#pragma unroll
        for (int r = 0; r < R; r++) b[r] -= lc[r] * aj;
#pragma unroll
        for (int r = 0; r <= R; r++) lc[r] = ln[r];
    }
}

// One wavefront: b <- L^-T b (m <= 64 ROWS).
template <int ROWS>
__device__ __forceinline__ void wave_solve_lt(const double* L, int m, double (&b)[ROWS]) {
    double rd[ROWS];
    wave_rdiag(L, m, rd);
    if constexpr (ROWS > 3) if (m > 192) solve_lt_seg<3>(L, m, b, rd);
    if constexpr (ROWS > 2) if (m > 128) solve_lt_seg<2>(L, m, b, rd);
    if constexpr (ROWS > 1) if (m > 64) solve_lt_seg<1>(L, m, b, rd);
    solve_lt_seg<0>(L, m, b, rd);
}

// Workgroup reductions through red[GP_WAVES] in LDS; every thread gets the
// same value (wave partials combined in wave order).  Callers keep red free
// for the duration of the call.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < GP_WAVES; w++) s += red[w];
    return s;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, WAVE));
    return v;
}

// NaN-ignoring maximum / minimum (fmax / fmin drop a NaN operand).
__device__ __forceinline__ double block_max(double v, double* red) {
    v = wave_max(v);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < GP_WAVES; w++) s = fmax(s, red[w]);
    return s;
}

__device__ __forceinline__ double block_min(double v, double* red) {
    v = wave_min(v);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int w = 1; w < GP_WAVES; w++) s = fmin(s, red[w]);
    return s;
}

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, WAVE));
    return v;
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, WAVE));
    return v;
}

__device__ __forceinline__ int block_min_int(int v, int* red) {
    v = wave_min_int(v);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    int s = red[0];
#pragma unroll
    for (int w = 1; w < GP_WAVES; w++) s = min(s, red[w]);
    return s;
}

__device__ __forceinline__ int block_max_int(int v, int* red) {
    v = wave_max_int(v);
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    int s = red[0];
#pragma unroll
    for (int w = 1; w < GP_WAVES; w++) s = max(s, red[w]);
    return s;
}

__device__ __forceinline__ int block_count(bool p, int* red) {
    int c = __popcll(__ballot(p));
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x / WAVE] = c;
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int w = 0; w < GP_WAVES; w++) s += red[w];
    return s;
}

// Index of the smallest value (the smallest index among equal values);
// values must not be NaN (map them to +inf).  Every thread gets the result.
// The lexicographic (value, index) minimum is associative and commutative,
// so the butterfly leaves the same pair in every lane.
__device__ __forceinline__ int block_argmin(double v, int idx, double* redv, int* redi) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(v, off, WAVE);
        int oi = __shfl_xor(idx, off, WAVE);
        bool take = (ov < v) || (ov == v && oi < idx);
        v = take ? ov : v;
        idx = take ? oi : idx;
    }
    __syncthreads();
    if (lane_id() == 0) {
        redv[threadIdx.x / WAVE] = v;
        redi[threadIdx.x / WAVE] = idx;
    }
    __syncthreads();
    double bv = redv[0];
    int bi = redi[0];
#pragma unroll
    for (int w = 1; w < GP_WAVES; w++) {
        double ov = redv[w];
        int oi = redi[w];
        if (ov < bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    return bi;
}

}  // namespace lcf
