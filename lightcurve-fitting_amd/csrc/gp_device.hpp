// gp_device.hpp -- device building blocks of the FP64 Gaussian-process
// kernels: packed-triangle Cholesky over a 256-thread workgroup, wave-level
// triangular solves with the right-hand side in registers, and workgroup
// reductions.
#pragma once
#include "lcf_device.hpp"

namespace lcf {

constexpr int GP_THREADS = 256;
constexpr int GP_WAVES = GP_THREADS / WAVE;
constexpr int GP_ROWS = 4;  // rows per lane in the wave-level solves (m <= 256)

// Packed lower triangle: element (i, j), j <= i.
__device__ __forceinline__ int tri(int i, int j) { return (i * (i + 1) >> 1) + j; }

// Row k's value from rows held as (lane + 64 r): wave-uniform k.
__device__ __forceinline__ double row_of(const double (&v)[GP_ROWS], int k) {
    const int r = k >> 6;
    double c = r == 0 ? v[0] : (r == 1 ? v[1] : (r == 2 ? v[2] : v[3]));
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

// One wavefront: b <- L^-1 b (rows lane + 64 r in registers); returns the
// squared norm of the solution (wave-uniform order j = 0..m-1).
__device__ __forceinline__ double wave_solve_l(const double* L, int m, double (&b)[GP_ROWS]) {
    const int lane = lane_id();
    double nrm = 0.0;
    for (int j = 0; j < m; j++) {
        double zj = row_of(b, j) / L[tri(j, j)];
        nrm += zj * zj;
#pragma unroll
        for (int r = 0; r < GP_ROWS; r++) {
            int i = lane + WAVE * r;
            if (i == j) b[r] = zj;
            else if (i > j && i < m) b[r] -= L[tri(i, j)] * zj;
        }
    }
    return nrm;
}

// One wavefront: b <- L^-T b.
__device__ __forceinline__ void wave_solve_lt(const double* L, int m, double (&b)[GP_ROWS]) {
    const int lane = lane_id();
    for (int j = m - 1; j >= 0; j--) {
        double aj = row_of(b, j) / L[tri(j, j)];
#pragma unroll
        for (int r = 0; r < GP_ROWS; r++) {
            int i = lane + WAVE * r;
            if (i == j) b[r] = aj;
            else if (i < j) b[r] -= L[tri(j, i)] * aj;
        }
    }
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
