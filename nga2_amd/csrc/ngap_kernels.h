// Device kernels of the pressure Poisson solver (gfx950).
//
// Storage of one multigrid level: cells numbered p = i + n0*(j + n1*k)
// (x fastest, the Fortran order of NGA2's arrays). The symmetric 7-point
// operator is held in four arrays: diagonal c and the coupling a_d(p)
// between p and its -d neighbour; the +d coupling of p is a_d(p + e_d).
// Periodic directions wrap; in a non-periodic direction the coupling across
// the boundary is 0 and the neighbour index is not formed.
#pragma once
#include <hip/hip_runtime.h>

namespace ngap {

constexpr int NT = 256;        // threads per block of streaming kernels
constexpr int MAX_PARTS = 8192; // partial sums per reduction (fixed order)

struct Geo {
    int n0, n1, n2;
    int p0, p1, p2;
    long long N;
};

// a * b + c rounded once, in either precision
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ int nbm(int i, int n, int per) { return i > 0 ? i - 1 : (per ? n - 1 : -1); }
__device__ __forceinline__ int nbp(int i, int n, int per) { return i < n - 1 ? i + 1 : (per ? 0 : -1); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

// Sum over a block of NT threads in a fixed order; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sm) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int q = 0; q < NT / 64; q++) s += sm[q];
    }
    __syncthreads();
    return s;
}

__device__ __forceinline__ double block_max(double v, double *sm) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int q = 0; q < NT / 64; q++) s = fmax(s, sm[q]);
    }
    __syncthreads();
    return s;
}

}  // namespace ngap
