// nb_split64.hpp -- device helpers shared by the FP64 split kernels (direct
// forces and the adjoint's pair pass).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace nbe {

constexpr int kWaves64 = 4;  // waves per workgroup in the FP64 split kernels

// FP64 reciprocal square root: v_rsq_f64 seed and two Newton steps
// y <- y (3 - s y^2)/2 (each step doubles the correct bits).
__device__ __forceinline__ double rsqrt64(double s) {
    double y = __builtin_amdgcn_rsq(s);
    double h = 0.5 * s;
    y = y * __builtin_fma(-h * y, y, 1.5);
    y = y * __builtin_fma(-h * y, y, 1.5);
    return y;
}

// Sum of C per-body components over nsplit partials [split][C][n] for the 64
// bodies of this workgroup: wave w adds splits w, w + kWaves64, ... in
// ascending order, then wave 0 adds the waves' sums in wave order.  Returns
// the totals in wave 0's lanes.
template <int C>
__device__ __forceinline__ void reduce_splits(const double *__restrict__ partial, int64_t n, int nsplit, int64_t i,
                                              double (&acc)[C], double (*red)[C][64]) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    if (i < n)
        for (int q = wave; q < nsplit; q += kWaves64) {
            const double *pp = partial + (size_t)q * C * n + i;
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] += pp[(size_t)c * n];
        }
    if (wave > 0)
#pragma unroll
        for (int c = 0; c < C; ++c) red[wave - 1][c][lane] = acc[c];
    __syncthreads();
    if (wave == 0)
#pragma unroll
        for (int w = 0; w < kWaves64 - 1; ++w)
#pragma unroll
            for (int c = 0; c < C; ++c) acc[c] += red[w][c][lane];
}

}  // namespace nbe
