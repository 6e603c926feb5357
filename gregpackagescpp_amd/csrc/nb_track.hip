// nb_track.hip -- body_tracker reductions over a subset of the bodies (gfx950).
//
// Reference: gtd::body_tracker (gregsys.hpp:3492-3611): total mass, centre of
// mass and COM velocity (the mass sum, then the ordered sums of m r and m v,
// divided by the mass unless it is zero), mean distance to a point, and mean
// pairwise separation (every pair outer < inner in the subset's order, divided
// by the integer k(k-1)/2).  Everything is FP64.  The O(k) sums run in the
// subset's order on one lane (the reference's association); the O(k^2)
// separation sums each outer body's row in parallel (inner ascending) and
// then adds the rows in order on one lane.
#include <algorithm>

#include "nb_system.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {

constexpr int kBlock = 256;

// out: mass, com[3], com_vel[3], mean distance to p (when has_point)
__global__ void k_track_serial(int64_t n, int64_t k, const int64_t *__restrict__ idx, const double *__restrict__ m,
                               const double *__restrict__ x, const double *__restrict__ v, double px, double py,
                               double pz, int has_point, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tm = 0;
    for (int64_t q = 0; q < k; ++q) tm += m[idx[q]];
    double mr[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
    for (int64_t q = 0; q < k; ++q) {
        const int64_t i = idx[q];
        for (int c = 0; c < 3; ++c) mr[c] += m[i] * x[c * n + i];
    }
    for (int64_t q = 0; q < k; ++q) {
        const int64_t i = idx[q];
        for (int c = 0; c < 3; ++c) mv[c] += m[i] * v[c * n + i];
    }
    out[0] = tm;
    for (int c = 0; c < 3; ++c) {
        out[1 + c] = (k == 0 || tm == 0) ? 0.0 : mr[c] / tm;
        out[4 + c] = (k == 0 || tm == 0) ? 0.0 : mv[c] / tm;
    }
    double sd = 0;
    if (has_point)
        for (int64_t q = 0; q < k; ++q) {
            const int64_t i = idx[q];
            const double dx = x[i] - px, dy = x[n + i] - py, dz = x[2 * n + i] - pz;
            sd += __builtin_sqrt(dx * dx + dy * dy + dz * dz);
        }
    out[7] = k ? sd / (double)(uint64_t)k : 0.0;
}

// row[a] = sum over b > a (ascending) of |r_idx[a] - r_idx[b]|
__global__ __launch_bounds__(kBlock) void k_track_rows(int64_t n, int64_t k, const int64_t *__restrict__ idx,
                                                       const double *__restrict__ x, double *__restrict__ row) {
    const int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (a >= k) return;
    const int64_t o = idx[a];
    const double xo = x[o], yo = x[n + o], zo = x[2 * n + o];
    double s = 0;
    for (int64_t b = a + 1; b < k; ++b) {
        const int64_t i = idx[b];
        const double dx = xo - x[i], dy = yo - x[n + i], dz = zo - x[2 * n + i];
        s += __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    }
    row[a] = s;
}

__global__ void k_track_sep(int64_t k, const double *__restrict__ row, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tot = 0;
    for (int64_t a = 0; a < k; ++a) tot += row[a];
    out[8] = k > 1 ? tot / (double)(((uint64_t)k * (uint64_t)(k - 1)) / 2) : 0.0;
}

}  // namespace

void track_bodies(nb_system *s, const int64_t *idx, int64_t k, const double *point, double *out9) {
    const int64_t n = s->n;
    for (int64_t q = 0; q < k; ++q)
        if (idx[q] < 0 || idx[q] >= n) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_track: body index out of range");
    hipStream_t st = s->stream;
    DBuf<int64_t> didx;
    DBuf<double> row, out;
    didx.ensure((size_t)std::max<int64_t>(k, 1));
    row.ensure((size_t)std::max<int64_t>(k, 1));
    out.ensure(9);
    if (k) NB_HIP(hipMemcpyAsync(didx.p, idx, (size_t)k * 8, hipMemcpyHostToDevice, st));
    const double px = point ? point[0] : 0.0, py = point ? point[1] : 0.0, pz = point ? point[2] : 0.0;
    hipLaunchKernelGGL(k_track_serial, dim3(1), dim3(64), 0, st, n, k, didx.p, s->m.p, s->x.p, s->v.p, px, py, pz,
                       point ? 1 : 0, out.p);
    if (k > 0)
        hipLaunchKernelGGL(k_track_rows, dim3((unsigned)((k + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, n, k,
                           didx.p, s->x.p, row.p);
    hipLaunchKernelGGL(k_track_sep, dim3(1), dim3(64), 0, st, k, row.p, out.p);
    NB_HIP(hipGetLastError());
    NB_HIP(hipMemcpyAsync(out9, out.p, 9 * 8, hipMemcpyDeviceToHost, st));
    NB_HIP(hipStreamSynchronize(st));
}

}  // namespace nbe
