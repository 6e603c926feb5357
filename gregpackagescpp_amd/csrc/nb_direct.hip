// nb_direct.hip -- direct-summation force and leapfrog-KDK update kernels (gfx950).
//
// Reference path: gtd::system::cumulative_acc (gregsys.hpp:352-357), loop_kdk
// (:477-513), evolve's KDK branch (:2946-2968) and the softened pair form of
// bh_acc (:371-375).
//
// Two force kernels:
//  * k_force_mixed: FP32 pair forces.  Every lane owns IPL i-bodies; the
//    j-body stream (x, y, z, G*m as float4) is wave-uniform, so it is read with
//    scalar loads into SGPRs and each VALU op takes its j operand straight from
//    an SGPR: no LDS, no barriers.  Per interaction: 3 sub, 3 fma (r^2+eps^2),
//    1 v_rsq_f32, 3 mul, 3 fma = 13 VALU slots.  Sums over a tile of j run in
//    plain FP32 and are folded into a Kahan-compensated (sum, comp) pair per
//    tile.  The j range is split over gridDim.y so that N = 64K fills 256 CUs;
//    the split partials are reduced in FP64 by k_update_mixed.
//  * k_force_fp64: FP64, one lane per i, j ascending, FP contraction off, the
//    exact operation order of the reference's pairwise update (so results are
//    bit-identical to an IEEE-double restatement of it).  For small n the
//    same sums run in two phases (k_pair_terms, k_ordered_sum) so that the
//    pair terms use the whole GPU.
#include <algorithm>

#include "nb_internal.hpp"
#include "nb_split64.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 64;  // j-bodies summed in plain FP32 before each Kahan fold
constexpr int kTermJ = 8;  // j per lane in k_pair_terms

__global__ void k_pack(int64_t n, const double *__restrict__ x, const double *__restrict__ m, double G,
                       float4 *__restrict__ pos32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos32[i] = make_float4((float)x[i], (float)x[n + i], (float)x[2 * n + i], (float)(G * m[i]));
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int IPL, bool SOFT>
__global__ __launch_bounds__(kBlock) void k_force_mixed(const float4 *__restrict__ pos, int n, int jchunk, float eps2,
                                                        float *__restrict__ partial) {
#pragma clang fp contract(fast)
    const int split = blockIdx.y;
    const int j0 = split * jchunk;
    const int j1 = min(n, j0 + jchunk);
    const int ibase = blockIdx.x * (kBlock * IPL) + threadIdx.x;
    float xi[IPL], yi[IPL], zi[IPL];
    float sx[IPL], sy[IPL], sz[IPL], cx[IPL], cy[IPL], cz[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        int i = ibase + k * kBlock;
        float4 p = i < n ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
        sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
    }
    for (int jt = j0; jt < j1; jt += kTile) {
        const int je = min(j1, jt + kTile);
        float px[IPL], py[IPL], pz[IPL];
#pragma unroll
        for (int k = 0; k < IPL; ++k) px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 8
        for (int j = jt; j < je; ++j) {
            const float4 pj = pos[j];
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                float dx = pj.x - xi[k];
                float dy = pj.y - yi[k];
                float dz = pj.z - zi[k];
                float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                if (!SOFT) inv = r2 > 0.f ? inv : 0.f;  // self pair at eps = 0
                float w = pj.w * inv;
                w = w * (inv * inv);
                px[k] = __builtin_fmaf(w, dx, px[k]);
                py[k] = __builtin_fmaf(w, dy, py[k]);
                pz[k] = __builtin_fmaf(w, dz, pz[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < IPL; ++k) {
            kahan_add(sx[k], cx[k], px[k]);
            kahan_add(sy[k], cy[k], py[k]);
            kahan_add(sz[k], cz[k], pz[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        int i = ibase + k * kBlock;
        if (i < n) {
            float *pp = partial + (size_t)split * 6 * n + i;
            pp[0] = sx[k];
            pp[(size_t)n] = sy[k];
            pp[2 * (size_t)n] = sz[k];
            pp[3 * (size_t)n] = cx[k];
            pp[4 * (size_t)n] = cy[k];
            pp[5 * (size_t)n] = cz[k];
        }
    }
}

// The same pair sum with the j stream staged through LDS: each workgroup
// copies kBlock j-bodies at a time into LDS (one coalesced float4 per lane)
// and every wave reads them back with broadcast ds_read_b128, so all operands
// of the 13 pair ops are VGPRs.  On gfx950 FP32 VALU ops whose operands are
// all VGPRs issue at up to twice the rate of ops that take an SGPR operand,
// which the scalar-load form above pays on 5 of its 13 ops (3 sub, the
// eps^2 fma, the mass multiply).  The tile and Kahan-fold boundaries are those
// of k_force_mixed, so both kernels give identical sums.
template <int IPL, bool SOFT>
__global__ __launch_bounds__(kBlock) void k_force_mixed_lds(const float4 *__restrict__ pos, int n, int jchunk,
                                                            float eps2, float *__restrict__ partial) {
#pragma clang fp contract(fast)
    __shared__ float4 tile[kBlock];
    const int split = blockIdx.y;
    const int j0 = split * jchunk;
    const int j1 = min(n, j0 + jchunk);
    const int ibase = blockIdx.x * (kBlock * IPL) + threadIdx.x;
    float e2 = eps2;
    asm volatile("" : "+v"(e2));  // keep eps^2 in a VGPR
    float xi[IPL], yi[IPL], zi[IPL];
    float sx[IPL], sy[IPL], sz[IPL], cx[IPL], cy[IPL], cz[IPL];
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        int i = ibase + k * kBlock;
        float4 p = i < n ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
        sx[k] = sy[k] = sz[k] = cx[k] = cy[k] = cz[k] = 0.f;
    }
    for (int jb = j0; jb < j1; jb += kBlock) {
        const int cnt = min(kBlock, j1 - jb);
        __syncthreads();  // the previous tile has been read by every wave
        if ((int)threadIdx.x < cnt) tile[threadIdx.x] = pos[jb + threadIdx.x];
        __syncthreads();
        for (int t = 0; t < cnt; t += kTile) {
            const int te = min(cnt, t + kTile);
            float px[IPL], py[IPL], pz[IPL];
#pragma unroll
            for (int k = 0; k < IPL; ++k) px[k] = py[k] = pz[k] = 0.f;
#pragma unroll 8
            for (int j = t; j < te; ++j) {
                const float4 pj = tile[j];
#pragma unroll
                for (int k = 0; k < IPL; ++k) {
                    float dx = pj.x - xi[k];
                    float dy = pj.y - yi[k];
                    float dz = pj.z - zi[k];
                    float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, e2)));
                    float inv = __builtin_amdgcn_rsqf(r2);
                    if (!SOFT) inv = r2 > 0.f ? inv : 0.f;  // self pair at eps = 0
                    float w = pj.w * inv;
                    w = w * (inv * inv);
                    px[k] = __builtin_fmaf(w, dx, px[k]);
                    py[k] = __builtin_fmaf(w, dy, py[k]);
                    pz[k] = __builtin_fmaf(w, dz, pz[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                kahan_add(sx[k], cx[k], px[k]);
                kahan_add(sy[k], cy[k], py[k]);
                kahan_add(sz[k], cz[k], pz[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        int i = ibase + k * kBlock;
        if (i < n) {
            float *pp = partial + (size_t)split * 6 * n + i;
            pp[0] = sx[k];
            pp[(size_t)n] = sy[k];
            pp[2 * (size_t)n] = sz[k];
            pp[3 * (size_t)n] = cx[k];
            pp[4 * (size_t)n] = cy[k];
            pp[5 * (size_t)n] = cz[k];
        }
    }
}

__global__ void k_update_mixed(int64_t n, const float *__restrict__ partial, int nsplit, double *__restrict__ a,
                               double *__restrict__ v, double *__restrict__ x, const double *__restrict__ m, double G,
                               double half_dt, double dt, int mode, float4 *__restrict__ pos32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double s = 0.0;
        for (int q = 0; q < nsplit; ++q) {
            const float *pp = partial + (size_t)q * 6 * n + i;
            s += (double)pp[(size_t)c * n] - (double)pp[(size_t)(c + 3) * n];
        }
        acc[c] = s;
        a[c * n + i] = s;
    }
    if (mode == UPD_STORE) return;
    double vv[3] = {v[i], v[n + i], v[2 * n + i]};
#pragma unroll
    for (int c = 0; c < 3; ++c) vv[c] += half_dt * acc[c];  // closing half kick (loop_kdk :499)
    if (mode == UPD_CLOSE_OPEN) {
        double xx[3] = {x[i], x[n + i], x[2 * n + i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vv[c] += half_dt * acc[c];  // next step: kick half (evolve :2950-2955)
            xx[c] += dt * vv[c];        // drift
            x[c * n + i] = xx[c];
        }
        pos32[i] = make_float4((float)xx[0], (float)xx[1], (float)xx[2], (float)(G * m[i]));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c * n + i] = vv[c];
}

// FP64 pairwise force, reference operation order: for body i the contributions
// of j = 0..n-1 (j != i) are added in ascending j, which is the order in which
// the reference's i<j double loop delivers them (gregsys.hpp:477-497).
// Contribution of body j to a_i in the reference's pairwise update, before it
// is added to the running sum (the reference forms the term, then adds it).
template <int FORM>
__device__ __forceinline__ void ref_pair_term(double xi, double yi, double zi, double xj, double yj, double zj,
                                              double mj, double G, double eps2, double &tx, double &ty, double &tz) {
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    if (FORM == 0) {
        // cumulative_acc: r^3 = |(d.d) d|, k = (G m_j) / r^3
        const double s = dx * dx + dy * dy + dz * dz;
        const double wx = s * dx, wy = s * dy, wz = s * dz;
        const double r3 = __builtin_sqrt(wx * wx + wy * wy + wz * wz);
        const double k = (G * mj) / r3;
        tx = k * dx;
        ty = k * dy;
        tz = k * dz;
    } else {
        // bh_acc form: G*(m/s) * (d/sqrt(s)), s = d.d + eps^2
        const double s = dx * dx + dy * dy + dz * dz + eps2;
        const double sq = __builtin_sqrt(s);
        const double k = G * (mj / s);
        tx = k * (dx / sq);
        ty = k * (dy / sq);
        tz = k * (dz / sq);
    }
}

// FP64 pairwise force, reference operation order: for body i the contributions
// of j = 0..n-1 (j != i) are added in ascending j, which is the order in which
// the reference's i<j double loop delivers them (gregsys.hpp:477-497).  One
// lane per i; used for large n, where n lanes fill the GPU.
template <int FORM>
__global__ __launch_bounds__(kBlock) void k_force_fp64(int64_t n, const double *__restrict__ x,
                                                       const double *__restrict__ m, double G, double eps2,
                                                       double *__restrict__ a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const int64_t ii = live ? i : 0;
    const double xi = x[ii], yi = x[n + ii], zi = x[2 * n + ii];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        if (j == ii) continue;
        double tx, ty, tz;
        ref_pair_term<FORM>(xi, yi, zi, x[j], x[n + j], x[2 * n + j], m[j], G, eps2, tx, ty, tz);
        ax += tx;
        ay += ty;
        az += tz;
    }
    if (live) {
        a[i] = ax;
        a[n + i] = ay;
        a[2 * n + i] = az;
    }
}

// The same sums in two phases, for small n (where one lane per i leaves the
// GPU idle): k_pair_terms evaluates the terms of a block of j for all i in
// parallel into terms[j - j0][3][n] (0 for j == i), then k_ordered_sum adds
// them to the running sums in ascending j, one lane per (component, i).  The
// additions happen in the same order on the same values, so the result is
// bit-identical to k_force_fp64 (adding the +0 of j == i leaves a sum that
// started at +0 unchanged).
template <int FORM>
__global__ __launch_bounds__(kBlock) void k_pair_terms(int64_t n, const double *__restrict__ x,
                                                       const double *__restrict__ m, double G, double eps2,
                                                       int64_t j0, int64_t jcnt, double *__restrict__ terms) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t jb = j0 + (int64_t)blockIdx.y * kTermJ;
    const int64_t je = min(j0 + jcnt, jb + kTermJ);
    const double xi = x[i], yi = x[n + i], zi = x[2 * n + i];
    for (int64_t j = jb; j < je; ++j) {
        double tx = 0.0, ty = 0.0, tz = 0.0;
        if (j != i) ref_pair_term<FORM>(xi, yi, zi, x[j], x[n + j], x[2 * n + j], m[j], G, eps2, tx, ty, tz);
        double *t = terms + (size_t)(j - j0) * 3 * n + i;
        t[0] = tx;
        t[n] = ty;
        t[2 * n] = tz;
    }
}

__global__ __launch_bounds__(kBlock) void k_ordered_sum(int64_t n3, int64_t jcnt, const double *__restrict__ terms,
                                                        double *__restrict__ a, int first) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n3) return;
    double s = first ? 0.0 : a[t];
    const double *p = terms + t;
#pragma unroll 16
    for (int64_t j = 0; j < jcnt; ++j) s += p[(size_t)j * n3];
    a[t] = s;
}

// FP64 pair forces with the j range split twice: over gridDim.y, and inside a
// workgroup over its kWaves64 waves, which share the same 64 i-bodies (one per
// lane).  The waves' sums are combined in LDS in wave order and one partial
// per (split, i) is written, [split][3][n]; splits are combined in fixed order
// by k_update_fp64_split, so the result is deterministic (but not the
// reference's summation order).
template <bool SOFT>
__global__ __launch_bounds__(64 * kWaves64) void k_force_fp64_split(int64_t n, const double *__restrict__ x,
                                                                    const double *__restrict__ m, int jchunk,
                                                                    double eps2, double *__restrict__ partial) {
#pragma clang fp contract(fast)
    __shared__ double red[kWaves64 - 1][3][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t c0 = (int64_t)blockIdx.y * jchunk;
    const int64_t c1 = min(n, c0 + jchunk);
    const int64_t wq = (jchunk + kWaves64 - 1) / kWaves64;
    const int64_t j0 = c0 + wave * wq;
    const int64_t j1 = min(c1, j0 + wq);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const int64_t ii = i < n ? i : 0;
    const double xi = x[ii], yi = x[n + ii], zi = x[2 * n + ii];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = j0; j < j1; ++j) {
        const double dx = x[j] - xi, dy = x[n + j] - yi, dz = x[2 * n + j] - zi;
        const double s = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
        double inv = rsqrt64(s);
        if (!SOFT) inv = s > 0.0 ? inv : 0.0;
        const double w = m[j] * (inv * inv * inv);
        ax = __builtin_fma(w, dx, ax);
        ay = __builtin_fma(w, dy, ay);
        az = __builtin_fma(w, dz, az);
    }
    if (wave > 0) {
        red[wave - 1][0][lane] = ax;
        red[wave - 1][1][lane] = ay;
        red[wave - 1][2][lane] = az;
    }
    __syncthreads();
    if (wave == 0 && i < n) {
#pragma unroll
        for (int w = 0; w < kWaves64 - 1; ++w) {
            ax += red[w][0][lane];
            ay += red[w][1][lane];
            az += red[w][2][lane];
        }
        double *pp = partial + (size_t)blockIdx.y * 3 * n + i;
        pp[0] = ax;
        pp[n] = ay;
        pp[2 * n] = az;
    }
}

// Split reduction, then the KDK update for UPD_CLOSE / UPD_CLOSE_OPEN.  The
// drift reads xin and writes xout (the same array for an in-place step); when
// snap_x is set, the synchronised state (x before the drift, v after the
// closing kick) is also written there and to snap_v.
__global__ __launch_bounds__(64 * kWaves64) void k_update_fp64_split(int64_t n, const double *__restrict__ partial,
                                                                     int nsplit, double G, double *__restrict__ a,
                                                                     double *__restrict__ v, const double *xin,
                                                                     double *xout, double half_dt, double dt,
                                                                     int mode, double *__restrict__ snap_x,
                                                                     double *__restrict__ snap_v) {
    __shared__ double red[kWaves64 - 1][3][64];
    const int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    double acc[3];
    reduce_splits<3>(partial, n, nsplit, i, acc, red);
    if (threadIdx.x >= 64 || i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc[c] = G * acc[c];
        a[c * n + i] = acc[c];
    }
    if (mode == UPD_STORE) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double vv = v[c * n + i] + half_dt * acc[c];
        if (mode == UPD_CLOSE_OPEN) {
            const double xx = xin[c * n + i];
            if (snap_x) {
                snap_x[c * n + i] = xx;
                snap_v[c * n + i] = vv;
            }
            vv += half_dt * acc[c];
            xout[c * n + i] = xx + dt * vv;
        }
        v[c * n + i] = vv;
    }
}

// Per-body potential energy, reference order (cumulative_pe :358-363 then the
// per-body halving of calc_energy :620): the pair term -(G*m_lo*m_hi)/r uses
// the lower index's mass first, as the reference's b1/b2 roles do.
template <int FORM>
__global__ __launch_bounds__(kBlock) void k_pe_fp64(int64_t n, const double *__restrict__ x,
                                                    const double *__restrict__ m, double G, double eps2,
                                                    double *__restrict__ pe) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const int64_t ii = live ? i : 0;
    const double xi = x[ii], yi = x[n + ii], zi = x[2 * n + ii], mi = m[ii];
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        const double xj = x[j], yj = x[n + j], zj = x[2 * n + j], mj = m[j];
        if (j == ii) continue;
        double dx = xj - xi, dy = yj - yi, dz = zj - zi;
        double r = FORM == 0 ? __builtin_sqrt(dx * dx + dy * dy + dz * dz)
                             : __builtin_sqrt(dx * dx + dy * dy + dz * dz + eps2);
        double gmm = j < ii ? (G * mj) * mi : (G * mi) * mj;
        acc += -gmm / r;
    }
    if (live) pe[i] = acc / 2.0;
}

__global__ void k_kick(int64_t n, const double *__restrict__ a, double *__restrict__ v, double coef) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    v[i] += coef * a[i];
}

__global__ void k_kick_drift(int64_t n, const double *__restrict__ a, double *__restrict__ v,
                             double *__restrict__ x, double half_dt, double dt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    double vv = v[i] + half_dt * a[i];
    v[i] = vv;
    x[i] += dt * vv;
}

__global__ void k_drift(int64_t n, const double *__restrict__ v, double *__restrict__ x, double dt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    x[i] += dt * v[i];
}

// The other integrators of evolve, per component (3n entries), in the
// reference's operation order (gregsys.hpp:380-419).
// take_euler_step: x += v*dt ; v += a*dt
__global__ void k_euler(int64_t n, const double *__restrict__ a, double *__restrict__ v, double *__restrict__ x,
                        double dt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const double vv = v[i];
    x[i] += vv * dt;
    v[i] = vv + a[i] * dt;
}

// predicted state of modified Euler (h = dt) and midpoint (h = dt/2):
// px = x + h*v ; pv = v + h*a
__global__ void k_predict(int64_t n, const double *__restrict__ x, const double *__restrict__ v,
                          const double *__restrict__ a, double h, double *__restrict__ px, double *__restrict__ pv) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    px[i] = x[i] + h * v[i];
    pv[i] = v[i] + h * a[i];
}

// take_modified_euler_step: x += half_dt*(v + pv) ; v += half_dt*(a + pa)
__global__ void k_modeuler(int64_t n, const double *__restrict__ a, const double *__restrict__ pv,
                           const double *__restrict__ pa, double *__restrict__ v, double *__restrict__ x,
                           double half_dt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    const double vv = v[i];
    x[i] += half_dt * (vv + pv[i]);
    v[i] = vv + half_dt * (a[i] + pa[i]);
}

// take_midpoint_step: x += dt*pv ; v += dt*pa
__global__ void k_midpoint(int64_t n, const double *__restrict__ pv, const double *__restrict__ pa,
                           double *__restrict__ v, double *__restrict__ x, double dt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    x[i] += dt * pv[i];
    v[i] += dt * pa[i];
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

int choose_nsplit(int64_t n, int ipl, int waves_per_simd) {
    // waves in flight = ceil(n / (256 ipl)) * 4 * nsplit; 256 CUs x 4 SIMDs
    if (ipl <= 0) ipl = 2;
    if (waves_per_simd <= 0) waves_per_simd = 8;
    const int64_t wgs = (n + (int64_t)kBlock * ipl - 1) / ((int64_t)kBlock * ipl);
    const int64_t want = (int64_t)waves_per_simd * 1024 / 4;  // workgroups of 4 waves
    int64_t sp = (want + wgs - 1) / wgs;
    sp = std::max<int64_t>(1, std::min<int64_t>(sp, 64));
    while (sp > 1 && n / sp < kTile) --sp;  // at least one full Kahan tile of j per split
    return (int)sp;
}

void launch_pack(hipStream_t s, int64_t n, const double *x, const double *m, double G, float4 *pos32) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, m, G, pos32);
    NB_HIP(hipGetLastError());
}

template <int IPL>
static void launch_force_mixed_ipl(hipStream_t s, int64_t n, const float4 *pos32, float eps2, bool soft,
                                   float *partial, int nsplit, int variant) {
    int jchunk = (int)((n + nsplit - 1) / nsplit);
    jchunk = (jchunk + kTile - 1) / kTile * kTile;
    dim3 grid((unsigned)((n + kBlock * IPL - 1) / (kBlock * IPL)), (unsigned)nsplit);
    if (variant == 1) {
        if (soft)
            hipLaunchKernelGGL((k_force_mixed_lds<IPL, true>), grid, dim3(kBlock), 0, s, pos32, (int)n, jchunk,
                               eps2, partial);
        else
            hipLaunchKernelGGL((k_force_mixed_lds<IPL, false>), grid, dim3(kBlock), 0, s, pos32, (int)n, jchunk,
                               eps2, partial);
    } else if (soft) {
        hipLaunchKernelGGL((k_force_mixed<IPL, true>), grid, dim3(kBlock), 0, s, pos32, (int)n, jchunk, eps2,
                           partial);
    } else {
        hipLaunchKernelGGL((k_force_mixed<IPL, false>), grid, dim3(kBlock), 0, s, pos32, (int)n, jchunk, eps2,
                           partial);
    }
}

void launch_force_mixed(hipStream_t s, int64_t n, const float4 *pos32, float eps2, bool soft, float *partial,
                        int nsplit, int ipl, int variant) {
    if (n <= 0) return;
    if (ipl == 1) launch_force_mixed_ipl<1>(s, n, pos32, eps2, soft, partial, nsplit, variant);
    else if (ipl == 4) launch_force_mixed_ipl<4>(s, n, pos32, eps2, soft, partial, nsplit, variant);
    else launch_force_mixed_ipl<2>(s, n, pos32, eps2, soft, partial, nsplit, variant);
    NB_HIP(hipGetLastError());
}

void launch_update_mixed(hipStream_t s, int64_t n, const float *partial, int nsplit, double *a, double *v, double *x,
                         const double *m, double G, double half_dt, double dt, int mode, float4 *pos32) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_update_mixed, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, partial, nsplit, a, v, x, m, G,
                       half_dt, dt, mode, pos32);
    NB_HIP(hipGetLastError());
}

int64_t fp64_terms_block(int64_t n) {
    if (n <= 1 || n > kTermsMaxBodies) return 0;
    return std::min<int64_t>(n, std::max<int64_t>(kTermJ, kTermsBudget / (3 * n)));
}

void launch_force_fp64(hipStream_t s, int64_t n, const double *x, const double *m, double G, double eps2, int form,
                       double *a, double *terms) {
    if (n <= 0) return;
    const int64_t jb = fp64_terms_block(n);
    if (jb > 0 && terms) {
        const int64_t n3 = 3 * n;
        for (int64_t j0 = 0; j0 < n; j0 += jb) {
            const int64_t cnt = std::min<int64_t>(jb, n - j0);
            const dim3 grid(blocks_for(n), (unsigned)((cnt + kTermJ - 1) / kTermJ));
            if (form == 0)
                hipLaunchKernelGGL(k_pair_terms<0>, grid, dim3(kBlock), 0, s, n, x, m, G, eps2, j0, cnt, terms);
            else
                hipLaunchKernelGGL(k_pair_terms<1>, grid, dim3(kBlock), 0, s, n, x, m, G, eps2, j0, cnt, terms);
            hipLaunchKernelGGL(k_ordered_sum, dim3(blocks_for(n3)), dim3(kBlock), 0, s, n3, cnt, terms, a,
                               j0 == 0 ? 1 : 0);
        }
    } else if (form == 0) {
        hipLaunchKernelGGL(k_force_fp64<0>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, m, G, eps2, a);
    } else {
        hipLaunchKernelGGL(k_force_fp64<1>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, m, G, eps2, a);
    }
    NB_HIP(hipGetLastError());
}

// Splits of the j range for the FP64 split kernels: enough workgroups (64
// i-bodies x kWaves64 waves each) to fill the GPU, with at least 16 j per wave.
int choose_nsplit64(int64_t n) {
    const int64_t groups = (n + 63) / 64;
    int64_t sp = (2048 + groups - 1) / groups;
    sp = std::max<int64_t>(1, std::min<int64_t>(sp, 64));
    while (sp > 1 && n / (sp * kWaves64) < 16) --sp;
    return (int)sp;
}

void launch_force_fp64_split(hipStream_t s, int64_t n, const double *x, const double *m, double eps2, bool soft,
                             double *partial, int nsplit) {
    if (n <= 0) return;
    const int jchunk = (int)((n + nsplit - 1) / nsplit);
    dim3 grid((unsigned)((n + 63) / 64), (unsigned)nsplit);
    if (soft)
        hipLaunchKernelGGL(k_force_fp64_split<true>, grid, dim3(64 * kWaves64), 0, s, n, x, m, jchunk, eps2,
                           partial);
    else
        hipLaunchKernelGGL(k_force_fp64_split<false>, grid, dim3(64 * kWaves64), 0, s, n, x, m, jchunk, eps2,
                           partial);
    NB_HIP(hipGetLastError());
}

void launch_update_fp64_split(hipStream_t s, int64_t n, const double *partial, int nsplit, double G, double *a,
                              double *v, double *x, double half_dt, double dt, int mode) {
    launch_update_fp64_split_ex(s, n, partial, nsplit, G, a, v, x, x, half_dt, dt, mode, nullptr, nullptr);
}

void launch_update_fp64_split_ex(hipStream_t s, int64_t n, const double *partial, int nsplit, double G, double *a,
                                 double *v, const double *xin, double *xout, double half_dt, double dt, int mode,
                                 double *snap_x, double *snap_v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_update_fp64_split, dim3((unsigned)((n + 63) / 64)), dim3(64 * kWaves64), 0, s, n, partial,
                       nsplit, G, a, v, xin, xout, half_dt, dt, mode, snap_x, snap_v);
    NB_HIP(hipGetLastError());
}

void launch_pe_fp64(hipStream_t s, int64_t n, const double *x, const double *m, double G, double eps2, int form,
                    double *pe) {
    if (n <= 0) return;
    if (form == 0)
        hipLaunchKernelGGL(k_pe_fp64<0>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, m, G, eps2, pe);
    else
        hipLaunchKernelGGL(k_pe_fp64<1>, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, m, G, eps2, pe);
    NB_HIP(hipGetLastError());
}

void launch_kick(hipStream_t s, int64_t n, const double *a, double *v, double coef) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_kick, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, a, v, coef);
    NB_HIP(hipGetLastError());
}

void launch_kick_drift(hipStream_t s, int64_t n, const double *a, double *v, double *x, double half_dt, double dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_kick_drift, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, a, v, x, half_dt, dt);
    NB_HIP(hipGetLastError());
}

void launch_euler(hipStream_t s, int64_t n, const double *a, double *v, double *x, double dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_euler, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, a, v, x, dt);
    NB_HIP(hipGetLastError());
}

void launch_predict(hipStream_t s, int64_t n, const double *x, const double *v, const double *a, double h,
                    double *px, double *pv) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_predict, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, x, v, a, h, px, pv);
    NB_HIP(hipGetLastError());
}

void launch_modeuler(hipStream_t s, int64_t n, const double *a, const double *pv, const double *pa, double *v,
                     double *x, double half_dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_modeuler, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, a, pv, pa, v, x, half_dt);
    NB_HIP(hipGetLastError());
}

void launch_midpoint(hipStream_t s, int64_t n, const double *pv, const double *pa, double *v, double *x, double dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_midpoint, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, pv, pa, v, x, dt);
    NB_HIP(hipGetLastError());
}

void launch_drift(hipStream_t s, int64_t n, const double *v, double *x, double dt) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_drift, dim3(blocks_for(3 * n)), dim3(kBlock), 0, s, n, v, x, dt);
    NB_HIP(hipGetLastError());
}

}  // namespace nbe
