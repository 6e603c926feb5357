// nb_adjoint.hip -- reverse-time adjoint of the direct-summation KDK step (gfx950).
//
// No counterpart exists in the reference; the forward it differentiates is
// evolve's KDK branch (gregsys.hpp:2946-2968) with the softened pair force of
// bh_acc (:371-375) (or the unsoftened form at eps = 0).
//
// One KDK step:  v' = v + h a(x) ; x1 = x + dt v' ; v1 = v' + h a(x1),  h = dt/2,
// with a_i = G sum_j m_j d_ij s_ij^{-3/2}, d_ij = x_j - x_i, s = |d|^2 + eps^2.
// Its adjoint, for cotangents (lx, lv) of (x1, v1):
//   lx += h J(x1)^T lv ;  lv += dt lx ;  lx += h J(x)^T lv      (and lm alike)
// where, with w = m_k u_j - m_j u_k,
//   (J^T u)_k = G sum_j [ s^{-3/2} w - 3 s^{-5/2} d_kj (d_kj . w) ]
//   (dA/dm)^T u |_k = -G sum_j s^{-3/2} (u_j . d_kj).
// The opening VJP of step s and the closing VJP of step s-1 sit at the same
// positions with the same lv, so they are fused into one VJP with 2h: one pair
// pass per reverse step.  Forward states are checkpointed every `ckpt` steps;
// each segment is recomputed forward (positions stored) before its reverse
// sweep.  All arithmetic is FP64; the pair passes split the j range over the
// GPU and combine the splits in fixed order.
#include <chrono>
#include <vector>

#include "nb_split64.hpp"
#include "nb_system.hpp"

namespace nbe {

namespace {
constexpr int kBlock = 256;

// One VJP pair pass: for the 64 k-bodies of a workgroup (one per lane) the
// split's j range is divided over its kWaves64 waves; the waves' sums are
// combined in LDS in wave order and one partial per (split, k) is written,
// [split][4][n] = (J^T u)_xyz and (dA/dm)^T u.
template <bool SOFT>
__global__ __launch_bounds__(64 * kWaves64) void k_vjp_split(int64_t n, const double *__restrict__ x,
                                                             const double *__restrict__ m,
                                                             const double *__restrict__ u, int jchunk, double eps2,
                                                             double *__restrict__ partial) {
    __shared__ double red[kWaves64 - 1][4][64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t c0 = (int64_t)blockIdx.y * jchunk;
    const int64_t c1 = min(n, c0 + jchunk);
    const int64_t wq = (jchunk + kWaves64 - 1) / kWaves64;
    const int64_t j0 = c0 + wave * wq;
    const int64_t j1 = min(c1, j0 + wq);
    const int64_t k = (int64_t)blockIdx.x * 64 + lane;
    const int64_t kk = k < n ? k : 0;
    const double xk = x[kk], yk = x[n + kk], zk = x[2 * n + kk], mk = m[kk];
    const double ukx = u[kk], uky = u[n + kk], ukz = u[2 * n + kk];
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j = j0; j < j1; ++j) {
        const double dx = x[j] - xk, dy = x[n + j] - yk, dz = x[2 * n + j] - zk;
        const double mj = m[j];
        const double ujx = u[j], ujy = u[n + j], ujz = u[2 * n + j];
        const double s = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
        double inv = rsqrt64(s);
        if (!SOFT) inv = s > 0.0 ? inv : 0.0;
        const double inv2 = inv * inv;
        const double inv3 = inv * inv2;
        const double inv5x3 = 3.0 * inv3 * inv2;
        const double wx = __builtin_fma(mk, ujx, -mj * ukx);
        const double wy = __builtin_fma(mk, ujy, -mj * uky);
        const double wz = __builtin_fma(mk, ujz, -mj * ukz);
        const double dw = __builtin_fma(dz, wz, __builtin_fma(dy, wy, dx * wx));
        const double t = inv5x3 * dw;
        g[0] = __builtin_fma(inv3, wx, __builtin_fma(-t, dx, g[0]));
        g[1] = __builtin_fma(inv3, wy, __builtin_fma(-t, dy, g[1]));
        g[2] = __builtin_fma(inv3, wz, __builtin_fma(-t, dz, g[2]));
        const double ud = __builtin_fma(ujz, dz, __builtin_fma(ujy, dy, ujx * dx));
        g[3] = __builtin_fma(-inv3, ud, g[3]);
    }
    if (wave > 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) red[wave - 1][c][lane] = g[c];
    __syncthreads();
    if (wave == 0 && k < n) {
#pragma unroll
        for (int w = 0; w < kWaves64 - 1; ++w)
#pragma unroll
            for (int c = 0; c < 4; ++c) g[c] += red[w][c][lane];
        double *pp = partial + (size_t)blockIdx.y * 4 * n + k;
#pragma unroll
        for (int c = 0; c < 4; ++c) pp[(size_t)c * n] = g[c];
    }
}

// lx += scale * sum(partial xyz) ; lm += scale * sum(partial m), fixed split order
__global__ __launch_bounds__(64 * kWaves64) void k_vjp_apply(int64_t n, const double *__restrict__ partial,
                                                             int nsplit, double scale, double *__restrict__ lx,
                                                             double *__restrict__ lm) {
    __shared__ double red[kWaves64 - 1][4][64];
    const int64_t k = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    double acc[4];
    reduce_splits<4>(partial, n, nsplit, k, acc, red);
    if (threadIdx.x >= 64 || k >= n) return;
    for (int c = 0; c < 3; ++c) lx[c * n + k] += scale * acc[c];
    lm[k] += scale * acc[3];
}

__global__ void k_axpy(int64_t len, double a, const double *__restrict__ xv, double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) y[i] += a * xv[i];
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct AdjWork {
    DBuf<double> ck, traj, lx, lv, lm, vpart, xT, vT;
};

}  // namespace

// Forward KDK pieces on the FP64 split kernels.
static void fast_forces(nb_system *s) {
    const int64_t n = s->n;
    s->dw.nsplit64 = choose_nsplit64(n);
    s->dw.partial64.ensure((size_t)s->dw.nsplit64 * 3 * (size_t)n);
    launch_force_fp64_split(s->stream, n, s->x.p, s->m.p, s->eps * s->eps, s->eps > 0, s->dw.partial64.p,
                            s->dw.nsplit64);
    launch_update_fp64_split(s->stream, n, s->dw.partial64.p, s->dw.nsplit64, s->G, s->a.p, s->v.p, s->x.p,
                             s->half_dt, s->dt, UPD_STORE);
}

static void fast_step(nb_system *s) {
    const int64_t n = s->n;
    launch_kick_drift(s->stream, n, s->a.p, s->v.p, s->x.p, s->half_dt, s->dt);
    launch_force_fp64_split(s->stream, n, s->x.p, s->m.p, s->eps * s->eps, s->eps > 0, s->dw.partial64.p,
                            s->dw.nsplit64);
    launch_update_fp64_split(s->stream, n, s->dw.partial64.p, s->dw.nsplit64, s->G, s->a.p, s->v.p, s->x.p,
                             s->half_dt, s->dt, UPD_CLOSE);
}

void adjoint_run(nb_system *s, uint64_t steps, uint64_t ckpt, const double *gxT, const double *gvT, double *gx0,
                 double *gv0, double *gm, int64_t *fwd_ns, int64_t *bwd_ns) {
    const int64_t n = s->n;
    const size_t n3 = 3 * (size_t)n;
    hipStream_t st = s->stream;
    if (ckpt > steps) ckpt = steps;
    const uint64_t nseg = (steps + ckpt - 1) / ckpt;
    AdjWork w;
    w.ck.ensure(nseg * 2 * n3);
    w.traj.ensure((ckpt + 1) * n3);
    w.lx.ensure(n3);
    w.lv.ensure(n3);
    w.lm.ensure((size_t)n);
    w.xT.ensure(n3);
    w.vT.ensure(n3);
    const int nsplit = choose_nsplit64(n);
    w.vpart.ensure((size_t)nsplit * 4 * (size_t)n);
    const int jchunk = (int)((n + nsplit - 1) / nsplit);
    const dim3 vgrid((unsigned)((n + 63) / 64), (unsigned)nsplit);
    const double h = s->half_dt, dt = s->dt, eps2 = s->eps * s->eps;
    const bool soft = s->eps > 0;

    // ---- forward with checkpoints ----
    NB_HIP(hipStreamSynchronize(st));
    auto t0 = std::chrono::high_resolution_clock::now();
    fast_forces(s);
    for (uint64_t q = 0; q <= steps; ++q) {
        if (q % ckpt == 0 && q < steps) {
            double *c = w.ck.p + (q / ckpt) * 2 * n3;
            NB_HIP(hipMemcpyAsync(c, s->x.p, n3 * 8, hipMemcpyDeviceToDevice, st));
            NB_HIP(hipMemcpyAsync(c + n3, s->v.p, n3 * 8, hipMemcpyDeviceToDevice, st));
        }
        if (q < steps) {
            ScopedTimer tm(s->prof, "force", st);
            fast_step(s);
        }
    }
    NB_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::high_resolution_clock::now();
    NB_HIP(hipMemcpyAsync(w.xT.p, s->x.p, n3 * 8, hipMemcpyDeviceToDevice, st));
    NB_HIP(hipMemcpyAsync(w.vT.p, s->v.p, n3 * 8, hipMemcpyDeviceToDevice, st));

    // ---- cotangents in ----
    std::vector<double> tmp(n3);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) tmp[(size_t)(c * n + i)] = gxT[3 * i + c];
    NB_HIP(hipMemcpyAsync(w.lx.p, tmp.data(), n3 * 8, hipMemcpyHostToDevice, st));
    NB_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) tmp[(size_t)(c * n + i)] = gvT[3 * i + c];
    NB_HIP(hipMemcpyAsync(w.lv.p, tmp.data(), n3 * 8, hipMemcpyHostToDevice, st));
    NB_HIP(hipMemsetAsync(w.lm.p, 0, (size_t)n * 8, st));
    NB_HIP(hipStreamSynchronize(st));

    auto vjp = [&](const double *xs, double coef) {
        ScopedTimer tm(s->prof, "adjoint", st);
        if (soft)
            hipLaunchKernelGGL(k_vjp_split<true>, vgrid, dim3(64 * kWaves64), 0, st, n, xs, s->m.p, w.lv.p, jchunk,
                               eps2, w.vpart.p);
        else
            hipLaunchKernelGGL(k_vjp_split<false>, vgrid, dim3(64 * kWaves64), 0, st, n, xs, s->m.p, w.lv.p, jchunk,
                               eps2, w.vpart.p);
        hipLaunchKernelGGL(k_vjp_apply, dim3((unsigned)((n + 63) / 64)), dim3(64 * kWaves64), 0, st, n, w.vpart.p,
                           nsplit, coef * s->G, w.lx.p, w.lm.p);
    };

    // ---- reverse sweep, segment by segment ----
    auto t2 = std::chrono::high_resolution_clock::now();
    bool first = true;
    for (uint64_t seg = nseg; seg-- > 0;) {
        const uint64_t start = seg * ckpt;
        const uint64_t len = std::min<uint64_t>(ckpt, steps - start);
        const double *c = w.ck.p + seg * 2 * n3;
        NB_HIP(hipMemcpyAsync(s->x.p, c, n3 * 8, hipMemcpyDeviceToDevice, st));
        NB_HIP(hipMemcpyAsync(s->v.p, c + n3, n3 * 8, hipMemcpyDeviceToDevice, st));
        fast_forces(s);
        NB_HIP(hipMemcpyAsync(w.traj.p, s->x.p, n3 * 8, hipMemcpyDeviceToDevice, st));
        for (uint64_t q = 1; q <= len; ++q) {
            ScopedTimer tm(s->prof, "force", st);
            fast_step(s);
            NB_HIP(hipMemcpyAsync(w.traj.p + q * n3, s->x.p, n3 * 8, hipMemcpyDeviceToDevice, st));
        }
        for (uint64_t q = len; q-- > 0;) {
            const uint64_t sg = start + q;
            if (first) {
                vjp(w.traj.p + len * n3, h);  // closing kick of the last step, at x_T
                first = false;
            }
            hipLaunchKernelGGL(k_axpy, dim3(blocks((int64_t)n3)), dim3(kBlock), 0, st, (int64_t)n3, dt, w.lx.p,
                               w.lv.p);
            vjp(w.traj.p + q * n3, sg > 0 ? 2.0 * h : h);  // opening kick of sg (+ closing kick of sg-1)
        }
    }
    NB_HIP(hipGetLastError());
    NB_HIP(hipStreamSynchronize(st));
    auto t3 = std::chrono::high_resolution_clock::now();

    // ---- gradients out; state back to time T ----
    NB_HIP(hipMemcpy(tmp.data(), w.lx.p, n3 * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) gx0[3 * i + c] = tmp[(size_t)(c * n + i)];
    NB_HIP(hipMemcpy(tmp.data(), w.lv.p, n3 * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) gv0[3 * i + c] = tmp[(size_t)(c * n + i)];
    NB_HIP(hipMemcpy(gm, w.lm.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    NB_HIP(hipMemcpyAsync(s->x.p, w.xT.p, n3 * 8, hipMemcpyDeviceToDevice, st));
    NB_HIP(hipMemcpyAsync(s->v.p, w.vT.p, n3 * 8, hipMemcpyDeviceToDevice, st));
    fast_forces(s);
    NB_HIP(hipStreamSynchronize(st));
    s->prof.collect();
    s->acc_valid = s->precision == NB_PREC_FP64_FAST;
    s->evolved = true;
    s->acc_method = NB_LEAPFROG_KDK;
    if (fwd_ns) *fwd_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    if (bwd_ns) *bwd_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t3 - t2).count();
}

}  // namespace nbe
