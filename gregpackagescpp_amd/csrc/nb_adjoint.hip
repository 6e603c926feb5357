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

// No contraction: the kick and drift recomputed here (v + h a, x + dt v) must
// round as the forward step's update kernels do, so that recomputed segments
// land on the forward trajectory bit for bit.  The pair passes below state
// their fused multiply-adds explicitly.
#pragma clang fp contract(off)

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

// lx += scale * sum(partial xyz) ; lm += scale * sum(partial m), fixed split
// order; then, when lv is set, lv += dt * lx (the next reverse step's kick).
__global__ __launch_bounds__(64 * kWaves64) void k_vjp_apply(int64_t n, const double *__restrict__ partial,
                                                             int nsplit, double scale, double *__restrict__ lx,
                                                             double *__restrict__ lm, double *__restrict__ lv,
                                                             double dt) {
    __shared__ double red[kWaves64 - 1][4][64];
    const int64_t k = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    double acc[4];
    reduce_splits<4>(partial, n, nsplit, k, acc, red);
    if (threadIdx.x >= 64 || k >= n) return;
    for (int c = 0; c < 3; ++c) {
        const double l = lx[c * n + k] + scale * acc[c];
        lx[c * n + k] = l;
        if (lv) lv[c * n + k] += dt * l;
    }
    lm[k] += scale * acc[3];
}

__global__ void k_copy(int64_t len, const double *__restrict__ src, double *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) dst[i] = src[i];
}

// Opening half kick and drift into another position array:
// v += h a ; xout = xin + dt v   (3n entries).
__global__ void k_kick_drift_to(int64_t len, const double *__restrict__ a, double *__restrict__ v,
                                const double *__restrict__ xin, double *__restrict__ xout, double h, double dt) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= len) return;
    const double vv = v[i] + h * a[i];
    v[i] = vv;
    xout[i] = xin[i] + dt * vv;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct AdjWork {
    DBuf<double> ck, traj, lx, lv, lm, vpart, xT, vT, stage;
    StepGraph fwd_seg, rev_seg;
};

}  // namespace

// Adjoint driver.  Launch structure (all on the context's stream):
//   forward:  a(x0); snapshot 0; opening kick+drift; then per step one force
//             pass and one update that closes the step, snapshots the state
//             at segment ends, and opens the next step (2 launches/step);
//   reverse:  per segment, restore its snapshot, recompute its positions into
//             traj[0..len] (2 launches/step), then per step one VJP pair pass
//             and one apply that also performs the next step's lv += dt lx.
// With graphs on, every full forward segment and every interior reverse
// segment is one graph launch; segment snapshots pass through a fixed
// staging buffer so that the captured pointers do not change.
void adjoint_run(nb_system *s, uint64_t steps, uint64_t ckpt, const double *gxT, const double *gvT, double *gx0,
                 double *gv0, double *gm, int64_t *fwd_ns, int64_t *bwd_ns) {
    const int64_t n = s->n;
    const size_t n3 = 3 * (size_t)n;
    hipStream_t st = s->stream;
    if (ckpt > steps) ckpt = steps;
    const uint64_t nseg = (steps + ckpt - 1) / ckpt;
    const bool graphs = s->tune_step_graphs != 2;
    AdjWork w;
    w.ck.ensure(nseg * 2 * n3);
    w.traj.ensure((ckpt + 1) * n3);
    w.lx.ensure(n3);
    w.lv.ensure(n3);
    w.lm.ensure((size_t)n);
    w.xT.ensure(n3);
    w.vT.ensure(n3);
    w.stage.ensure(2 * n3);
    s->dw.nsplit64 = choose_nsplit64(n);
    s->dw.partial64.ensure((size_t)s->dw.nsplit64 * 3 * (size_t)n);
    const int nsplit = choose_nsplit64(n);
    w.vpart.ensure((size_t)nsplit * 4 * (size_t)n);
    const int jchunk = (int)((n + nsplit - 1) / nsplit);
    const dim3 vgrid((unsigned)((n + 63) / 64), (unsigned)nsplit);
    const double h = s->half_dt, dt = s->dt, eps2 = s->eps * s->eps;
    const bool soft = s->eps > 0;
    const unsigned b3 = blocks((int64_t)n3);

    auto force_at = [&](const double *xs) {
        launch_force_fp64_split(st, n, xs, s->m.p, eps2, soft, s->dw.partial64.p, s->dw.nsplit64);
    };
    auto update = [&](const double *xin, double *xout, int mode, double *sx, double *sv) {
        launch_update_fp64_split_ex(st, n, s->dw.partial64.p, s->dw.nsplit64, s->G, s->a.p, s->v.p, xin, xout, h,
                                    dt, mode, sx, sv);
    };
    auto copy = [&](const double *src, double *dst, size_t len) {
        hipLaunchKernelGGL(k_copy, dim3(blocks((int64_t)len)), dim3(kBlock), 0, st, (int64_t)len, src, dst);
    };
    // steps q0 .. q0+len-1 of the forward run (state opened for step q0);
    // the last step of a segment that is not the run's last writes the
    // synchronised state to the staging buffer
    auto fwd_steps = [&](uint64_t q0, uint64_t len) {
        for (uint64_t q = q0; q < q0 + len; ++q) {
            force_at(s->x.p);
            if (q + 1 == steps)
                update(s->x.p, s->x.p, UPD_CLOSE, nullptr, nullptr);
            else if (q + 1 == q0 + len)
                update(s->x.p, s->x.p, UPD_CLOSE_OPEN, w.stage.p, w.stage.p + n3);
            else
                update(s->x.p, s->x.p, UPD_CLOSE_OPEN, nullptr, nullptr);
        }
    };

    // ---- forward with checkpoints ----
    NB_HIP(hipStreamSynchronize(st));
    auto t0 = std::chrono::high_resolution_clock::now();
    force_at(s->x.p);
    update(s->x.p, s->x.p, UPD_STORE, nullptr, nullptr);
    copy(s->x.p, w.ck.p, n3);
    copy(s->v.p, w.ck.p + n3, n3);
    hipLaunchKernelGGL(k_kick_drift_to, dim3(b3), dim3(kBlock), 0, st, (int64_t)n3, s->a.p, s->v.p, s->x.p, s->x.p,
                       h, dt);
    for (uint64_t seg = 0; seg < nseg; ++seg) {
        const uint64_t q0 = seg * ckpt, len = std::min<uint64_t>(ckpt, steps - q0);
        const bool last = seg + 1 == nseg;
        if (graphs && !last && len == ckpt) {
            if (w.fwd_seg.valid() || w.fwd_seg.capture(st, [&]() { fwd_steps(q0, len); }))
                w.fwd_seg.launch(st);
            else
                fwd_steps(q0, len);  // capture invalidated from outside: the same launches, eagerly
        } else {
            fwd_steps(q0, len);
        }
        if (!last) copy(w.stage.p, w.ck.p + (seg + 1) * 2 * n3, 2 * n3);
    }
    NB_HIP(hipGetLastError());
    NB_HIP(hipStreamSynchronize(st));
    auto t1 = std::chrono::high_resolution_clock::now();
    copy(s->x.p, w.xT.p, n3);
    copy(s->v.p, w.vT.p, n3);

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

    // lx += coef G J(xs)^T lv, lm += coef G (dA/dm)^T lv; then, when `open`,
    // lv += dt lx for the next reverse step
    auto vjp = [&](const double *xs, double coef, bool open) {
        if (soft)
            hipLaunchKernelGGL(k_vjp_split<true>, vgrid, dim3(64 * kWaves64), 0, st, n, xs, s->m.p, w.lv.p, jchunk,
                               eps2, w.vpart.p);
        else
            hipLaunchKernelGGL(k_vjp_split<false>, vgrid, dim3(64 * kWaves64), 0, st, n, xs, s->m.p, w.lv.p, jchunk,
                               eps2, w.vpart.p);
        hipLaunchKernelGGL(k_vjp_apply, dim3((unsigned)((n + 63) / 64)), dim3(64 * kWaves64), 0, st, n, w.vpart.p,
                           nsplit, coef * s->G, w.lx.p, w.lm.p, open ? w.lv.p : nullptr, dt);
    };
    // recompute of a segment from its restored state (traj[0], v) and its
    // reverse sweep; `first` adds the closing VJP at x_T
    auto rev_segment = [&](uint64_t start, uint64_t len, bool first) {
        double *tr = w.traj.p;
        force_at(tr);
        update(tr, tr, UPD_STORE, nullptr, nullptr);
        hipLaunchKernelGGL(k_kick_drift_to, dim3(b3), dim3(kBlock), 0, st, (int64_t)n3, s->a.p, s->v.p, tr, tr + n3,
                           h, dt);
        for (uint64_t q = 1; q <= len; ++q) {
            force_at(tr + q * n3);
            if (q < len)
                update(tr + q * n3, tr + (q + 1) * n3, UPD_CLOSE_OPEN, nullptr, nullptr);
            else
                update(tr + q * n3, tr + q * n3, UPD_CLOSE, nullptr, nullptr);
        }
        if (first) vjp(tr + len * n3, h, true);  // closing kick of the last step, at x_T
        for (uint64_t q = len; q-- > 0;) {
            const uint64_t sg = start + q;
            vjp(tr + q * n3, sg > 0 ? 2.0 * h : h, sg > 0);  // opening kick of sg (+ closing kick of sg-1)
        }
    };

    // ---- reverse sweep, segment by segment ----
    auto t2 = std::chrono::high_resolution_clock::now();
    for (uint64_t seg = nseg; seg-- > 0;) {
        const uint64_t start = seg * ckpt;
        const uint64_t len = std::min<uint64_t>(ckpt, steps - start);
        const double *c = w.ck.p + seg * 2 * n3;
        copy(c, w.traj.p, n3);
        copy(c + n3, s->v.p, n3);
        const bool first = seg + 1 == nseg;
        if (graphs && !first && seg > 0 && len == ckpt) {
            if (w.rev_seg.valid() || w.rev_seg.capture(st, [&]() { rev_segment(start, len, false); }))
                w.rev_seg.launch(st);
            else
                rev_segment(start, len, false);
        } else {
            rev_segment(start, len, first);
        }
    }
    NB_HIP(hipGetLastError());
    NB_HIP(hipStreamSynchronize(st));
    auto t3 = std::chrono::high_resolution_clock::now();

    // ---- gradients out; state back to time T ----
    copy_wait(tmp.data(), w.lx.p, n3 * 8, hipMemcpyDeviceToHost, st);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) gx0[3 * i + c] = tmp[(size_t)(c * n + i)];
    copy_wait(tmp.data(), w.lv.p, n3 * 8, hipMemcpyDeviceToHost, st);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) gv0[3 * i + c] = tmp[(size_t)(c * n + i)];
    copy_wait(gm, w.lm.p, (size_t)n * 8, hipMemcpyDeviceToHost, st);
    copy(w.xT.p, s->x.p, n3);
    copy(w.vT.p, s->v.p, n3);
    force_at(s->x.p);
    update(s->x.p, s->x.p, UPD_STORE, nullptr, nullptr);
    NB_HIP(hipStreamSynchronize(st));
    s->acc_valid = s->precision == NB_PREC_FP64_FAST;
    s->evolved = true;
    s->acc_method = NB_LEAPFROG_KDK;
    if (fwd_ns) *fwd_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    if (bwd_ns) *bwd_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t3 - t2).count();
}

}  // namespace nbe
