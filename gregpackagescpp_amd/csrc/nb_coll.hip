// nb_coll.hip -- body-body collision detect / bounce / merge (gfx950), for one
// system and for batched ensembles of independent small systems.
//
// Reference path: system::s_coll (gregsys.hpp:690-914) with and without
// -DGREGSYS_MERGERS, the merge constructor and operator+= of gtd::body
// (gregbod.hpp:933-940, 1065-1082, 1094-1150), and the KDK driver that calls
// s_coll after each step's closing kick (FUNC_TEMPL_SELECT :16-57).
//
// s_coll is serial and order dependent (ordered map keyed by distance with
// equal keys dropped, axis momenta captured at scan time, one max_mom per call,
// a full rescan after every merge), so it is replayed exactly, by one
// workgroup per system:
//  * every pair is first tested in parallel (the reference's own FP64 test
//    r_i + r_j > |p_j - p_i|); outer bodies with no overlapping j > i cannot
//    change anything and are skipped (positions do not move during the pass
//    until a merge, after which every later outer body is scanned);
//  * a scanned outer body's j range is split over the threads; the merge
//    candidate is the LAST j reaching the largest curr_mom (the sequential
//    "curr >= max_mom" record rule), found by a block reduction;
//  * bounce entries are sorted by (distance, j) -- equal distances keep the
//    smallest j, as the map's emplace does -- and applied in order by one
//    thread, since each entry updates the outer body's velocity.
// Arithmetic is FP64 with contraction off; the merged radius uses a fixed
// Newton cube root (bit-reproducible, within 1 ulp of cbrt).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "nb_system.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxEntries = 256;   // overlapping partners of one outer body
constexpr int kBatchMaxN = 1024;   // bodies per system in a batch (LDS positions)

struct BodyView {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m, *r, *e;
    int64_t *lab;
    __device__ BodyView at(int64_t off) const {
        return BodyView{x + off, y + off, z + off, vx + off, vy + off, vz + off, ax + off, ay + off, az + off,
                        m + off, r + off, e + off, lab + off};
    }
};

struct Body {
    double x, y, z, vx, vy, vz, ax, ay, az, m, r, e;
    int64_t lab;
};

__device__ __forceinline__ Body load_body(const BodyView &B, int i) {
    return Body{B.x[i], B.y[i], B.z[i], B.vx[i], B.vy[i], B.vz[i], B.ax[i], B.ay[i], B.az[i],
                B.m[i], B.r[i], B.e[i], B.lab[i]};
}
__device__ __forceinline__ void store_body(const BodyView &B, int i, const Body &b) {
    B.x[i] = b.x; B.y[i] = b.y; B.z[i] = b.z;
    B.vx[i] = b.vx; B.vy[i] = b.vy; B.vz[i] = b.vz;
    B.ax[i] = b.ax; B.ay[i] = b.ay; B.az[i] = b.az;
    B.m[i] = b.m; B.r[i] = b.r; B.e[i] = b.e;
    B.lab[i] = b.lab;
}

__device__ double newton_cbrt(double a) {
    if (!(a > 0) || isinf(a)) return cbrt(a);
    int e;
    double f = frexp(a, &e);
    int r = ((e % 3) + 3) % 3;
    f = ldexp(f, r);
    e -= r;
    double y = 1.0;
    for (int k = 0; k < 7; ++k) {
        double y2 = y * y;
        y = y - (y2 * y - f) / (3.0 * y2);
    }
    return ldexp(y, e / 3);
}

// (m1 x1 + m2 x2)/(m1 + m2): com / vel_com / acc_com (gregbod.hpp:1094-1150)
__device__ __forceinline__ double mw(double m1, double a, double m2, double b) {
    return (m1 * a + m2 * b) / (m1 + m2);
}

// merge ctor body(one, two) (gregbod.hpp:933-940)
__device__ Body merge_new(const Body &o, const Body &t, int64_t label) {
    Body b;
    b.m = o.m + t.m;
    b.r = newton_cbrt(o.r * o.r * o.r + t.r * t.r * t.r);
    b.x = mw(o.m, o.x, t.m, t.x); b.y = mw(o.m, o.y, t.m, t.y); b.z = mw(o.m, o.z, t.m, t.z);
    b.vx = mw(o.m, o.vx, t.m, t.vx); b.vy = mw(o.m, o.vy, t.m, t.vy); b.vz = mw(o.m, o.vz, t.m, t.vz);
    b.ax = mw(o.m, o.ax, t.m, t.ax); b.ay = mw(o.m, o.ay, t.m, t.ay); b.az = mw(o.m, o.az, t.m, t.az);
    b.e = (o.e + t.e) / 2;
    b.lab = label;
    return b;
}

// body::operator+= (gregbod.hpp:1065-1082)
__device__ void merge_into(Body &s, const Body &t) {
    double x = mw(s.m, s.x, t.m, t.x), y = mw(s.m, s.y, t.m, t.y), z = mw(s.m, s.z, t.m, t.z);
    double vx = mw(s.m, s.vx, t.m, t.vx), vy = mw(s.m, s.vy, t.m, t.vy), vz = mw(s.m, s.vz, t.m, t.vz);
    double ax = mw(s.m, s.ax, t.m, t.ax), ay = mw(s.m, s.ay, t.m, t.ay), az = mw(s.m, s.az, t.m, t.az);
    s.x = x; s.y = y; s.z = z; s.vx = vx; s.vy = vy; s.vz = vz; s.ax = ax; s.ay = ay; s.az = az;
    s.e = (s.e + t.e) / 2;
    s.m += t.m;
    s.r = newton_cbrt(s.r * s.r * s.r + t.r * t.r * t.r);
}

// Geometry of one (outer o, inner j) pair exactly as s_coll computes it
// (gregsys.hpp:731-743): r12 = p_j - p_o, dist = |r12|, n = r12/dist.
struct PairGeo {
    double dist, nx, ny, nz;
    bool overlap;
};
__device__ __forceinline__ PairGeo pair_geo(const Body &o, double xj, double yj, double zj, double rj) {
    PairGeo g;
    double rad = o.r + rj;
    double dx = xj - o.x, dy = yj - o.y, dz = zj - o.z;
    g.dist = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
    g.nx = dx / g.dist;
    g.ny = dy / g.dist;
    g.nz = dz / g.dist;
    g.overlap = rad > g.dist;
    return g;
}

struct ReplayOut {
    int64_t *log;      // 4 int64 per record
    int64_t log_cap;
    int64_t *log_len;
    int *err;
};

// The whole s_coll pass for one system by one workgroup.  n, next_label and
// the merge log are updated in place.  flags: >= n bytes; when `prefiltered`
// it holds, per body i, whether a conservative test found a j > i that may
// overlap (only those bodies are tested exactly), else it is scratch.
__device__ void coll_replay(const BodyView &B, int &n_ref, bool merge, double min_mom, int64_t step,
                            int64_t &next_label, unsigned char *flags, bool prefiltered, const ReplayOut &out) {
    __shared__ int s_any, s_nent, s_outer, s_n, s_first, s_dirty;
    __shared__ double s_maxmom;
    __shared__ Body s_ob;  // the outer body (or the body it has merged into)
    __shared__ int s_ent[kMaxEntries];
    __shared__ double s_dist[kMaxEntries], s_amo[kMaxEntries], s_ami[kMaxEntries];
    __shared__ double s_nx[kMaxEntries], s_ny[kMaxEntries], s_nz[kMaxEntries];
    __shared__ int s_idx[kMaxEntries], s_keep[kMaxEntries];
    __shared__ double s_rv[kThreads];
    __shared__ int s_rj[kThreads];
    __shared__ int64_t s_outer_label;
    const int t = threadIdx.x;
    const int n0 = n_ref;
    if (t == 0) {
        s_any = 0;
        s_n = n0;
    }
    if (!prefiltered)
        for (int i = t; i < n0; i += kThreads) flags[i] = 1;
    __syncthreads();
    // exact overlap test (the reference's r_o + r_j > |p_j - p_o|) of every
    // pair i < j whose i is flagged; a squared-distance test with a relative
    // margin skips the square root for pairs that are clearly apart
    for (int i = 0; i < n0 - 1; ++i) {
        if (!flags[i]) continue;  // uniform: every thread reads the same flag
        const double xi = B.x[i], yi = B.y[i], zi = B.z[i], ri = B.r[i];
        bool hit = false;
        for (int j = i + 1 + t; j < n0; j += kThreads) {
            const double dx = B.x[j] - xi, dy = B.y[j] - yi, dz = B.z[j] - zi;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double rs = ri + B.r[j];
            if (d2 < rs * rs * (1.0 + 1e-9)) hit = hit || (rs > __builtin_sqrt(d2));
        }
        __syncthreads();  // every thread has read flags[i]
        if (t == 0) flags[i] = 0;
        __syncthreads();
        if (hit) {
            flags[i] = 1;
            s_any = 1;
        }
    }
    if (n0 > 0) {
        __syncthreads();
        if (t == 0) flags[n0 - 1] = 0;  // no j > n-1
    }
    __syncthreads();
    if (!s_any) return;

    if (t == 0) {
        s_maxmom = min_mom;
        s_outer = 0;
        s_dirty = 0;
    }
    __syncthreads();
    while (true) {
        __syncthreads();
        const int outer = s_outer;
        const int n = s_n;
        if (outer >= n) break;
        if (!s_dirty && !flags[outer]) {
            __syncthreads();
            if (t == 0) s_outer = outer + 1;
            continue;
        }
        if (t == 0) {
            s_ob = load_body(B, outer);
            s_outer_label = s_ob.lab;
            s_first = 0;
        }
        __syncthreads();
        bool rescan = false;
        while (true) {
            // ---- scan j (in parallel) ----
            if (t == 0) s_nent = 0;
            __syncthreads();
            const Body o = s_ob;
            const int cur_outer = s_outer;
            const int cur_n = s_n;
            double best = -INFINITY;
            int bestj = -1;
            for (int j = (rescan ? 0 : cur_outer + 1) + t; j < cur_n; j += kThreads) {
                if (j == cur_outer) continue;
                PairGeo g = pair_geo(o, B.x[j], B.y[j], B.z[j], B.r[j]);
                if (!g.overlap) continue;
                if (merge) {
                    const double mj = B.m[j];
                    const double pox = o.m * o.vx, poy = o.m * o.vy, poz = o.m * o.vz;
                    const double pjx = mj * B.vx[j], pjy = mj * B.vy[j], pjz = mj * B.vz[j];
                    const double amo = pox * g.nx + poy * g.ny + poz * g.nz;
                    const double ami = pjx * g.nx + pjy * g.ny + pjz * g.nz;
                    const double ms = o.m + mj;
                    const double vcx = (pox + pjx) / ms, vcy = (poy + pjy) / ms, vcz = (poz + pjz) / ms;
                    const double curr = amo - ami - (o.m - mj) * (vcx * g.nx + vcy * g.ny + vcz * g.nz);
                    if (curr >= best) {
                        best = curr;
                        bestj = j;
                    }
                }
                int k = atomicAdd(&s_nent, 1);
                if (k < kMaxEntries) s_ent[k] = j;
                else atomicExch(out.err, 1);
            }
            s_rv[t] = best;
            s_rj[t] = bestj;
            __syncthreads();
            for (int off = kThreads / 2; off > 0; off >>= 1) {
                if (t < off) {
                    double v2 = s_rv[t + off];
                    int j2 = s_rj[t + off];
                    if (j2 >= 0 && (s_rj[t] < 0 || v2 > s_rv[t] || (v2 == s_rv[t] && j2 > s_rj[t]))) {
                        s_rv[t] = v2;
                        s_rj[t] = j2;
                    }
                }
                __syncthreads();
            }
            const bool do_merge = merge && s_rj[0] >= 0 && s_rv[0] >= s_maxmom;
            if (!do_merge) break;
            // ---- merge outer with the candidate (gregsys.hpp:766-808) ----
            const int cand = s_rj[0];
            __syncthreads();
            if (t == 0) {
                s_maxmom = s_rv[0];
                Body c = load_body(B, cand);
                if (!s_first) {
                    s_ob = merge_new(s_ob, c, next_label++);
                    s_first = 1;
                } else {
                    merge_into(s_ob, c);
                }
                int64_t k = *out.log_len;
                if (k < out.log_cap) {
                    out.log[4 * k + 0] = step;
                    out.log[4 * k + 1] = s_outer_label;
                    out.log[4 * k + 2] = c.lab;
                    out.log[4 * k + 3] = s_ob.lab;
                }
                *out.log_len = k + 1;
            }
            // erase cand: shift [cand+1, n) down by one, chunk by chunk
            for (int base = cand; base < cur_n - 1; base += kThreads) {
                const int i = base + t;
                Body tmp;
                const bool mv = i < cur_n - 1;
                if (mv) tmp = load_body(B, i + 1);
                __syncthreads();
                if (mv) store_body(B, i, tmp);
                __syncthreads();
            }
            if (t == 0) {
                if (cand < cur_outer) s_outer = cur_outer - 1;
                s_n = cur_n - 1;
                s_dirty = 1;
            }
            rescan = true;
            __syncthreads();
        }
        // ---- bounce entries in (distance, j) order (gregsys.hpp:810-837) ----
        if (t == 0) {
            const int ne = min(s_nent, kMaxEntries);
            Body &o = s_ob;
            // capture every entry at scan state, then sort
            double *dist = s_dist, *amo = s_amo, *ami = s_ami, *nx = s_nx, *ny = s_ny, *nz = s_nz;
            int *idx = s_idx, *keep = s_keep;
            for (int k = 0; k < ne; ++k) idx[k] = s_ent[k];
            for (int k = 0; k < ne; ++k) {
                PairGeo g = pair_geo(o, B.x[idx[k]], B.y[idx[k]], B.z[idx[k]], B.r[idx[k]]);
                dist[k] = g.dist;
            }
            for (int a = 1; a < ne; ++a) {  // insertion sort by (dist, j)
                double dv = dist[a];
                int jv = idx[a];
                int b = a - 1;
                while (b >= 0 && (dist[b] > dv || (dist[b] == dv && idx[b] > jv))) {
                    dist[b + 1] = dist[b];
                    idx[b + 1] = idx[b];
                    --b;
                }
                dist[b + 1] = dv;
                idx[b + 1] = jv;
            }
            // axis momenta at scan time, before any entry is applied
            for (int k = 0; k < ne; ++k) {
                const int j = idx[k];
                keep[k] = !(k > 0 && dist[k] == dist[k - 1]);  // equal keys: the first emplace wins
                PairGeo g = pair_geo(o, B.x[j], B.y[j], B.z[j], B.r[j]);
                nx[k] = g.nx; ny[k] = g.ny; nz[k] = g.nz;
                const double mj = B.m[j];
                amo[k] = (o.m * o.vx) * g.nx + (o.m * o.vy) * g.ny + (o.m * o.vz) * g.nz;
                ami[k] = (mj * B.vx[j]) * g.nx + (mj * B.vy[j]) * g.ny + (mj * B.vz[j]) * g.nz;
            }
            for (int k = 0; k < ne; ++k) {
                if (!keep[k]) continue;
                const int j = idx[k];
                const double mj = B.m[j];
                const double o_vel = amo[k] / o.m;
                const double i_vel = ami[k] / mj;
                const double o_minus_i = o_vel - i_vel;
                const double avg_rest = (o.e + B.e[j]) / 2;
                if (o_vel > 0 || i_vel < 0) {
                    if (o_vel <= 0 && i_vel < 0)
                        if (o_vel < i_vel) continue;
                    if (o_vel > 0 && i_vel >= 0)
                        if (o_vel < i_vel) continue;
                    const double tot = amo[k] + ami[k];
                    const double ms = o.m + mj;
                    const double new_o = (tot - mj * avg_rest * o_minus_i) / ms;
                    const double new_i = (tot + o.m * avg_rest * o_minus_i) / ms;
                    const double fo = new_o - o_vel, fi = new_i - i_vel;
                    o.vx += fo * nx[k]; o.vy += fo * ny[k]; o.vz += fo * nz[k];
                    B.vx[j] += fi * nx[k]; B.vy[j] += fi * ny[k]; B.vz[j] += fi * nz[k];
                }
            }
            if (s_first) store_body(B, s_outer, o);
            else {
                B.vx[s_outer] = o.vx;
                B.vy[s_outer] = o.vy;
                B.vz[s_outer] = o.vz;
            }
            s_outer = s_outer + 1;
        }
        __syncthreads();
    }
    __syncthreads();
    if (t == 0) n_ref = s_n;
    __syncthreads();
}

// ---- one system: collision pass after a step ----
__global__ __launch_bounds__(kThreads) void k_coll_single(BodyView B, int *n_io, int merge, double min_mom,
                                                          int64_t step, int64_t *next_label, unsigned char *flags,
                                                          int64_t *log, int64_t log_cap, int64_t *log_len,
                                                          int *err) {
    __shared__ int s_nb;
    __shared__ int64_t s_next;
    if (threadIdx.x == 0) {
        s_nb = *n_io;
        s_next = *next_label;
    }
    __syncthreads();
    ReplayOut out{log, log_cap, log_len, err};
    coll_replay(B, s_nb, merge != 0, min_mom, step, s_next, flags, false, out);
    if (threadIdx.x == 0) {
        *n_io = s_nb;
        *next_label = s_next;
    }
}

// ---- batched ensembles: one workgroup per system, many steps per launch ----
struct BatchParams {
    double G, dt, half_dt, eps2;
    int coll;        // 0 or 1 (overlap check)
    int merge;
    double min_mom;
    int mixed;       // FP32 pair forces
};

__device__ __forceinline__ void kahan_add(float &s, float &c, float v) {
    float y = v - c;
    float t2 = s + y;
    c = (t2 - s) - y;
    s = t2;
}

__global__ __launch_bounds__(kThreads) void k_batch_steps(BodyView all, const int *__restrict__ order,
                                                          const int64_t *__restrict__ offsets, int *counts,
                                                          int64_t *next_labels, unsigned char *flags, int64_t *log,
                                                          int64_t *log_lens, int *err, BatchParams P,
                                                          int64_t step0, int64_t nsteps) {
    __shared__ float4 s_pos[kBatchMaxN];
    __shared__ float s_R[kBatchMaxN];
    __shared__ int s_nb;
    __shared__ int64_t s_next;
    const int sys = order[blockIdx.x];
    const int64_t off = offsets[sys];
    const BodyView B = all.at(off);
    const int t = threadIdx.x;
    if (t == 0) {
        s_nb = counts[sys];
        s_next = next_labels[sys];
    }
    __syncthreads();
    ReplayOut out{log + 4 * off, offsets[sys + 1] - off, log_lens + sys, err};
    for (int64_t st = 1; st <= nsteps; ++st) {
        const int n = s_nb;
        if (n == 0) break;
        // kick half, drift (evolve :2950-2955)
        for (int i = t; i < n; i += kThreads) {
            double vx = B.vx[i] + P.half_dt * B.ax[i];
            double vy = B.vy[i] + P.half_dt * B.ay[i];
            double vz = B.vz[i] + P.half_dt * B.az[i];
            B.vx[i] = vx; B.vy[i] = vy; B.vz[i] = vz;
            B.x[i] += P.dt * vx; B.y[i] += P.dt * vy; B.z[i] += P.dt * vz;
        }
        __syncthreads();
        if (P.mixed) {
            // FP32 positions relative to the last body (a nearby fragment).
            // s_R[i] = r_i (1 + 2^-18) + 2^-21 max|x'_i| bounds the radius plus
            // the FP32 position error, so d2 < (R_i + R_j)^2 is a conservative
            // overlap prefilter computed alongside the force.
            const double rx = B.x[n - 1], ry = B.y[n - 1], rz = B.z[n - 1];
            for (int i = t; i < n; i += kThreads) {
                const float4 p = make_float4((float)(B.x[i] - rx), (float)(B.y[i] - ry), (float)(B.z[i] - rz),
                                             (float)(P.G * B.m[i]));
                s_pos[i] = p;
                const float mag = fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z)));
                s_R[i] = (float)B.r[i] * (1.0f + 0x1.0p-18f) + 0x1.0p-21f * mag;
            }
            __syncthreads();
            const float eps2 = (float)P.eps2;
            for (int i = t; i < n; i += kThreads) {
                const float4 pi = s_pos[i];
                const float Ri = s_R[i];
                float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
                int hits = 0;
                for (int jt = 0; jt < n; jt += 64) {
                    const int je = min(n, jt + 64);
                    float px = 0.f, py = 0.f, pz = 0.f;
                    for (int j = jt; j < je; ++j) {
                        const float4 pj = s_pos[j];
                        const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        const float rs = Ri + s_R[j];
                        hits += d2 < rs * rs ? 1 : 0;  // counts the self pair once
                        const float r2 = d2 + eps2;
                        float inv = __builtin_amdgcn_rsqf(r2);
                        inv = r2 > 0.f ? inv : 0.f;
                        const float w = pj.w * inv * (inv * inv);
                        px = __builtin_fmaf(w, dx, px);
                        py = __builtin_fmaf(w, dy, py);
                        pz = __builtin_fmaf(w, dz, pz);
                    }
                    kahan_add(sx, cx, px);
                    kahan_add(sy, cy, py);
                    kahan_add(sz, cz, pz);
                }
                B.ax[i] = (double)sx - (double)cx;
                B.ay[i] = (double)sy - (double)cy;
                B.az[i] = (double)sz - (double)cz;
                flags[off + i] = hits > 1 ? 1 : 0;
            }
        } else {
            // FP64, reference order: j ascending (cumulative_acc :352-357 or bh_acc's softened form)
            for (int i = t; i < n; i += kThreads) {
                const double xi = B.x[i], yi = B.y[i], zi = B.z[i];
                double ax = 0.0, ay = 0.0, az = 0.0;
                for (int j = 0; j < n; ++j) {
                    if (j == i) continue;
                    const double dx = B.x[j] - xi, dy = B.y[j] - yi, dz = B.z[j] - zi;
                    if (P.eps2 == 0.0) {
                        const double s = dx * dx + dy * dy + dz * dz;
                        const double wx = s * dx, wy = s * dy, wz = s * dz;
                        const double r3 = __builtin_sqrt(wx * wx + wy * wy + wz * wz);
                        const double k = (P.G * B.m[j]) / r3;
                        ax += k * dx; ay += k * dy; az += k * dz;
                    } else {
                        const double s = dx * dx + dy * dy + dz * dz + P.eps2;
                        const double sq = __builtin_sqrt(s);
                        const double k = P.G * (B.m[j] / s);
                        ax += k * (dx / sq); ay += k * (dy / sq); az += k * (dz / sq);
                    }
                }
                B.ax[i] = ax; B.ay[i] = ay; B.az[i] = az;
            }
        }
        __syncthreads();
        // closing half kick (loop_kdk :499)
        for (int i = t; i < n; i += kThreads) {
            B.vx[i] += P.half_dt * B.ax[i];
            B.vy[i] += P.half_dt * B.ay[i];
            B.vz[i] += P.half_dt * B.az[i];
        }
        __syncthreads();
        if (P.coll)
            coll_replay(B, s_nb, P.merge != 0, P.min_mom, step0 + st, s_next, flags + off, P.mixed != 0, out);
        __syncthreads();
    }
    if (t == 0) {
        counts[sys] = s_nb;
        next_labels[sys] = s_next;
    }
}

inline BodyView view_of(nb_system *s, int64_t *lab) {
    const int64_t n = s->n;
    return BodyView{s->x.p, s->x.p + n, s->x.p + 2 * n, s->v.p, s->v.p + n, s->v.p + 2 * n,
                    s->a.p, s->a.p + n, s->a.p + 2 * n, s->m.p, s->r.p, s->e.p, lab};
}

}  // namespace

struct CollWork {
    DBuf<int64_t> lab;
    DBuf<unsigned char> flags;
    DBuf<int64_t> log, loglen, next;
    DBuf<int> n, err;
    int64_t log_cap = 0;
    int64_t synced_log = 0;
};

void coll_work_free(CollWork *w) { delete w; }

// s_coll after a step's closing kick.  Labels live on the device while
// collisions are on; the body count and new merge records come back here.
void collision_pass(nb_system *s, uint64_t step) {
    if (!s->coll) s->coll = new CollWork();
    CollWork &w = *s->coll;
    const int64_t n = s->n;
    hipStream_t st = s->stream;
    // labels and the next label travel with every pass (the host copies are authoritative)
    w.lab.ensure((size_t)std::max<int64_t>(n, 1));
    w.loglen.ensure(1);
    w.next.ensure(1);
    if (n) NB_HIP(hipMemcpyAsync(w.lab.p, s->labels.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    NB_HIP(hipMemcpyAsync(w.next.p, &s->next_label, 8, hipMemcpyHostToDevice, st));
    w.flags.ensure((size_t)std::max<int64_t>(n, 1));
    w.log_cap = std::max<int64_t>(n, 1);
    w.log.ensure(4 * (size_t)w.log_cap);
    w.n.ensure(1);
    w.err.ensure(1);
    int nn = (int)n;
    int zero_i = 0;
    int64_t zero = 0;
    NB_HIP(hipMemcpyAsync(w.n.p, &nn, sizeof(int), hipMemcpyHostToDevice, st));
    NB_HIP(hipMemcpyAsync(w.err.p, &zero_i, sizeof(int), hipMemcpyHostToDevice, st));
    NB_HIP(hipMemcpyAsync(w.loglen.p, &zero, 8, hipMemcpyHostToDevice, st));
    BodyView B = view_of(s, w.lab.p);
    {
        ScopedTimer tm(s->prof, "collide", st);
        hipLaunchKernelGGL(k_coll_single, dim3(1), dim3(kThreads), 0, st, B, w.n.p, s->mergers ? 1 : 0,
                           s->min_tot_com_mom, (int64_t)step, w.next.p, w.flags.p, w.log.p, w.log_cap, w.loglen.p,
                           w.err.p);
    }
    NB_HIP(hipGetLastError());
    int n_out = 0, errv = 0;
    int64_t nlog = 0, next = 0;
    NB_HIP(hipMemcpyAsync(&n_out, w.n.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(&errv, w.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(&nlog, w.loglen.p, 8, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(&next, w.next.p, 8, hipMemcpyDeviceToHost, st));
    NB_HIP(hipStreamSynchronize(st));
    if (errv) throw ApiError(NB_ERR_CAPACITY, "collision pass: more than 256 overlapping partners for one body");
    if (nlog > 0) {
        std::vector<int64_t> rec(4 * (size_t)nlog);
        NB_HIP(hipMemcpy(rec.data(), w.log.p, rec.size() * 8, hipMemcpyDeviceToHost));
        s->merge_log.insert(s->merge_log.end(), rec.begin(), rec.end());
        w.synced_log = (int64_t)(s->merge_log.size() / 4);
    }
    s->next_label = next;
    if (n_out != n) {
        // the SoA blocks are laid out with stride n: repack to the new count
        const int64_t m2 = n_out;
        std::vector<double> buf;
        for (DBuf<double> *d : {&s->x, &s->v, &s->a}) {
            buf.resize(3 * (size_t)n);
            NB_HIP(hipMemcpy(buf.data(), d->p, buf.size() * 8, hipMemcpyDeviceToHost));
            std::vector<double> packed(3 * (size_t)m2);
            for (int c = 0; c < 3; ++c)
                std::copy(buf.begin() + (size_t)c * n, buf.begin() + (size_t)c * n + m2,
                          packed.begin() + (size_t)c * m2);
            NB_HIP(hipMemcpy(d->p, packed.data(), packed.size() * 8, hipMemcpyHostToDevice));
        }
        s->n = m2;
    }
    s->labels.resize((size_t)s->n);
    if (s->n) NB_HIP(hipMemcpy(s->labels.data(), w.lab.p, (size_t)s->n * 8, hipMemcpyDeviceToHost));
}

}  // namespace nbe

// ---------------- batched ensembles (C ABI) ----------------
using namespace nbe;

struct nb_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t nsys = 0, total = 0;
    std::vector<int64_t> offsets;
    std::vector<int> order;
    DBuf<double> x, v, a, m, r, e;
    DBuf<int64_t> lab, off, next, log, loglen;
    DBuf<int> cnt, ord, errf;
    DBuf<unsigned char> flags;
    BatchParams P{6.6743e-11, 1.0, 0.5, 0.0, 3, 1, 1e10, 1};
    int64_t step = 0;
    ~nb_batch() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
thread_local std::string g_batch_err;

int bfail(nb_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    else g_batch_err = msg;
    return code;
}

template <typename F>
int bguard(nb_batch *b, F &&f) {
    try {
        if (b) NB_HIP(hipSetDevice(b->device));
        return f();
    } catch (const ApiError &e) {
        return bfail(b, e.code, e.what);
    } catch (const DeviceError &e) {
        return bfail(b, NB_ERR_DEVICE, e.what);
    } catch (const std::bad_alloc &) {
        return bfail(b, NB_ERR_CAPACITY, "host allocation failed");
    }
}
}  // namespace

extern "C" {

int nb_batch_create(int device, int64_t nsys, const int64_t *offsets, const double *m, const double *r,
                    const double *pos, const double *vel, const double *rest_c, nb_batch **out) {
    if (!out) return bfail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_batch_create: out is NULL");
    *out = nullptr;
    return bguard(nullptr, [&]() {
        if (nsys <= 0 || !offsets || !m || !pos || !vel)
            throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_batch_create: bad arguments");
        int count = 0;
        NB_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw ApiError(NB_ERR_INVALID_ARGUMENT, "device index out of range");
        NB_HIP(hipSetDevice(device));
        const int64_t total = offsets[nsys];
        if (offsets[0] != 0 || total < 0) throw ApiError(NB_ERR_INVALID_ARGUMENT, "offsets must start at 0");
        for (int64_t s = 0; s < nsys; ++s) {
            int64_t k = offsets[s + 1] - offsets[s];
            if (k < 0) throw ApiError(NB_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
            if (k > kBatchMaxN) throw ApiError(NB_ERR_CAPACITY, "a batched system holds at most 1024 bodies");
        }
        for (int64_t i = 0; i < total; ++i) {
            if (m[i] < 0) throw ApiError(NB_ERR_NEGATIVE_MASS, "negative mass");
            if (r && r[i] < 0) throw ApiError(NB_ERR_NEGATIVE_RADIUS, "negative radius");
        }
        nb_batch *b = new nb_batch();
        b->device = device;
        hipError_t he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (he != hipSuccess) {
            delete b;
            throw DeviceError(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(he));
        }
        b->nsys = nsys;
        b->total = total;
        b->offsets.assign(offsets, offsets + nsys + 1);
        const size_t T = (size_t)std::max<int64_t>(total, 1);
        b->x.ensure(3 * T); b->v.ensure(3 * T); b->a.ensure(3 * T);
        b->m.ensure(T); b->r.ensure(T); b->e.ensure(T); b->lab.ensure(T); b->log.ensure(4 * T);
        b->flags.ensure(T);
        b->off.ensure((size_t)nsys + 1); b->next.ensure((size_t)nsys); b->loglen.ensure((size_t)nsys);
        b->cnt.ensure((size_t)nsys); b->ord.ensure((size_t)nsys); b->errf.ensure(1);
        std::vector<double> tmp(3 * T, 0.0);
        auto up3 = [&](DBuf<double> &d, const double *src) {
            for (int64_t i = 0; i < total; ++i)
                for (int c = 0; c < 3; ++c) tmp[(size_t)c * T + (size_t)i] = src ? src[3 * i + c] : 0.0;
            NB_HIP(hipMemcpy(d.p, tmp.data(), 3 * T * 8, hipMemcpyHostToDevice));
        };
        up3(b->x, pos);
        up3(b->v, vel);
        up3(b->a, nullptr);
        std::vector<double> one(T, 1.0);
        NB_HIP(hipMemcpy(b->m.p, m, (size_t)total * 8, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->r.p, r ? r : one.data(), (size_t)total * 8, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->e.p, rest_c ? rest_c : one.data(), (size_t)total * 8, hipMemcpyHostToDevice));
        std::vector<int64_t> lab(T), next((size_t)nsys), zeros((size_t)nsys, 0);
        std::vector<int> cnt((size_t)nsys);
        for (int64_t s = 0; s < nsys; ++s) {
            int64_t k = offsets[s + 1] - offsets[s];
            cnt[(size_t)s] = (int)k;
            next[(size_t)s] = k;
            for (int64_t q = 0; q < k; ++q) lab[(size_t)(offsets[s] + q)] = q;
        }
        // longest systems first (n^2 work): the dispatcher hands out workgroups in order
        b->order.resize((size_t)nsys);
        std::iota(b->order.begin(), b->order.end(), 0);
        std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return cnt[(size_t)a] > cnt[(size_t)c]; });
        NB_HIP(hipMemcpy(b->lab.p, lab.data(), (size_t)total * 8, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->off.p, offsets, (size_t)(nsys + 1) * 8, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->next.p, next.data(), (size_t)nsys * 8, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->loglen.p, zeros.data(), (size_t)nsys * 8, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->cnt.p, cnt.data(), (size_t)nsys * 4, hipMemcpyHostToDevice));
        NB_HIP(hipMemcpy(b->ord.p, b->order.data(), (size_t)nsys * 4, hipMemcpyHostToDevice));
        NB_HIP(hipMemset(b->errf.p, 0, sizeof(int)));
        *out = b;
        return NB_OK;
    });
}

int nb_batch_destroy(nb_batch *b) {
    if (!b) return NB_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return NB_OK;
}

const char *nb_batch_last_error(const nb_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int nb_batch_set_params(nb_batch *b, double G, double dt, double eps, int mode, int mergers, double min_mom,
                        int precision) {
    if (!b) return bfail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL batch");
    if (!(G > 0) || !(dt > 0) || eps < 0 || min_mom < 0)
        return bfail(b, NB_ERR_INVALID_ARGUMENT, "G and dt must be positive, eps and min_tot_com_mom non-negative");
    if (mode != NB_NO_COLL_CHECK && mode != NB_OVERLAP_COLL_CHECK)
        return bfail(b, NB_ERR_NOT_SUPPORTED, "only collisions = 0 or 3 (overlap check) are supported");
    if (precision != NB_PREC_FP64 && precision != NB_PREC_MIXED)
        return bfail(b, NB_ERR_INVALID_ARGUMENT, "batched systems run NB_PREC_FP64 or NB_PREC_MIXED");
    b->P = BatchParams{G, dt, dt / 2, eps * eps, mode == NB_OVERLAP_COLL_CHECK ? 1 : 0, mergers ? 1 : 0, min_mom,
                       precision == NB_PREC_MIXED ? 1 : 0};
    return NB_OK;
}

int nb_batch_evolve(nb_batch *b, uint64_t steps, int64_t *elapsed_ns) {
    if (!b) return bfail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL batch");
    if (elapsed_ns) *elapsed_ns = 0;
    return bguard(b, [&]() {
        auto t0 = std::chrono::high_resolution_clock::now();
        // initial accelerations: one launch with zero steps would not compute
        // them, so the first step's opening kick uses a from a force-only pass
        const BodyView all{b->x.p, b->x.p + b->total, b->x.p + 2 * b->total, b->v.p, b->v.p + b->total,
                           b->v.p + 2 * b->total, b->a.p, b->a.p + b->total, b->a.p + 2 * b->total,
                           b->m.p, b->r.p, b->e.p, b->lab.p};
        // evolve semantics: a(x0) first (calc_acc_and_e), then `steps` steps
        BatchParams P0 = b->P;
        {
            // a(x0): run the force part only by a zero-dt step without collisions
            BatchParams Pf = P0;
            Pf.dt = 0.0;
            Pf.half_dt = 0.0;
            Pf.coll = 0;
            hipLaunchKernelGGL(k_batch_steps, dim3((unsigned)b->nsys), dim3(kThreads), 0, b->stream, all, b->ord.p,
                               b->off.p, b->cnt.p, b->next.p, b->flags.p, b->log.p, b->loglen.p, b->errf.p, Pf,
                               b->step, (int64_t)1);
        }
        const int64_t chunk = 1000;  // bound the duration of one launch
        for (uint64_t done = 0; done < steps;) {
            const int64_t k = (int64_t)std::min<uint64_t>(chunk, steps - done);
            hipLaunchKernelGGL(k_batch_steps, dim3((unsigned)b->nsys), dim3(kThreads), 0, b->stream, all, b->ord.p,
                               b->off.p, b->cnt.p, b->next.p, b->flags.p, b->log.p, b->loglen.p, b->errf.p, P0,
                               b->step, k);
            NB_HIP(hipGetLastError());
            b->step += k;
            done += (uint64_t)k;
        }
        int errv = 0;
        NB_HIP(hipMemcpyAsync(&errv, b->errf.p, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        NB_HIP(hipStreamSynchronize(b->stream));
        if (errv) throw ApiError(NB_ERR_CAPACITY, "collision pass: more than 256 overlapping partners for one body");
        auto t1 = std::chrono::high_resolution_clock::now();
        if (elapsed_ns) *elapsed_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        return NB_OK;
    });
}

int nb_batch_get_state(const nb_batch *cb, int64_t cap, int64_t *counts, double *m, double *r, double *pos,
                       double *vel, int64_t *ids) {
    nb_batch *b = const_cast<nb_batch *>(cb);
    if (!b || !counts) return bfail(b, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return bguard(b, [&]() {
        NB_HIP(hipStreamSynchronize(b->stream));
        const int64_t T = b->total;
        std::vector<int> cnt((size_t)b->nsys);
        NB_HIP(hipMemcpy(cnt.data(), b->cnt.p, cnt.size() * 4, hipMemcpyDeviceToHost));
        int64_t live = 0;
        for (int c : cnt) live += c;
        if (cap < live) throw ApiError(NB_ERR_CAPACITY, "nb_batch_get_state: capacity too small");
        std::vector<double> hx(3 * (size_t)T), hv(3 * (size_t)T), hm((size_t)T), hr((size_t)T);
        std::vector<int64_t> hl((size_t)T);
        NB_HIP(hipMemcpy(hx.data(), b->x.p, hx.size() * 8, hipMemcpyDeviceToHost));
        NB_HIP(hipMemcpy(hv.data(), b->v.p, hv.size() * 8, hipMemcpyDeviceToHost));
        NB_HIP(hipMemcpy(hm.data(), b->m.p, hm.size() * 8, hipMemcpyDeviceToHost));
        NB_HIP(hipMemcpy(hr.data(), b->r.p, hr.size() * 8, hipMemcpyDeviceToHost));
        NB_HIP(hipMemcpy(hl.data(), b->lab.p, hl.size() * 8, hipMemcpyDeviceToHost));
        int64_t k = 0;
        for (int64_t s = 0; s < b->nsys; ++s) {
            counts[s] = cnt[(size_t)s];
            for (int q = 0; q < cnt[(size_t)s]; ++q, ++k) {
                const size_t i = (size_t)(b->offsets[(size_t)s] + q);
                if (m) m[k] = hm[i];
                if (r) r[k] = hr[i];
                for (int c = 0; c < 3; ++c) {
                    if (pos) pos[3 * k + c] = hx[(size_t)c * T + i];
                    if (vel) vel[3 * k + c] = hv[(size_t)c * T + i];
                }
                if (ids) ids[k] = hl[i];
            }
        }
        return NB_OK;
    });
}

int nb_batch_merge_log(const nb_batch *cb, int64_t cap, int64_t *records, int64_t *count) {
    nb_batch *b = const_cast<nb_batch *>(cb);
    if (!b || !count) return bfail(b, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return bguard(b, [&]() {
        NB_HIP(hipStreamSynchronize(b->stream));
        std::vector<int64_t> lens((size_t)b->nsys), lg(4 * (size_t)std::max<int64_t>(b->total, 1));
        NB_HIP(hipMemcpy(lens.data(), b->loglen.p, lens.size() * 8, hipMemcpyDeviceToHost));
        NB_HIP(hipMemcpy(lg.data(), b->log.p, lg.size() * 8, hipMemcpyDeviceToHost));
        int64_t tot = 0;
        for (auto l : lens) tot += l;
        *count = tot;
        if (!records) return NB_OK;
        if (cap < tot) throw ApiError(NB_ERR_CAPACITY, "nb_batch_merge_log: capacity too small");
        int64_t k = 0;
        for (int64_t s = 0; s < b->nsys; ++s)
            for (int64_t q = 0; q < lens[(size_t)s]; ++q, ++k) {
                const size_t base = 4 * (size_t)(b->offsets[(size_t)s] + q);
                records[5 * k + 0] = s;
                for (int c = 0; c < 4; ++c) records[5 * k + 1 + c] = lg[base + c];
            }
        return NB_OK;
    });
}

}  // extern "C"
