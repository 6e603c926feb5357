// nb_coll_dev.hpp -- device code of the collision pass and of the batched
// ensemble kernel (see nb_coll.hip for the reference path and the design).
//
// Kept in a header so that the same source compiles in two places: inside
// nb_coll.hip for gfx950 (HIP), and in a CPU build that runs each workgroup's
// threads as cooperatively scheduled contexts with bounds-checked buffers
// (tests/emu/), which is how the kernel logic is checked without a GPU.  The
// includer provides the HIP vocabulary (threadIdx, __syncthreads,
// __syncthreads_or, atomicAdd, float4, __shared__ ...) and opens the namespace.
// Arithmetic is FP64 with contraction off (the includer sets it).
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
