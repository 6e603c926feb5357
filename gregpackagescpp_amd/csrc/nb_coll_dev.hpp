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

// Element i of a body array, addressed as the array base plus a 32-bit byte
// offset: the loads and stores then take the uniform base in SGPRs and one
// VGPR of offset for every array, instead of a 64-bit address per array.
// Bodies per system stay below 2^28: collision_pass checks it, and a batched
// system holds at most kBatchMaxN.
template <class T>
__device__ __forceinline__ T &el(T *base, int i) {
    return *(T *)((char *)base + (unsigned)i * (unsigned)sizeof(T));
}

struct Body {
    double x, y, z, vx, vy, vz, ax, ay, az, m, r, e;
    int64_t lab;
};

__device__ __forceinline__ Body load_body(const BodyView &B, int i) {
    return Body{el(B.x, i), el(B.y, i), el(B.z, i), el(B.vx, i), el(B.vy, i), el(B.vz, i), el(B.ax, i), el(B.ay, i), el(B.az, i),
                el(B.m, i), el(B.r, i), el(B.e, i), el(B.lab, i)};
}
__device__ __forceinline__ void store_body(const BodyView &B, int i, const Body &b) {
    el(B.x, i) = b.x; el(B.y, i) = b.y; el(B.z, i) = b.z;
    el(B.vx, i) = b.vx; el(B.vy, i) = b.vy; el(B.vz, i) = b.vz;
    el(B.ax, i) = b.ax; el(B.ay, i) = b.ay; el(B.az, i) = b.az;
    el(B.m, i) = b.m; el(B.r, i) = b.r; el(B.e, i) = b.e;
    el(B.lab, i) = b.lab;
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

struct ReplayOut {
    int64_t *log;      // 4 int64 per record
    int64_t log_cap;
    int64_t *log_len;
    int *err;
};

// FP64 body field f of B: 0-2 position, 3-5 velocity, 6-8 acceleration,
// 9 mass, 10 radius, 11 restitution.
__device__ __forceinline__ double *body_field(const BodyView &B, int f) {
    switch (f) {
        case 0: return B.x;
        case 1: return B.y;
        case 2: return B.z;
        case 3: return B.vx;
        case 4: return B.vy;
        case 5: return B.vz;
        case 6: return B.ax;
        case 7: return B.ay;
        case 8: return B.az;
        case 9: return B.m;
        case 10: return B.r;
        default: return B.e;
    }
}

// First i in [from, n) with flag[i] set, else n.  Called by every thread of
// the workgroup with the same arguments; each round looks at 8 flags per
// thread and takes the smallest hit.
__device__ int next_flagged(const unsigned char *flag, int from, int n) {
    __shared__ int s_first_hit;
    for (int base = from; base < n; base += kThreads * 8) {
        if (threadIdx.x == 0) s_first_hit = n;
        __syncthreads();
        const int lo = base + (int)threadIdx.x * 8;
        const int hi = min(lo + 8, n);
        for (int q = lo; q < hi; ++q)
            if (flag[q]) {
                atomicMin(&s_first_hit, q);
                break;
            }
        __syncthreads();
        const int f = s_first_hit;
        __syncthreads();  // every thread has read it before the next round resets it
        if (f < n) return f;
    }
    return n;
}

// The whole s_coll pass for one system by one workgroup.  n, next_label and
// the merge log are updated in place.
//   pref:  NULL, or per body whether a conservative test found a j that may
//          overlap it (only those bodies get the exact test);
//   exact: >= n bytes of scratch: per body whether some j > i overlaps it.
// The outer bodies are visited in index order; a body with no overlapping
// j > i changes nothing until the first merge, after which every later outer
// body is scanned (the rescan moves bodies, so the exact flags go stale).
__device__ void coll_replay(const BodyView &B, int &n_ref, bool merge, double min_mom, int64_t step,
                            int64_t &next_label, const unsigned char *pref, unsigned char *exact,
                            const ReplayOut &out) {
    __shared__ int s_n, s_nent, s_dirty, s_first;
    __shared__ double s_maxmom;
    __shared__ double s_ob[12];  // the outer body (or the body it has merged into), fields as body_field
    __shared__ int64_t s_ob_lab, s_outer_label;
    __shared__ int s_ent[kMaxEntries];
    __shared__ double s_dist[kMaxEntries];
    __shared__ double s_rv[kThreads];
    __shared__ int s_rj[kThreads];
    const int t = threadIdx.x;
    const int n0 = n_ref;

    // 1. exact overlap test (the reference's r_o + r_j > |p_j - p_o|) of every
    // pair i < j whose i the prefilter marked; a squared-distance test with a
    // relative margin skips the square root for pairs that are clearly apart
    for (int i = t; i < n0; i += kThreads) exact[i] = 0;
    __syncthreads();
    int found = 0;
    for (int i = 0; i < n0 - 1; ++i) {
        if (pref && !pref[i]) continue;  // uniform: every thread reads the same flag
        const double xi = el(B.x, i), yi = el(B.y, i), zi = el(B.z, i), ri = el(B.r, i);
        bool hit = false;
        for (int j = i + 1 + t; j < n0; j += kThreads) {
            const double dx = el(B.x, j) - xi, dy = el(B.y, j) - yi, dz = el(B.z, j) - zi;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double rs = ri + el(B.r, j);
            if (d2 < rs * rs * (1.0 + 1e-9)) hit = hit || (rs > __builtin_sqrt(d2));
        }
        if (hit) {
            exact[i] = 1;  // every thread that writes writes 1
            found = 1;
        }
    }
    if (!__syncthreads_or(found)) return;

    // 2. outer bodies in order (gregsys.hpp:700-851)
    if (t == 0) {
        s_n = n0;
        s_dirty = 0;
        s_maxmom = min_mom;
    }
    __syncthreads();
    int outer = 0;  // the same in every thread
    while (true) {
        const int n = s_n;
        if (!s_dirty) outer = next_flagged(exact, outer, n);
        if (outer >= n) break;
        if (t < 12) s_ob[t] = body_field(B, t)[outer];
        if (t == 12) {
            s_ob_lab = el(B.lab, outer);
            s_outer_label = s_ob_lab;
            s_first = 0;
        }
        int cur = outer;  // the outer body's index: moves down when a body before it is erased
        bool rescan = false;
        while (true) {
            // ---- scan j in parallel: overlaps and the merge candidate; the
            // pair geometry as s_coll computes it (gregsys.hpp:731-743):
            // r12 = p_j - p_o, dist = |r12|, n = r12/dist ----
            if (t == 0) s_nent = 0;
            __syncthreads();
            const int cur_n = s_n;
            const double ox = s_ob[0], oy = s_ob[1], oz = s_ob[2];
            const double ovx = s_ob[3], ovy = s_ob[4], ovz = s_ob[5];
            const double om = s_ob[9], orad = s_ob[10];
            double best = -INFINITY;
            int bestj = -1;
            for (int j = (rescan ? 0 : cur + 1) + t; j < cur_n; j += kThreads) {
                if (j == cur) continue;
                const double rad = orad + el(B.r, j);
                const double dx = el(B.x, j) - ox, dy = el(B.y, j) - oy, dz = el(B.z, j) - oz;
                const double dist = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
                if (!(rad > dist)) continue;
                if (merge) {
                    const double nx = dx / dist, ny = dy / dist, nz = dz / dist;
                    const double mj = el(B.m, j);
                    const double pox = om * ovx, poy = om * ovy, poz = om * ovz;
                    const double pjx = mj * el(B.vx, j), pjy = mj * el(B.vy, j), pjz = mj * el(B.vz, j);
                    const double amo = pox * nx + poy * ny + poz * nz;
                    const double ami = pjx * nx + pjy * ny + pjz * nz;
                    const double ms = om + mj;
                    const double vcx = (pox + pjx) / ms, vcy = (poy + pjy) / ms, vcz = (poz + pjz) / ms;
                    const double curr = amo - ami - (om - mj) * (vcx * nx + vcy * ny + vcz * nz);
                    if (curr >= best) {
                        best = curr;
                        bestj = j;
                    }
                }
                const int k = atomicAdd(&s_nent, 1);
                if (k < kMaxEntries) s_ent[k] = j;
                else atomicExch(out.err, 1);
            }
            // the LAST j reaching the largest curr_mom (the sequential "curr >= max_mom" rule)
            s_rv[t] = best;
            s_rj[t] = bestj;
            __syncthreads();
            for (int off = kThreads / 2; off > 0; off >>= 1) {
                if (t < off) {
                    const double v2 = s_rv[t + off];
                    const int j2 = s_rj[t + off];
                    if (j2 >= 0 && (s_rj[t] < 0 || v2 > s_rv[t] || (v2 == s_rv[t] && j2 > s_rj[t]))) {
                        s_rv[t] = v2;
                        s_rj[t] = j2;
                    }
                }
                __syncthreads();
            }
            const int cand = s_rj[0];
            const double cand_mom = s_rv[0];
            if (!(merge && cand >= 0 && cand_mom >= s_maxmom)) break;  // uniform
            // ---- merge outer with the candidate (gregsys.hpp:766-808): the
            // merge constructor body(one, two) for the first (gregbod.hpp:
            // 933-940), operator+= after it (gregbod.hpp:1065-1082) -- the
            // same arithmetic field by field, one field per thread; only the
            // label differs (a new one for the first merge) ----
            double nv = 0.0;
            if (t < 12) {
                const double cm = el(B.m, cand);
                const double ov = s_ob[t], cv = body_field(B, t)[cand];
                if (t < 9) nv = mw(s_ob[9], ov, cm, cv);
                else if (t == 9) nv = ov + cm;
                else if (t == 10) nv = newton_cbrt(ov * ov * ov + cv * cv * cv);
                else nv = (ov + cv) / 2;
            }
            if (t == 12) {
                const int64_t clab = el(B.lab, cand);
                if (!s_first) {
                    s_ob_lab = next_label++;
                    s_first = 1;
                }
                const int64_t k = *out.log_len;
                if (k < out.log_cap) {
                    out.log[4 * k + 0] = step;
                    out.log[4 * k + 1] = s_outer_label;
                    out.log[4 * k + 2] = clab;
                    out.log[4 * k + 3] = s_ob_lab;
                }
                *out.log_len = k + 1;
            }
            __syncthreads();  // every read of the old outer body and of s_maxmom is done
            if (t < 12) s_ob[t] = nv;
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
                s_maxmom = cand_mom;
                s_n = cur_n - 1;
                s_dirty = 1;
            }
            if (cand < cur) --cur;
            rescan = true;
        }
        // ---- bounce entries in (distance, j) order (gregsys.hpp:810-837) ----
        if (t == 0) {
            const int ne = min(s_nent, kMaxEntries);
            const double ox = s_ob[0], oy = s_ob[1], oz = s_ob[2];
            const double om = s_ob[9], oe = s_ob[11];
            // the outer body's momentum as it stands when the entries are taken
            const double pox = om * s_ob[3], poy = om * s_ob[4], poz = om * s_ob[5];
            double vx = s_ob[3], vy = s_ob[4], vz = s_ob[5];
            for (int k = 0; k < ne; ++k) {
                const int j = s_ent[k];
                const double dx = el(B.x, j) - ox, dy = el(B.y, j) - oy, dz = el(B.z, j) - oz;
                s_dist[k] = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
            }
            for (int a = 1; a < ne; ++a) {  // insertion sort by (dist, j)
                const double dv = s_dist[a];
                const int jv = s_ent[a];
                int b = a - 1;
                while (b >= 0 && (s_dist[b] > dv || (s_dist[b] == dv && s_ent[b] > jv))) {
                    s_dist[b + 1] = s_dist[b];
                    s_ent[b + 1] = s_ent[b];
                    --b;
                }
                s_dist[b + 1] = dv;
                s_ent[b + 1] = jv;
            }
            for (int k = 0; k < ne; ++k) {
                if (k > 0 && s_dist[k] == s_dist[k - 1]) continue;  // equal keys: the first emplace wins
                const int j = s_ent[k];
                // each j appears once, so its velocity is still the one the scan saw
                const double dx = el(B.x, j) - ox, dy = el(B.y, j) - oy, dz = el(B.z, j) - oz;
                const double dist = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
                const double nx = dx / dist, ny = dy / dist, nz = dz / dist;
                const double mj = el(B.m, j);
                const double amo = pox * nx + poy * ny + poz * nz;
                const double ami = (mj * el(B.vx, j)) * nx + (mj * el(B.vy, j)) * ny + (mj * el(B.vz, j)) * nz;
                const double o_vel = amo / om;
                const double i_vel = ami / mj;
                const double o_minus_i = o_vel - i_vel;
                const double avg_rest = (oe + el(B.e, j)) / 2;
                if (o_vel > 0 || i_vel < 0) {
                    if (o_vel <= 0 && i_vel < 0)
                        if (o_vel < i_vel) continue;
                    if (o_vel > 0 && i_vel >= 0)
                        if (o_vel < i_vel) continue;
                    const double tot = amo + ami;
                    const double ms = om + mj;
                    const double new_o = (tot - mj * avg_rest * o_minus_i) / ms;
                    const double new_i = (tot + om * avg_rest * o_minus_i) / ms;
                    const double fo = new_o - o_vel, fi = new_i - i_vel;
                    vx += fo * nx; vy += fo * ny; vz += fo * nz;
                    el(B.vx, j) += fi * nx; el(B.vy, j) += fi * ny; el(B.vz, j) += fi * nz;
                }
            }
            if (s_first) {
                for (int f = 0; f < 12; ++f) body_field(B, f)[cur] = s_ob[f];
                el(B.lab, cur) = s_ob_lab;
            }
            el(B.vx, cur) = vx;
            el(B.vy, cur) = vy;
            el(B.vz, cur) = vz;
        }
        outer = cur + 1;
        __syncthreads();
    }
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
    coll_replay(B, s_nb, merge != 0, min_mom, step, s_next, nullptr, flags, out);
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

// Bodies per system that k_batch_steps<MIXED, MAXB> holds: each of the
// workgroup's 4 waves owns up to MAXB blocks of 64 consecutive bodies.
template <int MAXB>
constexpr int batch_capacity() {
    return kThreads * MAXB;
}

static_assert(batch_capacity<4>() == kBatchMaxN, "the largest instance holds the largest system");

// j bodies per group in the FP32 tile loop (even)
#ifndef NB_BATCH_UNROLL
#define NB_BATCH_UNROLL 4
#endif
// 0 runs the exact replay after every step as well (a CPU test build uses it
// to show that the prefilter skips only steps without overlaps).
#ifndef NB_BATCH_PREFILTER
#define NB_BATCH_PREFILTER 1
#endif
// Makes a value opaque to the optimiser at this point (no code is emitted).
#ifndef NB_OPAQUE
#define NB_OPAQUE(x) __asm__ volatile("" : "+v"(x))
#endif
// Margin of the FP32 overlap prefilter below 1 (see k_batch_steps).
constexpr float kPrefilterMargin = 0x1.0p-20f;

// One system per workgroup for nsteps KDK steps (evolve :2946-2968) with the
// collision pass after each (FUNC_TEMPL_SELECT :16-57).
//
// Lane `lane` of wave w owns bodies (b0 + 4k)*64 + lane, k < MAXB, with
// b0 = (w - blockIdx.x) mod 4: the waves that hold one block more than the
// others fall on different SIMDs from system to system.  Between steps each
// body's FP64 x, v and a stay in its owner's registers; global memory is read
// at the start, written at the end, and around a collision replay.
//
// MIXED: FP32 pair forces.  Positions are staged in LDS as FP32 offsets from
// body n-1 as it stood at the previous force pass (a nearby fragment), with
// G m; per 64-j tile the sums are plain FP32, across tiles Kahan-compensated.
// The collision prefilter rides on the force loop: with
// R_i = r_i (1 + 2^-18) + 2^-21 max|x'_i| (bounding the radius plus the FP32
// position and distance error), an overlap r_i + r_j > |p_j - p_i| implies
// (R_i + R_j)/d_ij > 1, and R_i max_j(1/d_ij) + max_j(R_j/d_ij) bounds that
// from above, so bodies below 1 - 2^-20 cannot overlap anything.  The replay
// runs only on steps where some body is above it.  The self pair sits in the
// one tile where j and i share a block and is masked there.  Two distinct
// bodies at the same FP32 position make the sums non-finite; such a body is
// summed again with those pairs skipped and is always tested exactly.  With
// softening the prefilter is the squared-distance count d2 < (R_i + R_j)^2.
//
// FP64: the reference's order (cumulative_acc :352-357, or bh_acc's softened
// form), j ascending, positions, masses and radii staged in LDS as FP64; the
// pair loop also takes the replay's own squared overlap test, so the replay
// runs on steps where some pair passes it.
//
// Registers: the 3-block instances are held to 4 waves per SIMD (128 VGPRs;
// the compiler spills a few values, reloaded once per owned block per step,
// outside the tile loops); the 4-block instance would spill inside the force
// code at that bound and runs at 3.
template <bool MIXED, int MAXB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(MAXB <= 3 ? 4 : 1))) void k_batch_steps(
    BodyView all, const int *__restrict__ order, const int64_t *__restrict__ offsets, int *counts,
    int64_t *next_labels, int64_t *log, int64_t *log_lens, int *err, BatchParams P, int64_t step0, int64_t nsteps) {
    constexpr int NMAX = batch_capacity<MAXB>();
    constexpr int kStage = MIXED ? 20 * NMAX : 40 * NMAX;
    __shared__ __align__(16) unsigned char s_stage[kStage];
    __shared__ unsigned char s_pref[NMAX], s_exact[NMAX];
    __shared__ double s_ref[3];
    __shared__ int s_nb;
    __shared__ int64_t s_next;
    float4 *const s_pos = (float4 *)s_stage;                 // MIXED: x', y', z', G m
    float *const s_R = (float *)(s_stage + 16 * NMAX);        // MIXED: prefilter radius
    double *const s_x = (double *)s_stage;                    // FP64: x, y, z, m, r blocks
    double *const s_y = s_x + NMAX, *const s_z = s_y + NMAX, *const s_m = s_z + NMAX, *const s_r = s_m + NMAX;

    const int sys = order[blockIdx.x];
    const int64_t off = offsets[sys];
    const BodyView B = all.at(off);
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int b0 = ((t >> 6) - (int)(blockIdx.x & 3)) & 3;
    if (t == 0) {
        s_nb = counts[sys];
        s_next = next_labels[sys];
    }
    __syncthreads();
    ReplayOut out{log + 4 * off, offsets[sys + 1] - off, log_lens + sys, err};
    int n = s_nb;

    double X[MAXB][3], V[MAXB][3], A[MAXB][3];
    float GM[MAXB], R0[MAXB];
    // owned state from global memory; the per-body constants to registers or LDS
    // Owned bodies are addressed by a 32-bit byte offset from the (uniform)
    // array bases.  The offset passes through NB_OPAQUE where it is used, so
    // the compiler cannot hoist a 64-bit address per array and body out of
    // the step loop (that would hold 2 VGPRs per array per owned body).
    auto at = [](double *base, unsigned byte_off) -> double & { return *(double *)((char *)base + byte_off); };
    auto load_state = [&]() {
#pragma unroll
        for (int k = 0; k < MAXB; ++k) {
            const int i = (b0 + 4 * k) * 64 + lane;
            unsigned o = (unsigned)i * 8u;
            NB_OPAQUE(o);
            // every register is written, so nothing held before a replay stays live across it
#pragma unroll
            for (int c = 0; c < 3; ++c) X[k][c] = V[k][c] = A[k][c] = 0.0;
            GM[k] = R0[k] = 0.f;
            if (i < n) {
                X[k][0] = at(B.x, o); X[k][1] = at(B.y, o); X[k][2] = at(B.z, o);
                V[k][0] = at(B.vx, o); V[k][1] = at(B.vy, o); V[k][2] = at(B.vz, o);
                A[k][0] = at(B.ax, o); A[k][1] = at(B.ay, o); A[k][2] = at(B.az, o);
                if (MIXED) {
                    GM[k] = (float)(P.G * at(B.m, o));
                    R0[k] = (float)at(B.r, o) * (1.0f + 0x1.0p-18f);
                    if (i == n - 1) {
                        s_ref[0] = X[k][0];
                        s_ref[1] = X[k][1];
                        s_ref[2] = X[k][2];
                    }
                } else {
                    s_m[i] = at(B.m, o);
                    s_r[i] = at(B.r, o);
                }
            }
        }
    };
    auto store_state = [&]() {
#pragma unroll
        for (int k = 0; k < MAXB; ++k) {
            const int i = (b0 + 4 * k) * 64 + lane;
            unsigned o = (unsigned)i * 8u;
            NB_OPAQUE(o);
            if (i < n) {
                at(B.x, o) = X[k][0]; at(B.y, o) = X[k][1]; at(B.z, o) = X[k][2];
                at(B.vx, o) = V[k][0]; at(B.vy, o) = V[k][1]; at(B.vz, o) = V[k][2];
                at(B.ax, o) = A[k][0]; at(B.ay, o) = A[k][1]; at(B.az, o) = A[k][2];
            }
        }
    };
    load_state();
    __syncthreads();
    for (int64_t st = 1; st <= nsteps; ++st) {
        if (n == 0) break;
        // opening half kick and drift (evolve :2950-2955), then stage positions
        const double rx = s_ref[0], ry = s_ref[1], rz = s_ref[2];
#pragma unroll
        for (int k = 0; k < MAXB; ++k) {
            const int i = (b0 + 4 * k) * 64 + lane;
            if (i < n) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    V[k][c] = V[k][c] + P.half_dt * A[k][c];
                    X[k][c] = X[k][c] + P.dt * V[k][c];
                }
                if (MIXED) {
                    const float px = (float)(X[k][0] - rx), py = (float)(X[k][1] - ry), pz = (float)(X[k][2] - rz);
                    s_pos[i] = make_float4(px, py, pz, GM[k]);
                    const float mag = fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz)));
                    s_R[i] = R0[k] + 0x1.0p-21f * mag;
                } else {
                    s_x[i] = X[k][0];
                    s_y[i] = X[k][1];
                    s_z[i] = X[k][2];
                }
            }
        }
        __syncthreads();
        int flagged = 0;
#pragma unroll
        for (int k = 0; k < MAXB; ++k) {
            const int blk = b0 + 4 * k;
            if (blk * 64 >= n) continue;  // uniform over the wave
            const int i = blk * 64 + lane;
            const bool valid = i < n;
            double ax, ay, az;
            if (MIXED) {
                const float4 pi = s_pos[i];
                const float Ri = s_R[i];
                float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
                int flag;
                if (P.eps2 == 0.0) {
                    float mi = 0.f, mr = 0.f;
                    for (int jt = 0; jt < n; jt += 64) {
                        const int je = min(n, jt + 64);
                        float px = 0.f, py = 0.f, pz = 0.f;
                        if (jt != blk * 64) {
                            // j in groups of NB_BATCH_UNROLL (independent chains for the
                            // latency of rsq and LDS; the maxima take three-operand maxes)
                            int j = jt;
                            for (; j + NB_BATCH_UNROLL - 1 < je; j += NB_BATCH_UNROLL) {
                                float inv[NB_BATCH_UNROLL], wq[NB_BATCH_UNROLL], dq[NB_BATCH_UNROLL][3];
#pragma unroll
                                for (int u = 0; u < NB_BATCH_UNROLL; ++u) {
                                    const float4 pj = s_pos[j + u];
                                    dq[u][0] = pj.x - pi.x;
                                    dq[u][1] = pj.y - pi.y;
                                    dq[u][2] = pj.z - pi.z;
                                    const float r2 = __builtin_fmaf(
                                        dq[u][2], dq[u][2], __builtin_fmaf(dq[u][1], dq[u][1], dq[u][0] * dq[u][0]));
                                    inv[u] = __builtin_amdgcn_rsqf(r2);
                                    wq[u] = pj.w * inv[u] * (inv[u] * inv[u]);
                                }
#pragma unroll
                                for (int u = 0; u < NB_BATCH_UNROLL; ++u) {
                                    px = __builtin_fmaf(wq[u], dq[u][0], px);
                                    py = __builtin_fmaf(wq[u], dq[u][1], py);
                                    pz = __builtin_fmaf(wq[u], dq[u][2], pz);
                                }
#pragma unroll
                                for (int u = 0; u < NB_BATCH_UNROLL; u += 2) {
                                    mi = fmaxf(mi, fmaxf(inv[u], inv[u + 1]));
                                    mr = fmaxf(mr, fmaxf(s_R[j + u] * inv[u], s_R[j + u + 1] * inv[u + 1]));
                                }
                            }
                            for (; j < je; ++j) {
                                const float4 pj = s_pos[j];
                                const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                                const float inv = __builtin_amdgcn_rsqf(r2);
                                const float w = pj.w * inv * (inv * inv);
                                px = __builtin_fmaf(w, dx, px);
                                py = __builtin_fmaf(w, dy, py);
                                pz = __builtin_fmaf(w, dz, pz);
                                mi = fmaxf(mi, inv);
                                mr = fmaxf(mr, s_R[j] * inv);
                            }
                        } else {
                            // the block's own tile: pairs of j, the self pair taking inv = 0
                            int j = jt;
                            for (; j + 1 < je; j += 2) {
                                float inv[2], wq[2], dq[2][3];
#pragma unroll
                                for (int u = 0; u < 2; ++u) {
                                    const float4 pj = s_pos[j + u];
                                    dq[u][0] = pj.x - pi.x;
                                    dq[u][1] = pj.y - pi.y;
                                    dq[u][2] = pj.z - pi.z;
                                    const float r2 = __builtin_fmaf(
                                        dq[u][2], dq[u][2], __builtin_fmaf(dq[u][1], dq[u][1], dq[u][0] * dq[u][0]));
                                    inv[u] = j + u == i ? 0.f : __builtin_amdgcn_rsqf(r2);
                                    wq[u] = pj.w * inv[u] * (inv[u] * inv[u]);
                                }
#pragma unroll
                                for (int u = 0; u < 2; ++u) {
                                    px = __builtin_fmaf(wq[u], dq[u][0], px);
                                    py = __builtin_fmaf(wq[u], dq[u][1], py);
                                    pz = __builtin_fmaf(wq[u], dq[u][2], pz);
                                }
                                mi = fmaxf(mi, fmaxf(inv[0], inv[1]));
                                mr = fmaxf(mr, fmaxf(s_R[j] * inv[0], s_R[j + 1] * inv[1]));
                            }
                            for (; j < je; ++j) {
                                const float4 pj = s_pos[j];
                                const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                                const float inv = j == i ? 0.f : __builtin_amdgcn_rsqf(r2);
                                const float w = pj.w * inv * (inv * inv);
                                px = __builtin_fmaf(w, dx, px);
                                py = __builtin_fmaf(w, dy, py);
                                pz = __builtin_fmaf(w, dz, pz);
                                mi = fmaxf(mi, inv);
                                mr = fmaxf(mr, s_R[j] * inv);
                            }
                        }
                        kahan_add(sx, cx, px);
                        kahan_add(sy, cy, py);
                        kahan_add(sz, cz, pz);
                    }
                    flag = __builtin_fmaf(Ri, mi, mr) >= 1.0f - kPrefilterMargin ? 1 : 0;
                    if (valid && !(isfinite(sx) && isfinite(sy) && isfinite(sz))) {
                        // a pair at zero FP32 distance: sum again without such pairs
                        sx = sy = sz = cx = cy = cz = 0.f;
                        for (int jt = 0; jt < n; jt += 64) {
                            const int je = min(n, jt + 64);
                            float px = 0.f, py = 0.f, pz = 0.f;
                            for (int j = jt; j < je; ++j) {
                                const float4 pj = s_pos[j];
                                const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
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
                        flag = 1;
                    }
                } else {
                    const float eps2 = (float)P.eps2;
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
                    flag = hits > 1 ? 1 : 0;
                }
                ax = (double)sx - (double)cx;
                ay = (double)sy - (double)cy;
                az = (double)sz - (double)cz;
                if (valid) {
                    s_pref[i] = (unsigned char)flag;
                    flagged |= flag;
                }
            } else {
                // The squared distance of each pair is the collision test's own
                // (the same operands and order), so the replay's squared test
                // d2 < (r_i + r_j)^2 (1 + 1e-9), which every overlap passes,
                // flags the bodies the exact test must look at.
                const double xi = X[k][0], yi = X[k][1], zi = X[k][2], ri = s_r[i];
                ax = 0.0;
                ay = 0.0;
                az = 0.0;
                bool hit = false;
                if (P.eps2 == 0.0) {
                    for (int j = 0; j < n; ++j) {
                        if (j == i) continue;
                        const double dx = s_x[j] - xi, dy = s_y[j] - yi, dz = s_z[j] - zi;
                        const double s = dx * dx + dy * dy + dz * dz;
                        const double rs = ri + s_r[j];
                        hit = hit || s < rs * rs * (1.0 + 1e-9);
                        const double wx = s * dx, wy = s * dy, wz = s * dz;
                        const double r3 = __builtin_sqrt(wx * wx + wy * wy + wz * wz);
                        const double kf = (P.G * s_m[j]) / r3;
                        ax += kf * dx; ay += kf * dy; az += kf * dz;
                    }
                } else {
                    for (int j = 0; j < n; ++j) {
                        if (j == i) continue;
                        const double dx = s_x[j] - xi, dy = s_y[j] - yi, dz = s_z[j] - zi;
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        const double rs = ri + s_r[j];
                        hit = hit || d2 < rs * rs * (1.0 + 1e-9);
                        const double s = d2 + P.eps2;
                        const double sq = __builtin_sqrt(s);
                        const double kf = P.G * (s_m[j] / s);
                        ax += kf * (dx / sq); ay += kf * (dy / sq); az += kf * (dz / sq);
                    }
                }
                if (valid) {
                    s_pref[i] = hit ? 1 : 0;
                    flagged |= hit ? 1 : 0;
                }
            }
            if (valid) {
                A[k][0] = ax;
                A[k][1] = ay;
                A[k][2] = az;
                // closing half kick (loop_kdk :499)
#pragma unroll
                for (int c = 0; c < 3; ++c) V[k][c] = V[k][c] + P.half_dt * A[k][c];
                if (MIXED && i == n - 1) {
                    s_ref[0] = X[k][0];
                    s_ref[1] = X[k][1];
                    s_ref[2] = X[k][2];
                }
            }
        }
        // every read of this step's staged positions is done before the next step writes them
        const int any = __syncthreads_or(flagged);
        constexpr bool use_pref = NB_BATCH_PREFILTER;
        if (P.coll && (any || !use_pref)) {
            store_state();
            __syncthreads();
            coll_replay(B, s_nb, P.merge != 0, P.min_mom, step0 + st, s_next, use_pref ? s_pref : nullptr, s_exact,
                        out);
            n = s_nb;
            load_state();
            __syncthreads();
        }
    }
    store_state();
    if (t == 0) {
        counts[sys] = s_nb;
        next_labels[sys] = s_next;
    }
}
