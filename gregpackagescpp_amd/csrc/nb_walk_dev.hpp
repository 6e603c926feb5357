// nb_walk_dev.hpp -- device code of the FP32 Barnes-Hut walks: the per-wave
// depth-first walk and the frontier walk (see nb_bh.hip for the tree build,
// the launches and the design).
//
// Kept in a header so that the same source compiles in two places: inside
// nb_bh.hip for gfx950 (HIP), and in a CPU build that runs each wave's 64
// lanes as cooperatively scheduled contexts, with the wave intrinsics
// (ballot, shuffles, readfirstlane) as exchange points and bounds-checked
// buffers (tests/emu/).  The includer provides the HIP vocabulary (float4,
// int4, __shared__, __ballot, __shfl*, atomicAdd ...), kBlock and
// NB_WAVES_PER_EU, and opens the namespace.

// Walk record of a node: FP32 COM and G*mass, then skip pointer, first sorted
// body, body count and the FP32 bits of the squared opening radius, in one
// 32-byte record so that a node visit reads one cache line.
struct __align__(32) WalkNode {
    float4 q;
    int4 r;
};

// ---- per-wave depth-first walk, FP32 monopoles ----
// The wave shares one node sequence.  A node is accepted when every lane's MAC
// holds; a rejected node with at most `bucket` bodies is not descended but
// summed body by body (its bodies are contiguous in sorted order), which
// replaces up to 2*bucket-1 dependent node visits by a streamed loop (the
// autotuner picks the bucket; 1 disables it).
__device__ __forceinline__ void monopole(float4 q, float4 p, float eps2, bool soft, float &ax, float &ay,
                                         float &az) {
    const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
    const float s2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
    float inv = __builtin_amdgcn_rsqf(s2);
    if (!soft) inv = s2 > 0.f ? inv : 0.f;  // own body at eps = 0
    float w = q.w * inv;
    w = w * (inv * inv);
    ax = __builtin_fmaf(w, dx, ax);
    ay = __builtin_fmaf(w, dy, ay);
    az = __builtin_fmaf(w, dz, az);
}

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (the LDS unit serves one wave's requests in order; the fences stop the
// compiler from moving the accesses across this point).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// With a gate, the walk runs only when *gate is set (the frontier walk's
// overflow flag), so it can follow that walk on the stream unconditionally.
template <bool SOFT>
__global__ __launch_bounds__(kBlock) void k_walk_wave(int64_t n, int32_t total, const float4 *__restrict__ pos,
                                                      const WalkNode *__restrict__ wnode, int bucket, float eps2,
                                                      double *__restrict__ accs,
                                                      unsigned long long *__restrict__ icount,
                                                      const int *__restrict__ gate) {
#pragma clang fp contract(fast)
    if (gate && __builtin_amdgcn_readfirstlane(*gate) == 0) return;
    // Workgroups are dealt to the 8 XCDs round robin; renumber them so that
    // each XCD walks one contiguous Morton range (its own part of the tree
    // stays in its L2).  Applies when the grid splits evenly.
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t lb = (g % 8u == 0u) ? (b % 8u) * (g / 8u) + b / 8u : b;
    const int64_t s = (int64_t)lb * kBlock + threadIdx.x;
    const bool live = s < n;
    const float4 p = live ? pos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t active = __ballot(live);
    if (active == 0) return;
    __shared__ float4 s_leaf[kBlock / 64][64];
    const unsigned lane = __lane_id(), wv = threadIdx.x / 64;
    float ax = 0.f, ay = 0.f, az = 0.f;
    unsigned cnt = 0;
    int32_t node = 1;
    while (node < total) {
        node = __builtin_amdgcn_readfirstlane(node);
        const WalkNode nd = wnode[node];
        const float4 q = nd.q;
        const int4 r = nd.r;
        const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        const bool acc_lane = d2 > __int_as_float(r.w);
        if ((__ballot(acc_lane) & active) == active) {
            // the monopole of the accepted node from the MAC's d and |d|^2
            const float s2 = d2 + eps2;
            float inv = __builtin_amdgcn_rsqf(s2);
            if (!SOFT) inv = s2 > 0.f ? inv : 0.f;
            const float w = (q.w * inv) * (inv * inv);
            ax = __builtin_fmaf(w, dx, ax);
            ay = __builtin_fmaf(w, dy, ay);
            az = __builtin_fmaf(w, dz, az);
            ++cnt;
            node = r.x;
        } else if (r.z <= bucket) {
            // the node's bodies, 64 at a time: one coalesced load by the
            // lanes into this wave's LDS slot, then the same-address reads
            // broadcast each body, so the monopoles take VGPR operands (an
            // SGPR operand halves the VALU issue rate).  Same order as a
            // body-by-body loop.
            for (int32_t base = 0; base < r.z; base += 64) {
                const int32_t k = min(64, r.z - base);
                if ((int32_t)lane < k) s_leaf[wv][lane] = pos[r.y + base + (int32_t)lane];
                wave_lds_sync();
                int32_t j = 0;
                for (; j + 4 <= k; j += 4) {
                    const float4 b0 = s_leaf[wv][j], b1 = s_leaf[wv][j + 1], b2 = s_leaf[wv][j + 2],
                                 b3 = s_leaf[wv][j + 3];
                    monopole(b0, p, eps2, SOFT, ax, ay, az);
                    monopole(b1, p, eps2, SOFT, ax, ay, az);
                    monopole(b2, p, eps2, SOFT, ax, ay, az);
                    monopole(b3, p, eps2, SOFT, ax, ay, az);
                }
                for (; j < k; ++j) monopole(s_leaf[wv][j], p, eps2, SOFT, ax, ay, az);
                wave_lds_sync();  // every lane has read the slot before it is refilled
            }
            cnt += (unsigned)r.z;
            node = r.x;
        } else {
            node = node + 1;
        }
    }
    if (live) {
        accs[s] = ax;
        accs[n + s] = ay;
        accs[2 * n + s] = az;
    }
    // cnt is wave-uniform: one atomic per wave
    if (__lane_id() == (unsigned)__ffsll((unsigned long long)active) - 1)
        atomicAdd(icount, (unsigned long long)cnt * (unsigned long long)__popcll(active));
}

// ---- frontier walk, FP32 monopoles (the default) ----
// One wave per 64 Morton-adjacent bodies (a group), as in the wave walk, and
// the same cells accepted: a cell is taken as a monopole when every body of
// the group satisfies d^2 > r^2_open, a rejected cell with at most `bucket`
// bodies is summed body by body, and other rejected cells are opened.  The
// traversal differs: each step pops up to 64 cells off the wave's stack and
// tests one per lane, first against the group's bounding box (distance from
// the cell's centre of mass to the box: if that passes, every body passes,
// in FP32 as well, since rounding is monotonic) and otherwise body by body.
// Accepted cells and bucket ranges go to per-wave LDS lists that every lane
// then sums with VGPR operands; opened cells push their children, which the
// grouped record layout keeps adjacent.  Each XCD takes one contiguous range
// of the group order, which lists the groups of that range heaviest first by
// the previous walk's counts.  A stack that would overflow sets *werr; the
// depth-first wave walk then redoes the whole step (its launch follows,
// gated on that flag).
constexpr int kIL = 128;        // accepted cells buffered per wave (LDS fits 8 waves per SIMD)
constexpr int kBL = 64;         // bucket ranges buffered per wave
constexpr int kFWaves = kBlock / 64;
constexpr int kFStackDefault = 4096;

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_incl_scan(int v, unsigned lane) {
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o);
        v += (int)lane >= o ? t : 0;
    }
    return v;
}
__device__ __forceinline__ int lanes_below(uint64_t m, unsigned lane) {
    return __popcll(m & ((1ull << lane) - 1ull));
}

template <bool SOFT>
__global__ __launch_bounds__(kBlock) NB_WAVES_PER_EU(8) void k_walk_frontier(
    int64_t n, int32_t total, const float4 *__restrict__ pos, const WalkNode *__restrict__ wgrp, int bucket,
    float eps2, double *__restrict__ accs, unsigned long long *__restrict__ icount, int32_t *__restrict__ stacks,
    int32_t cap, const int32_t *__restrict__ order, uint32_t *__restrict__ cost, int *__restrict__ werr) {
#pragma clang fp contract(fast)
    __shared__ float4 s_il[kFWaves][kIL];
    __shared__ int2 s_bl[kFWaves][kBL];
    __shared__ float4 s_body[kFWaves][64];
    __shared__ float4 s_grp[kFWaves][64];
    const unsigned lane = __lane_id(), wv = threadIdx.x >> 6;
    // One wave per group, no loop over groups and no lane-0-only code in any
    // loop: a group handed out by one lane's atomic and broadcast by
    // readfirstlane inside a persistent loop was compiled into a loop that
    // lanes 1-63 never left.  Workgroups are renumbered so that each XCD
    // takes one contiguous range of the order (heaviest groups first).
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t lb = (g % 8u == 0u) ? (b % 8u) * (g / 8u) + b / 8u : b;
    const int64_t ngroups = (n + 63) / 64;
    const int64_t gi = (int64_t)lb * kFWaves + wv;
    if (gi >= ngroups) return;
    int32_t *const stk = stacks + ((size_t)b * kFWaves + wv) * (size_t)cap;
    const int4 r0 = wgrp[0].r;
    const int32_t cs0 = r0.x, nc0 = min(8, r0.z & 15);
    {
        const int64_t grp = order ? (int64_t)order[gi] : gi;
        const int64_t s = grp * 64 + lane;
        const bool live = s < n;
        const float4 p = live ? pos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint64_t active = __ballot(live);
        const int nlive = __popcll(active);
        s_grp[wv][lane] = p;
        wave_lds_sync();
        const float lox = wave_min(live ? p.x : INFINITY), hix = wave_max(live ? p.x : -INFINITY);
        const float loy = wave_min(live ? p.y : INFINITY), hiy = wave_max(live ? p.y : -INFINITY);
        const float loz = wave_min(live ? p.z : INFINITY), hiz = wave_max(live ? p.z : -INFINITY);
        float ax = 0.f, ay = 0.f, az = 0.f;
        unsigned cnt = 0;
        int il = 0, bl = 0;
        auto flush_il = [&]() {
            wave_lds_sync();
            int j = 0;
            for (; j + 4 <= il; j += 4) {
                const float4 c0 = s_il[wv][j], c1 = s_il[wv][j + 1], c2 = s_il[wv][j + 2], c3 = s_il[wv][j + 3];
                monopole(c0, p, eps2, SOFT, ax, ay, az);
                monopole(c1, p, eps2, SOFT, ax, ay, az);
                monopole(c2, p, eps2, SOFT, ax, ay, az);
                monopole(c3, p, eps2, SOFT, ax, ay, az);
            }
            for (; j < il; ++j) monopole(s_il[wv][j], p, eps2, SOFT, ax, ay, az);
            cnt += (unsigned)il;
            il = 0;
            wave_lds_sync();
        };
        auto flush_bl = [&]() {
            wave_lds_sync();
            for (int q = 0; q < bl; ++q) {
                const int2 rg = s_bl[wv][q];
                const int32_t j0 = __builtin_amdgcn_readfirstlane(rg.x), nb = __builtin_amdgcn_readfirstlane(rg.y);
                for (int32_t base = 0; base < nb; base += 64) {
                    const int32_t k = min(64, nb - base);
                    if ((int32_t)lane < k) s_body[wv][lane] = pos[j0 + base + (int32_t)lane];
                    wave_lds_sync();
                    int32_t j = 0;
                    for (; j + 4 <= k; j += 4) {
                        const float4 b0 = s_body[wv][j], b1 = s_body[wv][j + 1], b2 = s_body[wv][j + 2],
                                     b3 = s_body[wv][j + 3];
                        monopole(b0, p, eps2, SOFT, ax, ay, az);
                        monopole(b1, p, eps2, SOFT, ax, ay, az);
                        monopole(b2, p, eps2, SOFT, ax, ay, az);
                        monopole(b3, p, eps2, SOFT, ax, ay, az);
                    }
                    for (; j < k; ++j) monopole(s_body[wv][j], p, eps2, SOFT, ax, ay, az);
                    wave_lds_sync();
                }
                cnt += (unsigned)nb;
            }
            bl = 0;
        };
        stk[lane] = cs0 + (int32_t)lane;  // entries past nc0 are never popped (cap >= 64)
        int32_t sp = nc0;
        bool over = false;
        int32_t steps = 0;  // each step pops at least one cell of the tree
        while (sp > 0 && steps++ < total) {
            // this wave's stack writes are complete before its lanes read them
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const int32_t take = min(64, sp);
            sp -= take;
            const bool valid = (int32_t)lane < take;
            const int32_t node = valid ? stk[sp + (int32_t)lane] : 0;
            const bool inb = valid && node > 0 && node < total;
            WalkNode rec;
            rec.q = make_float4(0.f, 0.f, 0.f, 0.f);
            rec.r = make_int4(0, 0, 0, 0);
            if (inb) rec = wgrp[node];
            const float dx = fmaxf(fmaxf(lox - rec.q.x, rec.q.x - hix), 0.f);
            const float dy = fmaxf(fmaxf(loy - rec.q.y, rec.q.y - hiy), 0.f);
            const float dz = fmaxf(fmaxf(loz - rec.q.z, rec.q.z - hiz), 0.f);
            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const float r2o = __int_as_float(rec.r.w);
            bool acc = inb && d2 > r2o;
            if (inb && !acc) {
                // the box test is only sufficient: decide body by body with
                // the wave walk's test
                bool all = true;
                for (int i = 0; i < nlive; ++i) {
                    const float4 b = s_grp[wv][i];
                    const float ex = rec.q.x - b.x, ey = rec.q.y - b.y, ez = rec.q.z - b.z;
                    const float e2 = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
                    if (!(e2 > r2o)) {
                        all = false;
                        break;
                    }
                }
                acc = all;
            }
            const int32_t nbod = rec.r.z >> 4;
            const bool leafb = inb && !acc && nbod <= bucket;
            const bool open = inb && !acc && !leafb;
            const uint64_t m_acc = __ballot(acc), m_leaf = __ballot(leafb);
            const int n_acc = __popcll(m_acc), n_leaf = __popcll(m_leaf);
            if (il + n_acc > kIL) flush_il();
            if (acc) s_il[wv][il + lanes_below(m_acc, lane)] = rec.q;
            il += n_acc;
            if (bl + n_leaf > kBL) flush_bl();
            if (leafb) s_bl[wv][bl + lanes_below(m_leaf, lane)] = make_int2(rec.r.y, nbod);
            bl += n_leaf;
            const int nch = open ? min(8, rec.r.z & 15) : 0;
            const int incl = wave_incl_scan(nch, lane);
            const int tot = __shfl(incl, 63);
            if (sp + tot > cap) {
                over = true;
                break;
            }
            for (int k = 0; k < nch; ++k) stk[sp + incl - nch + k] = rec.r.x + k;
            sp += tot;
        }
        if (il) flush_il();
        if (bl) flush_bl();
        if (over) werr[0] = 1;
        if (live) {
            accs[s] = ax;
            accs[n + s] = ay;
            accs[2 * n + s] = az;
        }
        if (cost) cost[grp] = cnt;  // every lane stores the same value
        if (lane == 0) atomicAdd(icount, (unsigned long long)cnt * (unsigned long long)nlive);
    }
}

