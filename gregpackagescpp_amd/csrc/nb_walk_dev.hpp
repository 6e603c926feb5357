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
// Without wnode (the build wrote only the frontier walk's grouped layout)
// each node's record is put together from its grouped record (wgrp at
// slot[node]) and its skip pointer: the same values, two loads per visit.
template <bool SOFT>
__global__ __launch_bounds__(kBlock) void k_walk_wave(int64_t n, int32_t total, const float4 *__restrict__ pos,
                                                      const WalkNode *__restrict__ wnode, int bucket, float eps2,
                                                      double *__restrict__ accs,
                                                      unsigned long long *__restrict__ icount,
                                                      const int *__restrict__ gate,
                                                      const WalkNode *__restrict__ wgrp,
                                                      const uint32_t *__restrict__ slot,
                                                      const int32_t *__restrict__ skip) {
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
        WalkNode nd;
        if (wnode) {
            nd = wnode[node];
        } else {
            const WalkNode g = wgrp[slot[node]];
            nd.q = g.q;
            nd.r = make_int4(skip[node], g.r.y, g.r.z >> 4, g.r.w);
        }
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
// tests one per lane against the group's bounding box (distance from the
// cell's centre of mass to the box: if that passes, every body passes, in
// FP32 as well, since rounding is monotonic); the cells that fail it are
// then tested one at a time against all the group's bodies at once, one
// lane per body and a ballot per cell.  Accepted cells and bucket ranges go to
// per-wave LDS lists that every lane then sums with VGPR operands; opened
// cells push their children, which the grouped record layout keeps adjacent.
// The stack's top kLStack entries live in LDS; a push that would overflow
// them moves the bottom of the LDS part onto the wave's global stack (the
// entries nearest the top stay in LDS, and every pop takes the top of the
// whole stack), so global memory sees stack traffic only for deep frontiers.
// Each XCD takes one contiguous range of the group order (ranges of about
// equal cost), which lists its groups heaviest first by the previous walk's
// counts.  A stack that
// would exceed `cap` entries in all sets *werr; the depth-first wave walk
// then redoes the whole step (its launch follows, gated on that flag).
constexpr int kIL = 96;         // accepted cells buffered per wave (also the bucket bodies' staging)
constexpr int kBL = 64;         // bucket ranges buffered per wave
static_assert(kIL >= 64 && kBL >= 64, "one step adds up to 64 entries to either list after a flush");
#ifndef NB_FRONTIER_LDS_STACK
#define NB_FRONTIER_LDS_STACK 512  // 8 waves per SIMD fit the LDS (5120 B per wave); a test build sets a small one
#endif
constexpr int kLStack = NB_FRONTIER_LDS_STACK;  // stack entries per wave in LDS
static_assert(kLStack >= 8, "the root's children start in LDS");
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
__device__ __forceinline__ int lanes_below(uint64_t m, unsigned lane) {
    return __popcll(m & ((1ull << lane) - 1ull));
}
// Exclusive prefix sum over the wave of v in [0, 15] and its total, from one
// ballot per bit (no cross-lane data movement through LDS).
__device__ __forceinline__ int wave_excl_scan_small(int v, unsigned lane, int &total) {
    int excl = 0;
    total = 0;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
        const uint64_t m = __ballot((v >> bit) & 1);
        excl += lanes_below(m, lane) << bit;
        total += __popcll(m) << bit;
    }
    return excl;
}
// squared distance from c to the box [lo, hi] (0 inside)
__device__ __forceinline__ float box_d2(float cx, float cy, float cz, float lox, float loy, float loz, float hix,
                                        float hiy, float hiz) {
    const float dx = fmaxf(fmaxf(lox - cx, cx - hix), 0.f);
    const float dy = fmaxf(fmaxf(loy - cy, cy - hiy), 0.f);
    const float dz = fmaxf(fmaxf(loz - cz, cz - hiz), 0.f);
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

template <bool SOFT>
__global__ __launch_bounds__(kBlock) NB_WAVES_PER_EU(8) void k_walk_frontier(
    int64_t n, int32_t total, const float4 *__restrict__ pos, const WalkNode *__restrict__ wgrp, int bucket,
    float eps2, double *__restrict__ accs, unsigned long long *__restrict__ icount, int32_t *__restrict__ stacks,
    int32_t cap, const int32_t *__restrict__ order, const int64_t *__restrict__ xr, uint32_t *__restrict__ cost,
    int *__restrict__ werr) {
#pragma clang fp contract(fast)
    __shared__ float4 s_il[kFWaves][kIL];
    __shared__ int2 s_bl[kFWaves][kBL];
    __shared__ float4 s_fc[kFWaves][64];  // the step's cells that fail the group box: COM, r^2_open
    __shared__ int32_t s_stk[kFWaves][kLStack];
    const unsigned lane = __lane_id(), wv = threadIdx.x >> 6;
    // One wave per group, no loop over groups and no lane-0-only code in any
    // loop: a group handed out by one lane's atomic and broadcast by
    // readfirstlane inside a persistent loop was compiled into a loop that
    // lanes 1-63 never left.  Workgroups are dealt to the 8 XCDs round robin;
    // with xr (9 range bounds) XCD b % 8 takes the groups order[xr[x] ..
    // xr[x+1]) -- one contiguous Morton range of about equal cost, heaviest
    // groups first -- and its workgroups past that range end at once.
    // Without xr the grid is one range.
    const uint32_t b = blockIdx.x;
    const int64_t ngroups = (n + 63) / 64;
    int64_t gi;
    if (xr) {
        const uint32_t x = b % 8u;
        const int64_t lo = xr[x], hi = xr[x + 1];
        gi = lo + (int64_t)(b / 8u) * kFWaves + wv;
        if (gi >= hi) return;
    } else {
        gi = (int64_t)b * kFWaves + wv;
    }
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
        // the group's box (live bodies only), wave-uniform
        auto uni = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
        const float glx = uni(wave_min(live ? p.x : INFINITY)), ghx = uni(wave_max(live ? p.x : -INFINITY));
        const float gly = uni(wave_min(live ? p.y : INFINITY)), ghy = uni(wave_max(live ? p.y : -INFINITY));
        const float glz = uni(wave_min(live ? p.z : INFINITY)), ghz = uni(wave_max(live ? p.z : -INFINITY));
        if (lane < (unsigned)nc0) s_stk[wv][lane] = cs0 + (int32_t)lane;
        wave_lds_sync();
        float ax = 0.f, ay = 0.f, az = 0.f;
        unsigned cnt = 0;
        int il = 0, bl = 0;
        auto flush_il = [&]() {
            wave_lds_sync();
            int j = 0;
            for (; j + 2 <= il; j += 2) {
                const float4 c0 = s_il[wv][j], c1 = s_il[wv][j + 1];
                monopole(c0, p, eps2, SOFT, ax, ay, az);
                monopole(c1, p, eps2, SOFT, ax, ay, az);
            }
            if (j < il) monopole(s_il[wv][j], p, eps2, SOFT, ax, ay, az);
            cnt += (unsigned)il;
            il = 0;
            wave_lds_sync();
        };
        // bucket ranges, their bodies concatenated in list order and staged
        // 64 at a time through the cell list's LDS (emptied first): each lane
        // finds the range holding its body by a search over the ranges'
        // starts, so a round of loads serves several short ranges.  The
        // bodies are summed in the same order as range by range.
        auto flush_bl = [&]() {
            if (il) flush_il();
            wave_lds_sync();
            const int2 rg = (int)lane < bl ? s_bl[wv][lane] : make_int2(0, 0);
            int incl = rg.y;  // inclusive prefix sum of the range lengths
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(incl, o);
                if ((int)lane >= o) incl += y;
            }
            const int total_b = __shfl(incl, 63);
            wave_lds_sync();  // every lane has read its range before the starts replace the lengths
            if ((int)lane < bl) s_bl[wv][lane] = make_int2(rg.x, incl - rg.y);
            wave_lds_sync();
            for (int base = 0; base < total_b; base += 64) {
                const int g = base + (int)lane;
                if (g < total_b) {
                    int lo = 0, hi = bl - 1;  // the last range whose start is at most g
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_bl[wv][mid].y <= g) lo = mid;
                        else hi = mid - 1;
                    }
                    const int2 r = s_bl[wv][lo];
                    s_il[wv][lane] = pos[r.x + (g - r.y)];
                }
                wave_lds_sync();
                const int k = min(64, total_b - base);
                int j = 0;
                for (; j + 2 <= k; j += 2) {
                    const float4 b0 = s_il[wv][j], b1 = s_il[wv][j + 1];
                    monopole(b0, p, eps2, SOFT, ax, ay, az);
                    monopole(b1, p, eps2, SOFT, ax, ay, az);
                }
                if (j < k) monopole(s_il[wv][j], p, eps2, SOFT, ax, ay, az);
                wave_lds_sync();
            }
            cnt += (unsigned)total_b;
            bl = 0;
        };
        int32_t sp = nc0;  // entries in the LDS part
        int32_t gsp = 0;   // entries in the global part (below the LDS part)
        bool over = false;
        int32_t steps = 0;  // each step pops at least one cell of the tree
        while ((sp > 0 || gsp > 0) && steps++ < total) {
            // the top `take` entries of the whole stack (global part, then
            // LDS part), so each pop is the same whatever the split
            const int32_t take = min(64, sp + gsp);
            const int32_t kl = min(take, sp), kg = take - kl;
            int32_t node = 0;
            if (kg > 0) {
                // this wave's global stack writes are complete before its lanes read them
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                if ((int32_t)lane < kg) node = stk[gsp - kg + (int32_t)lane];
            }
            if ((int32_t)lane >= kg && (int32_t)lane < take) node = s_stk[wv][sp - kl + (int32_t)lane - kg];
            sp -= kl;
            gsp -= kg;
            const bool inb = (int32_t)lane < take && node > 0 && node < total;
            WalkNode rec;
            rec.q = make_float4(0.f, 0.f, 0.f, 0.f);
            rec.r = make_int4(0, 0, 0, 0);
            if (inb) rec = wgrp[node];
            const float r2o = __int_as_float(rec.r.w);
            bool acc = inb && box_d2(rec.q.x, rec.q.y, rec.q.z, glx, gly, glz, ghx, ghy, ghz) > r2o;
            // the box is only sufficient: a cell that fails it is accepted
            // when every body passes the wave walk's test, evaluated for all
            // the group's bodies at once (lane l, body l; one ballot per cell)
            const bool test = inb && !acc;
            const uint64_t m_test = __ballot(test);
            if (m_test) {
                const int nt = __popcll(m_test);
                const int at = lanes_below(m_test, lane);
                if (test) s_fc[wv][at] = make_float4(rec.q.x, rec.q.y, rec.q.z, r2o);
                wave_lds_sync();
                uint64_t m_fail = 0;
                for (int k = 0; k < nt; ++k) {
                    const float4 c = s_fc[wv][k];
                    const float ex = c.x - p.x, ey = c.y - p.y, ez = c.z - p.z;
                    const float e2 = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
                    if (__ballot(live && !(e2 > c.w))) m_fail |= 1ull << k;
                }
                if (test) acc = !((m_fail >> at) & 1ull);
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
            int tot;
            const int excl = wave_excl_scan_small(nch, lane, tot);
            if (gsp + sp + tot > cap) {
                over = true;
                break;
            }
            const int32_t first = rec.r.x;
            if (sp + tot > kLStack && tot > kLStack) {
                // pushes that alone exceed the LDS part: all of it moves below
                // to the global part, and the pushes go straight after it
                for (int32_t k = (int32_t)lane; k < sp; k += 64) stk[gsp + k] = s_stk[wv][k];
                gsp += sp;
                sp = 0;
                for (int k = 0; k < nch; ++k) stk[gsp + excl + k] = first + k;
                gsp += tot;
            } else if (sp + tot > kLStack) {
                // the bottom of the LDS part moves below to the global part;
                // its top `keep` entries (popped soonest) slide down in LDS
                const int32_t keep = min(sp, (kLStack - tot) / 2), mv = sp - keep;
                for (int32_t k = (int32_t)lane; k < mv; k += 64) stk[gsp + k] = s_stk[wv][k];
                constexpr int kKeepPer = (kLStack / 2 + 63) / 64;
                int32_t kv[kKeepPer];
#pragma unroll
                for (int q = 0; q < kKeepPer; ++q) {
                    const int32_t k = (int32_t)lane + 64 * q;
                    kv[q] = k < keep ? s_stk[wv][mv + k] : 0;
                }
                wave_lds_sync();  // every kept entry is read before any is rewritten
#pragma unroll
                for (int q = 0; q < kKeepPer; ++q) {
                    const int32_t k = (int32_t)lane + 64 * q;
                    if (k < keep) s_stk[wv][k] = kv[q];
                }
                gsp += mv;
                sp = keep;
                for (int k = 0; k < nch; ++k) s_stk[wv][sp + excl + k] = first + k;
                sp += tot;
            } else {
                for (int k = 0; k < nch; ++k) s_stk[wv][sp + excl + k] = first + k;
                sp += tot;
            }
            wave_lds_sync();  // the pushes are visible to the next pop
        }
        if (il) flush_il();
        if (bl) flush_bl();
        if (over) werr[0] = 1;
        {
            // the group and slot again (not held in registers across the walk)
            const int64_t grp2 = order ? (int64_t)order[gi] : gi;
            const int64_t s2 = grp2 * 64 + lane;
            if (s2 < n) {
                accs[s2] = ax;
                accs[n + s2] = ay;
                accs[2 * n + s2] = az;
            }
            if (cost) cost[grp2] = cnt;  // every lane stores the same value
        }
        if (lane == 0) atomicAdd(icount, (unsigned long long)cnt * (unsigned long long)nlive);
    }
}
