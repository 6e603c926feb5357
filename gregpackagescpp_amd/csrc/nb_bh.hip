// nb_bh.hip -- Barnes-Hut octree build, centre-of-mass pass and tree walk (gfx950).
//
// Reference path: bh_tree::find_min_max / add_bodies (greg8tree.hpp:1012-1078),
// bh_cube::set_corners / dispatch (:81-210), calc_com (:211-282),
// cell_iterator (:702-755) with bh_acc (gregsys.hpp:371-375), and evolve's
// KDK + Barnes-Hut loop (gregsys.hpp:2893-2936).
//
// The reference inserts bodies one by one into a pointer octree whose leaves
// hold exactly one body.  That topology depends only on each body's sequence
// of octant digits under the midpoint recursion mid = (lo + hi)/2 from the
// padded root cube, so it is rebuilt here without insertion:
//  1. bounding cube (FP64 min/max reduction; pad = L/20 -- the reference's
//     long-double quotient rounded to double equals the double quotient);
//  2. per body, 42 levels of octant digits by the same FP64 recursion, packed
//     into two 63-bit keys (levels 1-21, 22-42);
//  3. radix sort by the high key, gather of the low key, and an insertion-sort
//     fix of the rare runs with equal high keys;
//  4. c[s] = shared digits of sorted neighbours s, s+1.  An internal node is a
//     maximal run sharing L >= 1 digits; nodes starting at s are levels
//     (c[s-1], c[s]] plus the leaf of s, so an exclusive scan of
//     1 + max(0, c[s]-c[s-1]) numbers every node in DFS pre-order with octants
//     0..7 -- the order cell_iterator visits them.  One thread per sorted
//     position writes the nodes starting there: a node's last position (the
//     end of its run) and the start of its parent's run come from doubling
//     and bisection searches over the sorted keys, its skip pointer (next
//     node outside its subtree) is the first node starting at last + 1;
//  5. the bodies move into the tree's order in the same pass and stay in it
//     from step to step (the next step's gathers then read nearly sequential
//     addresses, and the walk writes accelerations in place); the input order
//     is restored when an evolve returns;
//  6. node mass and COM bottom up in FP64, children summed in reverse octant
//     order as calc_com: one workgroup per subtree of at most kSubtree nodes
//     (depth by depth under workgroup barriers), then one workgroup for the
//     nodes above those subtrees; the walk records are written on the way;
//  7. the walk is stackless over (node, skip): descend = node + 1, accept or
//     skip = skip[node].  Walks: per lane with the reference's per-body MAC
//     in FP64 (parity), per wave with FP32 monopoles where a node is opened
//     for all 64 lanes when any lane needs it, and the frontier walk (the
//     default FP32 walk, nb_walk_dev.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "nb_sort.hpp"
#include "nb_system.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {
constexpr int kBlock = 256;
constexpr int kLevels = 42;   // octant digits per body
constexpr int kHiLevels = 21; // digits in the high key

enum { BH_ERR_NONE = 0, BH_ERR_PLACEMENT = 1, BH_ERR_INDIST = 2, BH_ERR_DEPTH = 3 };
}  // namespace

#define NB_WAVES_PER_EU(w) __attribute__((amdgpu_waves_per_eu(w)))
#include "nb_walk_dev.hpp"

// A node's FP64 mass and centre of mass in one 32-byte record, written for
// the subtree roots and the nodes above them, so that the top of the tree's
// pass reads each child from one cache line (the SoA arrays take four).
struct __align__(32) MassCom {
    double m, x, y, z;
};
// The nodes above the subtrees, when there are at most kTopOwn of them:
// each one's record by its index in the `top` list (written by k_children),
// and its children's MassCom records at 8 index + rank (k_children for leaf
// children, k_com_subtrees for subtree roots), so the one workgroup that
// sums them reads adjacent lines instead of one line per array and child.
constexpr int kTopOwn = 1024;
struct __align__(16) TopRec {
    double side;
    int32_t skip, first, last;
    uint8_t depth, nc;
    int32_t kids[8];  // children in octant order, -1 past nc
};

struct BHWork {
    int64_t cap_n = 0, cap_nodes = 0;
    DBuf<double> part;      // bbox partials
    DBuf<double> box;       // btm[3], top[3]
    // the build's device scalars, in one array so that one kernel clears them
    // and one copy brings the ones the host needs back: the node count, the
    // error (BH_ERR_*), the long-run flags of the sort levels and the root counts
    DBuf<uint32_t> scal;
    DBuf<uint64_t> khi, klo, klo_slot;  // high keys and low keys in sorted order; low keys by body slot
    DBuf<uint32_t> idx;     // sorted position -> body slot (the identity once the bodies are in tree order)
    DBuf<uint8_t> c;        // shared digits of sorted neighbours
    DBuf<uint8_t> cm1, cm2, cm3;  // minima of c over chunks of 64, 4096 and 262144 positions
    DBuf<uint32_t> cnt;     // nodes starting at s, then their exclusive scan
    DBuf<uint32_t> blk;     // scan block sums
    // the bodies in tree order: slot -> input index, and the other copy of
    // every per-body array
    bool permuted = false;
    bool acc_permuted = false;  // the accelerations too (a walk wrote them)
    DBuf<int32_t> perm, perm_alt;
    DBuf<double> xa, va, aa;
    // masses, radii and restitution in tree order (the context's own arrays
    // stay in input order: Barnes-Hut steps do not change them)
    DBuf<double> mt, rt, et, mt_alt, rt_alt, et_alt;
    bool re_tree = false;  // rt and et hold the current tree order (only the collision pass reads them)
    // nodes (DFS pre-order)
    DBuf<int32_t> first, last, skip, parent, body;
    DBuf<uint8_t> depth, nchild;
    DBuf<int8_t> octant;
    DBuf<double> mass, com, side;  // com: SoA with stride cap_nodes
    DBuf<WalkNode> wnode;   // walk record per node (one 32-byte line), DFS order
    DBuf<WalkNode> wgrp;    // the same records with each node's children adjacent
    DBuf<uint8_t> rank;     // index of a node among its parent's children
    DBuf<uint32_t> cbase;   // exclusive scan of nchild (children of x at 1 + cbase[x])
    DBuf<uint32_t> cbraw;   // nchild, then its scan within blocks of 1024 (cbase less cblk[x >> 10])
    DBuf<uint32_t> slot;    // position of each node in the grouped layout
    DBuf<int32_t> kid;      // node at each position of the grouped layout (children of x: 1 + cbase[x] ...)
    DBuf<uint32_t> cblk, ctot;
    DBuf<int32_t> roots, top;  // subtree roots and the nodes above them (bottom-up pass)
    DBuf<MassCom> mc;          // mass and COM of the subtree roots and the nodes above them
    DBuf<TopRec> ttop;         // records of the nodes above the subtrees (up to kTopOwn)
    DBuf<MassCom> tk;          // their children's mass and COM, 8 per node
    DBuf<int32_t> rdest;       // per subtree root: its record's index in tk, or -1
    DBuf<int32_t> fstack;   // frontier walk: one node stack per wave
    DBuf<int64_t> gxr;      // frontier walk: the XCD ranges of the group order
    DBuf<int> werr;         // frontier walk: a stack would have overflowed
    DBuf<uint32_t> gcost;   // interactions of each 64-body group in the last walk
    DBuf<int32_t> gorder;   // groups of each XCD's range, heaviest first
    int64_t gcost_n = 0;
    DBuf<unsigned long long> icount2;  // count of the gated wave walk
    int bucket = 32;        // rejected nodes with at most this many bodies are summed directly
    DBuf<float4> pos32s;    // body positions (FP32) and G*m in tree order
    DBuf<unsigned long long> icount;
    RadixSorter sorter;
    int64_t nodes = 0;
    uint32_t *h_scal = nullptr;  // kScHost leading entries of scal
    // the octree sort's passes over the high key (7 = all of it), chosen from
    // whether the last build had runs of more than 32 bodies sharing 12, 15
    // or 18 levels
    int sort_passes = 7;
    hipEvent_t ev_count = nullptr;  // the node count, error and flags have reached h_scal
    // The frontier walk's schedule (k_frontier_reset, k_group_order) runs on
    // a second stream while the tree is built: it reads only the previous
    // walk's group costs.  ev_sched_in follows that walk on the context's
    // stream, ev_sched_out the schedule, which the walk waits for.
    hipStream_t sched = nullptr;
    hipEvent_t ev_sched_in = nullptr, ev_sched_out = nullptr;
    // collision pass
    DBuf<double> nbox;      // node corners btm[3], top[3], padded to kBoxStride doubles
    DBuf<int32_t> leaf_of;  // sorted position -> leaf node
    DBuf<uint32_t> bc_cnt, bc_blk, bc_total;  // overlaps per sorted body, then their exclusive scan
    DBuf<int32_t> e_own, e_nb;                // overlaps: owner body, neighbour body
    DBuf<double> e_geo, e_mom;                // distance and normal; axis momenta
    DBuf<int32_t> c_nb;                       // first kCollCache overlaps per body: neighbour,
    DBuf<double> c_geo;                       //   distance and normal (from the counting pass)
    DBuf<double> rmax;                        // largest radius
    DBuf<int32_t> cc_par;                     // union-find parents over bodies
    DBuf<uint64_t> cc_keys;                   // (component, entry) sort keys
    DBuf<uint32_t> cc_vals;
    DBuf<int32_t> g_own, g_nb;                // entries grouped by component
    DBuf<double> g_geo;
    uint32_t *h_bc_total = nullptr;
    ~BHWork() {
        if (sched) {
            (void)hipStreamSynchronize(sched);
            (void)hipStreamDestroy(sched);
        }
        if (ev_sched_in) (void)hipEventDestroy(ev_sched_in);
        if (ev_sched_out) (void)hipEventDestroy(ev_sched_out);
        if (ev_count) (void)hipEventDestroy(ev_count);
        if (h_scal) (void)hipHostFree(h_scal);
        if (h_bc_total) (void)hipHostFree(h_bc_total);
    }
};

void bh_work_free(BHWork *w) { delete w; }

namespace {

// Octree sort passes (see BHWork::sort_passes): the candidates below the full
// 7 are 4, 5 and 6 (12, 15 and 18 shared levels); a candidate is used when
// the last measured runs at its level held at most 32 bodies (k_lcp_counts).
constexpr int kMinSortPasses = 4;
constexpr int kSortStats = 3;
// slots of BHWork::scal
constexpr int kScTotal = 0, kScErr = 1, kScLong = 2, kScRoots = kScLong + kSortStats, kScalars = kScRoots + 2;
constexpr int kScHost = kScRoots;  // total, error and long-run flags come back to the host
// subtrees of at most this many nodes get their centres of mass in one workgroup's LDS
constexpr int32_t kSubtree = 1024;

// ---- 1. bounding cube ----
constexpr int kBoxParts = 16384;  // most partial boxes of k_bbox_partial
// With DRIFT the step's drift x += dt v is applied first (and with KICK the
// previous step's closing kick v += kdt a before it), in the arithmetic of
// the separate kick and drift kernels, so the box is the drifted bodies'.
template <bool KICK, bool DRIFT>
__global__ __launch_bounds__(kBlock) void k_bbox_partial(int64_t n, double *__restrict__ x, double *__restrict__ v,
                                                         const double *__restrict__ a, double kdt, double dt,
                                                         double *__restrict__ part) {
    __shared__ double sm[6][kBlock];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        for (int c = 0; c < 3; ++c) {
            double xc = x[c * n + i];
            if (DRIFT) {
                double vc = v[c * n + i];
                if (KICK) {
                    vc += kdt * a[c * n + i];
                    v[c * n + i] = vc;
                }
                xc += dt * vc;
                x[c * n + i] = xc;
            }
            mn[c] = fmin(mn[c], xc);
            mx[c] = fmax(mx[c], xc);
        }
    for (int c = 0; c < 3; ++c) {
        sm[c][threadIdx.x] = mn[c];
        sm[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < 3; ++c) {
                sm[c][threadIdx.x] = fmin(sm[c][threadIdx.x], sm[c][threadIdx.x + off]);
                sm[3 + c][threadIdx.x] = fmax(sm[3 + c][threadIdx.x], sm[3 + c][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

// find_min_max (greg8tree.hpp:1012-1047): longest edge with the reference's
// tie rule, pad = L/20, btm = min - pad, L' = L + 2*pad, top = btm + L'.
__global__ __launch_bounds__(kBlock) void k_bbox_final(int nparts, const double *__restrict__ part,
                                                       double *__restrict__ box, uint32_t *__restrict__ scal) {
    __shared__ double sm[6][kBlock];
    double lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += kBlock)
        for (int c = 0; c < 3; ++c) {
            lmn[c] = fmin(lmn[c], part[b * 6 + c]);
            lmx[c] = fmax(lmx[c], part[b * 6 + 3 + c]);
        }
    for (int c = 0; c < 3; ++c) {
        sm[c][threadIdx.x] = lmn[c];
        sm[3 + c][threadIdx.x] = lmx[c];
    }
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int c = 0; c < 3; ++c) {
                sm[c][threadIdx.x] = fmin(sm[c][threadIdx.x], sm[c][threadIdx.x + off]);
                sm[3 + c][threadIdx.x] = fmax(sm[3 + c][threadIdx.x], sm[3 + c][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double mn[3] = {sm[0][0], sm[1][0], sm[2][0]}, mx[3] = {sm[3][0], sm[4][0], sm[5][0]};
    double xr = mx[0] - mn[0], yr = mx[1] - mn[1], zr = mx[2] - mn[2];
    double le = xr > yr ? (xr > zr ? xr : zr) : (yr > zr ? yr : zr);
    double pad = le / 20.0;
    for (int c = 0; c < 3; ++c) box[c] = mn[c] - pad;
    le += 2 * pad;
    for (int c = 0; c < 3; ++c) box[3 + c] = box[c] + le;
    for (int q = kScErr; q < kScalars; ++q) scal[q] = 0u;  // the build's counters and flags start at zero
}

// ---- 2. octant digits by the reference's midpoint recursion ----
__global__ __launch_bounds__(kBlock) void k_keys(int64_t n, const double *__restrict__ x,
                                                 const double *__restrict__ box, uint64_t *__restrict__ khi,
                                                 uint64_t *__restrict__ klo, uint32_t *__restrict__ idx,
                                                 int *__restrict__ err) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double p[3], lo[3], hi[3];
    bool inside = true;
    for (int c = 0; c < 3; ++c) {
        p[c] = x[c * n + i];
        lo[c] = box[c];
        hi[c] = box[3 + c];
        inside = inside && lo[c] <= p[c] && p[c] < hi[c];  // between (gregvec.hpp:2115-2120)
    }
    if (!inside) atomicMax(err, BH_ERR_PLACEMENT);
    // one level's octant digit, halving the cell towards p
    auto digit = [&]() {
        unsigned d = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double mid = (lo[c] + hi[c]) * 0.5;  // (btm + top)/2.0l, exact halving
            const bool up = p[c] >= mid;
            d |= (unsigned)up << c;
            lo[c] = up ? mid : lo[c];
            hi[c] = up ? hi[c] : mid;
        }
        return (uint64_t)d;
    };
    uint64_t h = 0, l = 0;
#pragma unroll
    for (int lev = 1; lev <= kHiLevels; ++lev) h = (h << 3) | digit();
#pragma unroll
    for (int lev = kHiLevels + 1; lev <= kLevels; ++lev) l = (l << 3) | digit();
    khi[i] = h;
    klo[i] = l;
    idx[i] = (uint32_t)i;
}

__global__ void k_gather_lo(int64_t n, const uint64_t *__restrict__ klo, const uint32_t *__restrict__ idx,
                            uint64_t *__restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s < n) out[s] = klo[idx[s]];
}

// The sort orders the high key's bits from lo_bit up; runs of bodies equal
// in those bits are put in (high key, low key, index) order here, by one
// insertion sort per run (the runs are short -- the pass count is chosen for
// that -- and nearly sorted, since the input is the previous tree's order).
// With the full high key sorted (lo_bit 0) only bodies sharing 21 levels
// form runs, and they are ordered by the low key.
__global__ void k_fix_ties(int64_t n, int lo_bit, uint64_t *__restrict__ khi, uint64_t *__restrict__ klo,
                           uint32_t *__restrict__ idx) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n - 1) return;
    const uint64_t top = khi[s] >> lo_bit;
    if ((khi[s + 1] >> lo_bit) != top || (s > 0 && (khi[s - 1] >> lo_bit) == top)) return;
    int64_t e = s + 1;
    while (e + 1 < n && (khi[e + 1] >> lo_bit) == top) ++e;
    for (int64_t a = s + 1; a <= e; ++a) {
        const uint64_t kh = khi[a], kl = klo[a];
        const uint32_t id = idx[a];
        int64_t b = a - 1;
        while (b >= s && (khi[b] > kh || (khi[b] == kh && (klo[b] > kl || (klo[b] == kl && idx[b] > id))))) {
            khi[b + 1] = khi[b];
            klo[b + 1] = klo[b];
            idx[b + 1] = idx[b];
            --b;
        }
        khi[b + 1] = kh;
        klo[b + 1] = kl;
        idx[b + 1] = id;
    }
}


__device__ __forceinline__ int shared_digits(uint64_t h0, uint64_t l0, uint64_t h1, uint64_t l1) {
    if (h0 != h1) return (__clzll(h0 ^ h1) - 1) / 3;  // keys use the low 63 bits
    if (l0 != l1) return kHiLevels + (__clzll(l0 ^ l1) - 1) / 3;
    return kLevels;
}

// ---- 4. shared digits, node counts ----
// c[s] = octant digits shared by sorted neighbours s and s+1 (0 for the
// last); the nodes starting at s are the levels (c[s-1], c[s]] plus the leaf
// of s, so cnt[s] = 1 + max(0, c[s] - c[s-1]).  Each thread also recomputes
// c[s-1] (the same function of the same keys) so that one launch does both.
//
// For the next builds' sort it also flags, per level L = 3P (P = 4..6
// passes), whether some aligned 64-position chunk holds 16 consecutive
// positions with c >= L: a run of bodies sharing L levels longer than 32 has
// 32 consecutive such positions, 16 of them inside one chunk, so no flag
// means every run at that level holds at most 32 bodies.
__global__ __launch_bounds__(kBlock) void k_lcp_counts(int64_t n, const uint64_t *__restrict__ khi,
                                                       const uint64_t *__restrict__ klo, const uint32_t *__restrict__ idx,
                                                       const double *__restrict__ x, uint8_t *__restrict__ c,
                                                       uint8_t *__restrict__ cm1, uint32_t *__restrict__ cnt,
                                                       int *__restrict__ err, uint32_t *__restrict__ longrun) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int cs = 255;  // past the end: no effect on the minima below
    if (s < n) {
        cs = 0;
        if (s < n - 1) {
            cs = shared_digits(khi[s], klo[s], khi[s + 1], klo[s + 1]);
            if (cs >= kLevels) {
                const uint32_t a = idx[s], b = idx[s + 1];
                const bool same = x[a] == x[b] && x[n + a] == x[n + b] && x[2 * n + a] == x[2 * n + b];
                atomicMax(err, same ? BH_ERR_INDIST : BH_ERR_DEPTH);  // indistinguishable_bodies (greg8tree.hpp:171)
                cs = kLevels - 1;
            }
        }
        int cp = 0;
        if (s > 0) cp = min(kLevels - 1, shared_digits(khi[s - 1], klo[s - 1], khi[s], klo[s]));
        c[s] = (uint8_t)cs;
        cnt[s] = 1u + (uint32_t)max(0, cs - cp);
    }
    // long runs at the candidate sort levels (cs is 255 past the end)
#pragma unroll
    for (int q = 0; q < kSortStats; ++q) {
        const int L = 3 * (kMinSortPasses + q);
        uint64_t r = __ballot(s < n && cs >= L);
        r &= r >> 1;
        r &= r >> 2;
        r &= r >> 4;
        r &= r >> 8;  // bit i: positions i .. i+15 of the chunk all share L levels
        if (r != 0 && (threadIdx.x & 63) == 0) longrun[q] = 1u;
    }
    // minimum of c over each aligned chunk of 64 positions (one wave)
    int m = cs;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0 && s < n) cm1[s >> 6] = (uint8_t)m;
}

// Exclusive scan in three launches (block scan, scan of block sums, add).
// Exclusive scan over a 1024-thread block: wave scans by shuffles, then the
// 16 wave totals scanned by the first wave (two barriers).  Returns the
// exclusive prefix of v within the block; *block_total gets the block's sum.
__device__ __forceinline__ uint32_t block_scan_1024(uint32_t v, uint32_t *s_w, uint32_t *block_total) {
    const unsigned lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= (unsigned)o) x += y;
    }
    if (lane == 63u) s_w[wv] = x;
    __syncthreads();
    if (wv == 0) {
        uint32_t t = lane < 16u ? s_w[lane] : 0u;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t y = __shfl_up(t, o);
            if (lane >= (unsigned)o) t += y;
        }
        if (lane < 16u) s_w[lane] = t;  // inclusive over the waves (every lane read before any wrote)
    }
    __syncthreads();
    *block_total = s_w[15];
    return (wv ? s_w[wv - 1] : 0u) + x - v;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(int64_t n, uint32_t *__restrict__ a, uint32_t *__restrict__ blk) {
    __shared__ uint32_t s_w[16];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    const uint32_t v = i < n ? a[i] : 0;
    uint32_t tot;
    const uint32_t ex = block_scan_1024(v, s_w, &tot);
    if (i < n) a[i] = ex;
    if (threadIdx.x == 1023) blk[blockIdx.x] = tot;
}

// Minimum of the 64 bytes a[64 i ...] below `valid` (255 past it); the
// array is padded to whole chunks of 64, so four 16-byte loads stay inside.
__device__ __forceinline__ uint8_t chunk_min64(const uint8_t *__restrict__ a, int64_t i, int64_t valid) {
    const uint4 *p = reinterpret_cast<const uint4 *>(a + (i << 6));
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3];
    const uint32_t wd[16] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                             w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
    const int64_t hi = valid - (i << 6);
    unsigned m = 255u;
#pragma unroll
    for (int k = 0; k < 64; ++k)
        if (k < hi) m = min(m, (wd[k >> 2] >> (8 * (k & 3))) & 0xFFu);
    return (uint8_t)m;
}

// The scan of the block sums; with cm1 also the minima of c over chunks of
// 4096 positions (cm2, from 64 entries of cm1 each) and of 262144 (cm3).
__global__ __launch_bounds__(1024) void k_scan_sums(int nb, uint32_t *__restrict__ blk, uint32_t *__restrict__ total,
                                                    const uint8_t *__restrict__ cm1 = nullptr, int64_t n1 = 0,
                                                    uint8_t *__restrict__ cm2 = nullptr, int64_t n2 = 0,
                                                    uint8_t *__restrict__ cm3 = nullptr, int64_t n3 = 0) {
    __shared__ uint32_t s_w[16];
    // sequential chunks per thread: nb may exceed 1024
    const int per = (nb + 1023) / 1024;
    const int b = threadIdx.x * per, e = min(nb, b + per);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += blk[i];
    uint32_t tot;
    uint32_t run = block_scan_1024(s, s_w, &tot);
    for (int i = b; i < e; ++i) {
        uint32_t v = blk[i];
        blk[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023) *total = 1u + tot;  // + root
    if (!cm1) return;
    for (int64_t i = threadIdx.x; i < n2; i += 1024) cm2[i] = chunk_min64(cm1, i, n1);
    __threadfence_block();
    __syncthreads();
    for (int64_t i = threadIdx.x; i < n3; i += 1024) cm3[i] = chunk_min64(cm2, i, n2);
}

__global__ __launch_bounds__(1024) void k_scan_add(int64_t n, uint32_t *__restrict__ a, const uint32_t *__restrict__ blk) {
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    if (i < n) a[i] += blk[blockIdx.x];
}

__device__ __forceinline__ unsigned digit_at(uint64_t hi, uint64_t lo, int level) {
    return level <= kHiLevels ? (unsigned)(hi >> (3 * (kHiLevels - level))) & 7u
                              : (unsigned)(lo >> (3 * (kLevels - level))) & 7u;
}

// Over sorted keys the digits shared by positions a < b are the minimum of c
// over [a, b), so the run of bodies sharing at least L digits with s ends at
// the first t >= s with c[t] < L (c[n-1] = 0 ends every run).  That search
// goes up a hierarchy of minima -- cm1 over chunks of 64 positions, cm2 over
// 4096, cm3 over 262144 -- and back down, reading 64 entries (four 16-byte
// loads) per step, so a long run costs a handful of loads instead of a
// search over the keys.

// First index in [start, 64 (chunk + 1)) of chunk `chunk` of a with
// a[i] < L (1 <= L <= 42), or -1.  Arrays are padded to whole chunks.  The
// padding needs no mask: a search starts at or before the last valid entry,
// and the entry of position n - 1 (c = 0, and the minimum 0 of every chunk
// holding it) lies before any padding, so the first entry below L is valid.
__device__ __forceinline__ int64_t first_below_in(const uint8_t *__restrict__ a, int64_t chunk, int64_t start, int L) {
    const int64_t b = chunk << 6;
    const uint4 *p = reinterpret_cast<const uint4 *>(a + b);
    const uint4 w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3];
    const uint32_t wd[16] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w,
                             w2.x, w2.y, w2.z, w2.w, w3.x, w3.y, w3.z, w3.w};
    const int lo = (int)max<int64_t>(0, start - b);
    const uint32_t sub = 0x01010101u * (uint32_t)L;
    int fq = -1;
    uint32_t fw = 0u;
#pragma unroll
    for (int q = 15; q >= 0; --q) {
        // per byte, (x | 0x80) - L keeps its high bit unless x < L (x < 128,
        // so no byte borrows from the next one); bytes before lo are dropped
        uint32_t f = ~((wd[q] | 0x80808080u) - sub) & 0x80808080u;
        const int sh = lo - 4 * q;
        f = sh <= 0 ? f : sh >= 4 ? 0u : f & (~0u << (8 * sh));
        if (f) {
            fw = f;
            fq = q;
        }
    }
    return fq < 0 ? -1 : b + 4 * fq + (__builtin_ctz(fw) >> 3);
}

struct RunIndex {
    const uint8_t *c, *cm1, *cm2, *cm3;
    int64_t n, n1, n2, n3;
};

// First position t >= s with c[t] < L (L >= 1), which is the last position
// of the run sharing at least L digits with s when c[s - 1] < L: up the
// levels through the rest of each chunk, then down inside the chunk found
// (one inlined chunk scan, so the code stays small).
__device__ int64_t run_end(const RunIndex &R, int64_t s, int L) {
    int lev = 0;
    int64_t start = s, chunk = s >> 6, found;
    for (;;) {
        const uint8_t *a = lev == 0 ? R.c : lev == 1 ? R.cm1 : lev == 2 ? R.cm2 : R.cm3;
        if (found = first_below_in(a, chunk, start, L); found >= 0) {
            if (lev == 0) return found;
            break;
        }
        if (lev < 3) {  // the entries after this chunk's at the next level up
            start = chunk + 1;
            chunk >>= 6;
            ++lev;
        } else {  // along the top level
            ++chunk;
            start = chunk << 6;
            if (start >= R.n3) return R.n - 1;  // unreachable: the chunk holding n-1 has minimum 0
        }
    }
    while (lev > 0) {  // down: the first entry below L inside the chunk found
        --lev;
        const uint8_t *a = lev == 0 ? R.c : lev == 1 ? R.cm1 : lev == 2 ? R.cm2 : R.cm3;
        found = first_below_in(a, found, found << 6, L);
        if (found < 0) return R.n - 1;  // unreachable: that chunk's minimum is below L
    }
    return found;
}


// k_build_nodes' handling of radii and restitutions
constexpr int kReSkip = 0, kReFromPrevious = 1, kReFromInput = 2;

// ---- nodes in DFS pre-order, and the bodies moved into tree order ----
// Node numbering: the root is 0 and the nodes starting at sorted position s
// are 1 + base[s] + j, internal levels (c[s-1], c[s]] first, then the leaf
// of s, so an exclusive scan of cnt numbers every node in pre-order with
// octants 0..7 (the order cell_iterator visits them).  base[s] holds that
// scan within s's block of 1024 and bblk[s >> 10] the blocks before it (the
// readers add the two; no pass writes the sum).  One thread per s
// writes the nodes starting at s:
//  * first = s, last = the end of the run sharing L digits (a node's bodies
//    are contiguous), skip = the first node starting at last + 1 (the next
//    node outside its subtree);
//  * depth, octant, side (top.x - btm.x replayed by set_corners' halving on
//    x along the key's digits);
//  * for the leaf: body = s and the leaf's mass and COM.
// It runs before the host knows the node count: with more nodes than the
// arrays hold (cap) it writes nothing, and the host launches it again.
// The bodies themselves move into tree order in the same pass (every
// per-body array, gathered through idx), with the FP32 (x, y, z, G m)
// stream of the walk; idx becomes the identity.  Between steps the bodies
// stay in the previous step's tree order, so these gathers read nearly
// sequential addresses.
__global__ __launch_bounds__(kBlock) void k_build_nodes(
    int64_t n, const RunIndex R, const uint32_t *__restrict__ base, const uint32_t *__restrict__ total_p,
    const uint32_t *__restrict__ bblk, const uint64_t *__restrict__ khi, const uint64_t *__restrict__ klo,
    const double *__restrict__ box, uint32_t *__restrict__ idx, const double *__restrict__ x, const double *__restrict__ v,
    const double *__restrict__ m, const double *__restrict__ r, const double *__restrict__ e, int re_mode,
    const int32_t *__restrict__ perm, double *__restrict__ xo, double *__restrict__ vo, double *__restrict__ mo,
    double *__restrict__ ro, double *__restrict__ eo, int32_t *__restrict__ permo, float4 *__restrict__ pos32,
    double G, int32_t *__restrict__ first, int32_t *__restrict__ last, int32_t *__restrict__ skip,
    int32_t *__restrict__ body, uint8_t *__restrict__ depth,
    int8_t *__restrict__ octant, double *__restrict__ side, double *__restrict__ mass, double *__restrict__ com,
    int64_t cstride) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int32_t total = (int32_t)*total_p;
    if (s >= n || (int64_t)total > cstride) return;  // cstride is the node capacity
    // bodies into tree order
    const uint32_t b = idx[s];
    const double px = x[b], py = x[n + b], pz = x[2 * n + b], pm = m[b];
    xo[s] = px; xo[n + s] = py; xo[2 * n + s] = pz;
    vo[s] = v[b]; vo[n + s] = v[n + b]; vo[2 * n + s] = v[2 * n + b];
    mo[s] = pm;
    const int32_t in = perm ? perm[b] : (int32_t)b;
    permo[s] = in;
    // radii and restitutions (read only by the collision pass): not moved,
    // moved along from the previous tree order, or taken from input order
    if (re_mode != kReSkip) {
        const int32_t q = re_mode == kReFromInput ? in : (int32_t)b;
        ro[s] = r[q];
        eo[s] = e[q];
    }
    pos32[s] = make_float4((float)px, (float)py, (float)pz, (float)(G * pm));
    idx[s] = (uint32_t)s;
    // nodes starting at s
    const int cs = R.c[s], cp = s > 0 ? R.c[s - 1] : 0;
    const uint64_t h = khi[s], l = klo[s];
    const int32_t n0 = 1 + (int32_t)(base[s] + bblk[s >> 10]);
    double lo = box[0], hi = box[3];
    if (s == 0) {
        first[0] = 0;
        last[0] = (int32_t)(n - 1);
        skip[0] = total;
        body[0] = -1;
        depth[0] = 0;
        octant[0] = -1;
        side[0] = hi - lo;
    }
    // run ends, deepest level first: the run at level L holds the one at
    // L + 1, which ends at re with c[re] <= L, so it ends at re too unless
    // c[re] == L, and then at the first t > re with c[t] < L
    int64_t re = -1;
    for (int L = cs; L > cp; --L) {
        if (re < 0 || R.c[re] == L) re = run_end(R, re < 0 ? s : re + 1, L);
        const int32_t node = n0 + (L - cp - 1);
        last[node] = (int32_t)re;
        skip[node] = re >= n - 1 ? total : 1 + (int32_t)(base[re + 1] + bblk[(re + 1) >> 10]);
    }
    int lev = 0;
    for (int L = cp + 1; L <= cs; ++L) {
        for (; lev < L; ++lev) {
            const double mid = (lo + hi) * 0.5;
            const bool up = digit_at(h, l, lev + 1) & 1u;
            lo = up ? mid : lo;
            hi = up ? hi : mid;
        }
        const int32_t node = n0 + (L - cp - 1);
        first[node] = (int32_t)s;
        body[node] = -1;
        depth[node] = (uint8_t)L;
        octant[node] = (int8_t)digit_at(h, l, L);
        side[node] = hi - lo;
    }
    const int ld = max(cs, cp) + 1;
    for (; lev < ld; ++lev) {
        const double mid = (lo + hi) * 0.5;
        const bool up = digit_at(h, l, lev + 1) & 1u;
        lo = up ? mid : lo;
        hi = up ? hi : mid;
    }
    const int32_t leaf = n0 + max(0, cs - cp);
    first[leaf] = (int32_t)s;
    last[leaf] = (int32_t)s;
    skip[leaf] = leaf + 1;
    body[leaf] = (int32_t)s;
    depth[leaf] = (uint8_t)ld;
    octant[leaf] = (int8_t)digit_at(h, l, ld);
    side[leaf] = hi - lo;
    mass[leaf] = pm;
    com[leaf] = px;
    com[cstride + leaf] = py;
    com[2 * cstride + leaf] = pz;
}

// Children of each internal node, in octant order along the skip chain:
// their count, each child's index among its siblings, and its parent.  The same pass lists the nodes above the subtrees (more
// than kSubtree nodes: `top`) and the subtree roots (internal nodes of at
// most kSubtree nodes under such a node, or the root itself when the whole
// tree is one subtree): a top node lists itself and its root children, and
// writes the mass and COM record of its leaf children, which no subtree
// pass sums.  The first kTopOwn top nodes also write their TopRec, their
// leaf children's records into tk, and each root child's tk index into
// rdest.  The lists take one atomic per workgroup and list (returning
// atomics on one address serialise in the L2, about 10 ns each); their order
// does not matter to the passes that read them.
__global__ __launch_bounds__(kBlock) void k_children(int32_t total, const int32_t *__restrict__ skip,
                                                     const int32_t *__restrict__ body, uint8_t *__restrict__ nchild,
                                                     uint8_t *__restrict__ rank, uint32_t *__restrict__ cbase,
                                                     int32_t *__restrict__ parent, int32_t *__restrict__ roots,
                                                     uint32_t *__restrict__ nroots, int32_t *__restrict__ top,
                                                     uint32_t *__restrict__ ntop, const double *__restrict__ mass,
                                                     const double *__restrict__ com, int64_t cstride,
                                                     MassCom *__restrict__ mc, const double *__restrict__ side,
                                                     const int32_t *__restrict__ first,
                                                     const int32_t *__restrict__ last,
                                                     const uint8_t *__restrict__ depth, TopRec *__restrict__ ttop,
                                                     MassCom *__restrict__ tk, int32_t *__restrict__ rdest) {
    __shared__ uint32_t s_nt, s_nr, s_bt, s_br;
    const int32_t x = blockIdx.x * kBlock + threadIdx.x;
    if (threadIdx.x == 0) {
        s_nt = 0u;
        s_nr = 0u;
    }
    __syncthreads();
    bool is_top = false;
    int nr = 0, nk = 0;
    int32_t rc[8], rk[8];
    if (x < total) {
        if (x == 0) {
            rank[0] = 0;
            parent[0] = -1;
        }
        if (body[x] >= 0) {
            nchild[x] = 0;
            cbase[x] = 0;
        } else {
            const int32_t end = min(skip[x], total);
            is_top = end - x > kSubtree;
            if (x == 0 && !is_top) {  // the whole tree is one subtree
                rk[nr] = -1;
                rc[nr++] = 0;
            }
            int32_t ch = x + 1, k = 0;
            while (ch < end && k < 8) {  // an octree node has at most 8 children
                rank[ch] = (uint8_t)k;
                parent[ch] = x;
                const int32_t next = max(skip[ch], ch + 1);
                if (is_top) {
                    if (body[ch] < 0) {
                        if (next - ch <= kSubtree) {
                            rk[nr] = k;
                            rc[nr++] = ch;
                        }
                    } else {
                        MassCom v;
                        v.m = mass[ch];
                        v.x = com[ch];
                        v.y = com[cstride + ch];
                        v.z = com[2 * cstride + ch];
                        mc[ch] = v;
                    }
                }
                ch = next;
                ++k;
            }
            nchild[x] = (uint8_t)k;
            cbase[x] = (uint32_t)k;  // scanned into the first child's slot
            nk = k;
        }
    }
    const uint32_t lt = is_top ? atomicAdd(&s_nt, 1u) : 0u;
    const uint32_t lr = nr ? atomicAdd(&s_nr, (uint32_t)nr) : 0u;
    __syncthreads();
    if (threadIdx.x == 0) {
        s_bt = s_nt ? atomicAdd(ntop, s_nt) : 0u;
        s_br = s_nr ? atomicAdd(nroots, s_nr) : 0u;
    }
    __syncthreads();
    const uint32_t ti = s_bt + lt;
    const bool own = is_top && ti < (uint32_t)kTopOwn;
    if (is_top) top[ti] = x;
    for (int i = 0; i < nr; ++i) {
        roots[s_br + lr + i] = rc[i];
        rdest[s_br + lr + i] = own && rk[i] >= 0 ? (int32_t)(8 * ti) + rk[i] : -1;
    }
    if (own) {  // the children again (top nodes are few): the record, and the leaves' sums
        TopRec tr;
        tr.side = side[x];
        tr.skip = skip[x];
        tr.first = first[x];
        tr.last = last[x];
        tr.depth = depth[x];
        tr.nc = (uint8_t)nk;
        int32_t ch = x + 1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            tr.kids[k] = k < nk ? ch : -1;
            if (k < nk) {
                if (body[ch] >= 0) {
                    MassCom v;
                    v.m = mass[ch];
                    v.x = com[ch];
                    v.y = com[cstride + ch];
                    v.z = com[2 * cstride + ch];
                    tk[8 * ti + k] = v;
                }
                ch = max(skip[ch], ch + 1);
            }
        }
        ttop[ti] = tr;
    }
}

// doubles per node in the corner array: a node's six corners in one aligned 64-byte segment
constexpr int kBoxStride = 8;

// ---- collision pass on the tree: system::bh_s_coll (gregsys.hpp:631-686) ----
// Node cube corners (btm xyz, top xyz), replaying set_corners' recursion on
// all three axes along the node's digits (the same FP64 midpoints as k_keys).
__global__ void k_node_box(int32_t total, const double *__restrict__ box, const int32_t *__restrict__ first,
                           const uint8_t *__restrict__ depth, const uint64_t *__restrict__ khi,
                           const uint64_t *__restrict__ klo, double *__restrict__ nbox) {
    const int32_t x = blockIdx.x * kBlock + threadIdx.x;
    if (x >= total) return;
    const int32_t s = first[x];
    const uint64_t h = khi[s], l = klo[s];
    double lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
    const int dep = depth[x];
    for (int L = 1; L <= dep; ++L) {
        const unsigned d = digit_at(h, l, L);
        for (int c = 0; c < 3; ++c) {
            double mid = (lo[c] + hi[c]) * 0.5;
            bool up = (d >> c) & 1u;
            lo[c] = up ? mid : lo[c];
            hi[c] = up ? hi[c] : mid;
        }
    }
    for (int c = 0; c < 3; ++c) {
        nbox[(size_t)kBoxStride * x + c] = lo[c];
        nbox[(size_t)kBoxStride * x + 3 + c] = hi[c];
    }
}

__global__ void k_leaf_of(int32_t total, const int32_t *__restrict__ body, const int32_t *__restrict__ first,
                          int32_t *__restrict__ leaf_of) {
    const int32_t x = blockIdx.x * kBlock + threadIdx.x;
    if (x < total && body[x] >= 0) leaf_of[first[x]] = x;
}

// nn_iterator (greg8tree.hpp:615-700) for the body at sorted position s: the
// leaves whose closed box touches the own leaf's box, in DFS pre-order with
// octants 0..7 (stackless over the skip pointers from the root's first child),
// the own leaf skipped.  For each, the reference's overlap test in FP64:
// r_o + r_j > |p_j - p_o|, |.| the square root of the ordered sum of squares,
// n = (p_j - p_o)/|.|.  FILL = false counts the overlaps of s; FILL = true
// writes them, in visit order, from out_off[s].
// A subtree whose box lies farther from p_o than r_o + r_max (with margins far
// above the rounding of |.| and of the gaps) holds no body that can overlap o, so
// it is skipped: only non-overlapping leaves drop out of the walk, and the
// overlaps and their order are those of the full walk.
// The counting pass also keeps each body's first kCollCache overlaps (in
// visit order, with the same arithmetic as the writing pass); the writing
// pass copies them for bodies with at most that many and walks again only
// for the others.
constexpr int kCollCache = 4;

template <bool FILL>
__global__ __launch_bounds__(kBlock) void k_bhc_scan(int64_t n, int32_t total, const uint32_t *__restrict__ idx,
                                                     const int32_t *__restrict__ leaf_of,
                                                     const double *__restrict__ nbox, const int32_t *__restrict__ skip,
                                                     const int32_t *__restrict__ body,
                                                     const int32_t *__restrict__ parent, const double *__restrict__ x,
                                                     const double *__restrict__ r, const double *__restrict__ rmax_p,
                                                     uint32_t *__restrict__ cnt, uint32_t nent,
                                                     int32_t *__restrict__ e_own, int32_t *__restrict__ e_nb,
                                                     double *__restrict__ e_geo, int32_t *__restrict__ c_nb,
                                                     double *__restrict__ c_geo) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int32_t o = (int32_t)idx[s];
    if (FILL) {
        const uint32_t c0 = cnt[s], c1 = s + 1 < n ? cnt[s + 1] : nent;
        if (c1 == c0) return;  // no overlaps: nothing to write
        if (c1 - c0 <= (uint32_t)kCollCache) {
            for (uint32_t q = 0; q < c1 - c0; ++q) {
                const size_t src = (size_t)s * kCollCache + q, dst = (size_t)c0 + q;
                e_own[dst] = o;
                e_nb[dst] = c_nb[src];
                for (int c = 0; c < 4; ++c) e_geo[4 * dst + c] = c_geo[4 * src + c];
            }
            return;
        }
    }
    const int32_t own = leaf_of[s];
    double bb[6];
    for (int k = 0; k < 6; ++k) bb[k] = nbox[(size_t)kBoxStride * own + k];
    const double xo = x[o], yo = x[n + o], zo = x[2 * n + o], ro = r[o];
    // margins: relative to the radii, and relative to the coordinates (the
    // rounding of the box gaps below)
    const double reach = (ro + *rmax_p) * (1.0 + 1e-9) + 1e-9 * (fabs(xo) + fabs(yo) + fabs(zo));
    const double reach2 = reach * reach;
    // every other body lies outside the own leaf's cube (one body per leaf),
    // so a body farther than `reach` from all six faces overlaps nothing
    const double face = fmin(fmin(fmin(xo - bb[0], bb[3] - xo), fmin(yo - bb[1], bb[4] - yo)),
                             fmin(zo - bb[2], bb[5] - zo));
    if (face > reach) {
        if (!FILL) cnt[s] = 0u;
        return;
    }
    // every body within reach lies in the lowest ancestor whose cube holds the
    // whole reach box around p_o, so only that subtree is walked (in the same
    // pre-order, so the overlaps come in the same order)
    int32_t top = parent[own];
    while (top > 0) {
        const double *cb = nbox + (size_t)kBoxStride * top;
        if (xo - cb[0] > reach && cb[3] - xo > reach && yo - cb[1] > reach && cb[4] - yo > reach &&
            zo - cb[2] > reach && cb[5] - zo > reach)
            break;
        top = parent[top];
    }
    const int32_t end = top > 0 ? skip[top] : total;
    uint32_t k = FILL ? cnt[s] : 0u;
    int32_t p = top + 1;
    while (p < end) {
        if (p == own) {
            p = skip[p];
            continue;
        }
        const double *cb = nbox + (size_t)kBoxStride * p;
        const bool touch = !(bb[3] < cb[0] || bb[0] > cb[3] || bb[4] < cb[1] || bb[1] > cb[4] || bb[5] < cb[2] ||
                             bb[2] > cb[5]);
        const double gx = fmax(fmax(cb[0] - xo, xo - cb[3]), 0.0);
        const double gy = fmax(fmax(cb[1] - yo, yo - cb[4]), 0.0);
        const double gz = fmax(fmax(cb[2] - zo, zo - cb[5]), 0.0);
        if (!touch || gx * gx + gy * gy + gz * gz > reach2) {
            p = skip[p];
            continue;
        }
        const int32_t j = body[p];
        if (j < 0) {  // internal: descend
            ++p;
            continue;
        }
        const double rad = ro + r[j];
        const double dx = x[j] - xo, dy = x[n + j] - yo, dz = x[2 * n + j] - zo;
        const double dist = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
        if (rad > dist) {
            if (FILL) {
                e_own[k] = o;
                e_nb[k] = j;
                e_geo[(size_t)4 * k + 0] = dist;
                e_geo[(size_t)4 * k + 1] = dx / dist;
                e_geo[(size_t)4 * k + 2] = dy / dist;
                e_geo[(size_t)4 * k + 3] = dz / dist;
            } else if (k < (uint32_t)kCollCache) {
                const size_t c = (size_t)s * kCollCache + k;
                c_nb[c] = j;
                c_geo[4 * c + 0] = dist;
                c_geo[4 * c + 1] = dx / dist;
                c_geo[4 * c + 2] = dy / dist;
                c_geo[4 * c + 3] = dz / dist;
            }
            ++k;
        }
        p = skip[p];
    }
    if (!FILL) cnt[s] = k;
}

// Largest radius (for the walk's reach), over a grid: radii are not
// negative, so the IEEE bit patterns order like the values and one integer
// atomicMax per workgroup combines them (*out starts at +0.0).
__global__ __launch_bounds__(kBlock) void k_rmax(int64_t n, const double *__restrict__ r, double *__restrict__ out) {
    __shared__ double sm[kBlock];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        m = fmax(m, r[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
        __syncthreads();
    }
    // -0.0 (allowed as a radius) would order above every positive pattern
    const double top = sm[0] > 0.0 ? sm[0] : 0.0;
    if (threadIdx.x == 0)
        atomicMax(reinterpret_cast<unsigned long long *>(out), (unsigned long long)__double_as_longlong(top));
}

// Connected components of the overlap graph (owner o -- neighbour j), by a
// lock-free union-find over the bodies that have overlaps: roots are hooked
// under smaller roots with a compare-and-swap, so parent indices only
// decrease and every read, stale or not, names an ancestor.
__device__ __forceinline__ int32_t cc_parent(int32_t *par, int32_t x) {
    return __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ int32_t cc_find(int32_t *par, int32_t x) {
    while (true) {
        const int32_t p = cc_parent(par, x);
        if (p == x) return x;
        const int32_t g = cc_parent(par, p);
        if (g != p) __hip_atomic_store(par + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // halve the path
        x = p;
    }
}

__global__ void k_cc_init(uint32_t nent, const int32_t *__restrict__ e_own, const int32_t *__restrict__ e_nb,
                          int32_t *__restrict__ par) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nent) return;
    par[e_own[k]] = e_own[k];
    par[e_nb[k]] = e_nb[k];
}

__global__ void k_cc_union(uint32_t nent, const int32_t *__restrict__ e_own, const int32_t *__restrict__ e_nb,
                           int32_t *par) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nent) return;
    int32_t a = cc_find(par, e_own[k]), b = cc_find(par, e_nb[k]);
    while (a != b) {
        if (a < b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        int32_t expected = a;
        if (__hip_atomic_compare_exchange_strong(par + a, &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            break;  // root a now hangs under b
        a = cc_find(par, expected);
        b = cc_find(par, b);
    }
}

// key = (component root << 32) | entry index: a stable order that keeps each
// component's entries together in owner (leaf) order and visit order.
__global__ void k_cc_keys(uint32_t nent, const int32_t *__restrict__ e_own, int32_t *par, uint64_t *__restrict__ keys,
                          uint32_t *__restrict__ vals) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= nent) return;
    keys[k] = ((uint64_t)(uint32_t)cc_find(par, e_own[k]) << 32) | k;
    vals[k] = k;
}

__global__ void k_cc_gather(uint32_t nent, const uint32_t *__restrict__ vals, const int32_t *__restrict__ e_own,
                            const int32_t *__restrict__ e_nb, const double *__restrict__ e_geo,
                            int32_t *__restrict__ g_own, int32_t *__restrict__ g_nb, double *__restrict__ g_geo) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= nent) return;
    const uint32_t k = vals[q];
    g_own[q] = e_own[k];
    g_nb[q] = e_nb[k];
    for (int c = 0; c < 4; ++c) g_geo[4 * (size_t)q + c] = e_geo[4 * (size_t)k + c];
}

// The bounces in order, one lane per component of the overlap graph: owners
// in leaf order (sorted order), each owner's overlaps sorted by distance with
// an equal distance keeping the first found (the ordered map's emplace), axis
// momenta m v . n taken from the velocities as they stand when the owner is
// reached (its scan reads them all before any of its bounces), then s_coll's
// bounce arithmetic (gregsys.hpp:660-684).  Owners of different components
// touch disjoint bodies, so running components side by side gives the serial
// result; within a component the order is the serial one.
__global__ void k_bhc_replay(int64_t n, uint32_t nent, const uint64_t *__restrict__ keys,
                             const int32_t *__restrict__ e_own, int32_t *__restrict__ e_nb,
                             double *__restrict__ e_geo, double *__restrict__ e_mom, const double *__restrict__ m,
                             const double *__restrict__ rc, double *__restrict__ v) {
    const uint32_t q0 = blockIdx.x * kBlock + threadIdx.x;
    if (q0 >= nent) return;
    const uint32_t comp = (uint32_t)(keys[q0] >> 32);
    if (q0 > 0 && (uint32_t)(keys[q0 - 1] >> 32) == comp) return;  // not the first entry of its component
    uint32_t cend = q0 + 1;
    while (cend < nent && (uint32_t)(keys[cend] >> 32) == comp) ++cend;
    uint32_t k = q0;
    while (k < cend) {
        const int32_t o = e_own[k];
        uint32_t ke = k + 1;
        while (ke < cend && e_own[ke] == o) ++ke;
        for (uint32_t a = k + 1; a < ke; ++a) {  // stable insertion sort by distance
            const int32_t jn = e_nb[a];
            double g[4];
            for (int c = 0; c < 4; ++c) g[c] = e_geo[4 * (size_t)a + c];
            uint32_t b = a;
            while (b > k && e_geo[4 * (size_t)(b - 1)] > g[0]) {
                e_nb[b] = e_nb[b - 1];
                for (int c = 0; c < 4; ++c) e_geo[4 * (size_t)b + c] = e_geo[4 * (size_t)(b - 1) + c];
                --b;
            }
            e_nb[b] = jn;
            for (int c = 0; c < 4; ++c) e_geo[4 * (size_t)b + c] = g[c];
        }
        const double mo = m[o], eo = rc[o];
        double vo[3] = {v[o], v[n + o], v[2 * n + o]};
        for (uint32_t q = k; q < ke; ++q) {
            const int32_t j = e_nb[q];
            const double *g = e_geo + 4 * (size_t)q;
            const double mj = m[j];
            e_mom[2 * (size_t)q] = (mo * vo[0]) * g[1] + (mo * vo[1]) * g[2] + (mo * vo[2]) * g[3];
            e_mom[2 * (size_t)q + 1] = (mj * v[j]) * g[1] + (mj * v[n + j]) * g[2] + (mj * v[2 * n + j]) * g[3];
        }
        for (uint32_t q = k; q < ke; ++q) {
            const double *g = e_geo + 4 * (size_t)q;
            if (q > k && g[0] == e_geo[4 * (size_t)(q - 1)]) continue;  // equal key: the first one stays
            const int32_t j = e_nb[q];
            const double amo = e_mom[2 * (size_t)q], ami = e_mom[2 * (size_t)q + 1];
            const double mj = m[j];
            const double o_vel = amo / mo;
            const double i_vel = ami / mj;
            const double o_minus_i = o_vel - i_vel;
            const double avg_rest = (eo + rc[j]) / 2;
            if (o_vel > 0 || i_vel < 0) {
                if (o_vel <= 0 && i_vel < 0)
                    if (o_vel < i_vel) continue;
                if (o_vel > 0 && i_vel >= 0)
                    if (o_vel < i_vel) continue;
                const double tot = amo + ami;
                const double ms = mo + mj;
                const double new_o = (tot - mj * avg_rest * o_minus_i) / ms;
                const double new_i = (tot + mo * avg_rest * o_minus_i) / ms;
                const double fo = new_o - o_vel, fi = new_i - i_vel;
                vo[0] += fo * g[1]; vo[1] += fo * g[2]; vo[2] += fo * g[3];
                v[j] += fi * g[1]; v[n + j] += fi * g[2]; v[2 * n + j] += fi * g[3];
            }
        }
        v[o] = vo[0]; v[n + o] = vo[1]; v[2 * n + o] = vo[2];
        k = ke;
    }
}

// ---- 5. mass and COM bottom up ----
// Walk records of node x: q = (com, G*mass) in FP32 and r = (skip, first
// sorted body, body count, r^2_open as float bits; -1 for leaves).  The
// grouped copy puts the root at 0 and the children of y at 1 + cbase[y] +
// rank, in octant order, with r = (first child's slot, first sorted body,
// min(count, 2^27 - 1) << 4 | child count, r^2_open).
// The walk records to write: the depth-first layout (wnode, indexed by node)
// and/or the grouped layout (wgrp, children adjacent); neither when null.
struct WalkRecs {
    WalkNode *wnode, *wgrp;
    const uint32_t *slot, *cbase;
    const uint8_t *nchild;
    double G, theta;
    int32_t total;
    __device__ bool on() const { return wnode || wgrp; }
};
// The walk records of node x from its values: FP64 mass and COM, whether it
// is a leaf, its side, skip, first and last body, child count and base, and
// its slot in the grouped layout.
__device__ __forceinline__ void put_walk_recs(const WalkRecs &W, int32_t x, double m, double cx, double cy,
                                              double cz, bool leaf, double sd, int32_t sk, int32_t fi, int32_t la,
                                              int nc, uint32_t cb, uint32_t at) {
    WalkNode w;
    w.q = make_float4((float)cx, (float)cy, (float)cz, (float)(W.G * m));
    float r2o;
    if (leaf) r2o = -1.0f;
    else if (W.theta > 0) {
        const double r = sd / W.theta;
        r2o = (float)(r * r);
    } else r2o = INFINITY;
    w.r = make_int4(sk, fi, la - fi + 1, __float_as_int(r2o));
    if (W.wnode) W.wnode[x] = w;
    if (!W.wgrp) return;
    const int32_t cnt = min(la - fi + 1, (1 << 27) - 1);
    WalkNode g;
    g.q = w.q;
    g.r = make_int4(nc ? 1 + (int32_t)cb : 0, fi, (cnt << 4) | nc, __float_as_int(r2o));
    if (at < (uint32_t)W.total) W.wgrp[at] = g;
}

__device__ void write_walk_recs(const WalkRecs &W, int32_t x, int64_t cstride, const double *__restrict__ mass,
                                const double *__restrict__ com, const double *__restrict__ side,
                                const int32_t *__restrict__ body, const int32_t *__restrict__ skip,
                                const int32_t *__restrict__ first, const int32_t *__restrict__ last) {
    put_walk_recs(W, x, mass[x], com[x], com[cstride + x], com[2 * cstride + x], body[x] >= 0, side[x], skip[x],
                  first[x], last[x], W.nchild[x], W.cbase[x], W.slot[x]);
}

// The bottom-up pass is split at subtrees of at most kSubtree nodes: one
// workgroup sums such a subtree entirely in LDS, depth by depth under
// workgroup barriers (its internal nodes sorted by depth with an LDS counting
// sort), and one workgroup then sums the nodes above those subtrees.  Each
// node's slot in the grouped walk layout (kid[] lists the children of x at
// 1 + cbase[x] in octant order), the subtree roots (internal nodes of at
// most kSubtree nodes under a parent of more) and the nodes above them are
// found first (k_children, k_link); the leaves' walk records are written on
// the way.
constexpr int kDepths = kLevels + 1;  // node depths 0..42

// Exclusive offsets of the per-depth counts s_cur (kDepths <= 64 entries)
// into s_off[0..kDepths], by the first wave with one shuffle scan; s_cur is
// reset to the offsets (the running insertion points).  The caller
// synchronises before and after.
__device__ __forceinline__ void depth_offsets(uint32_t *s_cur, uint32_t *s_off) {
    static_assert(kDepths <= 64, "one depth per lane");
    if (threadIdx.x < 64) {
        const unsigned lane = threadIdx.x;
        const uint32_t v = lane < (unsigned)kDepths ? s_cur[lane] : 0u;
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            x += lane >= (unsigned)o ? y : 0u;
        }
        if (lane < (unsigned)kDepths) {
            s_off[lane] = x - v;
            s_cur[lane] = x - v;
        }
        if (lane == (unsigned)kDepths - 1) s_off[kDepths] = x;
    }
}

// It also completes the scan of the child counts: cbase[x] = cbraw[x] (the
// scan within x's block of 1024) + cblk[x >> 10] (the blocks before it).
__global__ void k_link(int32_t total, const int32_t *__restrict__ parent, const uint8_t *__restrict__ rank,
                       const uint32_t *__restrict__ cbraw, const uint32_t *__restrict__ cblk,
                       uint32_t *__restrict__ cbase, uint32_t *__restrict__ slot, int32_t *__restrict__ kid,
                       WalkRecs W, const int32_t *__restrict__ body, const float4 *__restrict__ pos32) {
    const int32_t x = blockIdx.x * kBlock + threadIdx.x;
    if (x >= total) return;
    cbase[x] = cbraw[x] + cblk[x >> 10];
    const int32_t p = parent[x];
    const uint32_t at = x == 0 ? 0u : (p >= 0 && p < total ? 1u + cbraw[p] + cblk[p >> 10] + rank[x] : 0u);
    slot[x] = at;
    if (at < (uint32_t)total && (x == 0 || at > 0)) kid[at] = x;
    const int32_t b = body[x];
    if (b >= 0) {
        // A leaf's records are put_walk_recs' for body b: its COM and G m in
        // FP32 are the body's FP32 position stream entry, skip x + 1, one
        // body, no children, and an opening radius that accepts it always.
        if (W.on()) {
            const float4 q = pos32[b];
            const int r2o = __float_as_int(-1.0f);
            if (W.wnode) W.wnode[x] = WalkNode{q, make_int4(x + 1, b, 1, r2o)};
            if (W.wgrp && at < (uint32_t)total) W.wgrp[at] = WalkNode{q, make_int4(0, b, 1 << 4, r2o)};
        }
    }
}

// calc_com (greg8tree.hpp:211-282) for internal node p from its children
// (kid[1 + cbase[p] ...], octant order), summed in reverse octant order
// (:262-266), with the data in global memory.
__device__ __forceinline__ void com_node_kids(int64_t cstride, int32_t p, const uint8_t *__restrict__ nchild,
                                              const uint32_t *__restrict__ cbase, const int32_t *__restrict__ kid,
                                              int32_t total, double *__restrict__ mass, double *__restrict__ com) {
    const int nc = min((int)nchild[p], 8);
    const uint32_t c0 = 1u + cbase[p];
    double ms = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = nc - 1; k >= 0; --k) {
        const int32_t ch = kid[c0 + k];
        if (ch <= p || ch >= total) continue;
        const double mc = mass[ch];
        ms += mc;
        sx += mc * com[ch];
        sy += mc * com[cstride + ch];
        sz += mc * com[2 * cstride + ch];
    }
    mass[p] = ms;
    com[p] = sx / ms;
    com[cstride + p] = sy / ms;
    com[2 * cstride + p] = sz / ms;
}

__global__ __launch_bounds__(kBlock) void k_com_subtrees(
    int32_t total, const int32_t *__restrict__ roots, const uint32_t *__restrict__ nroots,
    const int32_t *__restrict__ skip, const int32_t *__restrict__ body, const uint8_t *__restrict__ depth,
    const uint8_t *__restrict__ nchild, const uint32_t *__restrict__ cbase, const int32_t *__restrict__ kid,
    double *__restrict__ mass, double *__restrict__ com, int64_t cstride, WalkRecs W, const double *__restrict__ side,
    const int32_t *__restrict__ first, const int32_t *__restrict__ last, MassCom *__restrict__ mc,
    const int32_t *__restrict__ rdest, MassCom *__restrict__ tk) {
    constexpr int kPer = kSubtree / kBlock;  // nodes per thread
    __shared__ double s_m[kSubtree], s_x[kSubtree], s_y[kSubtree], s_z[kSubtree];
    __shared__ int16_t s_kid[kSubtree];   // children of the subtree's nodes, local indices, grouped layout order
    __shared__ uint32_t s_list[kSubtree]; // internal nodes by depth: local index | first child << 11 | count << 22
    __shared__ uint32_t s_cur[kDepths], s_off[kDepths + 1];
    const uint32_t cnt = *nroots;
    const int t = threadIdx.x;
    for (uint32_t q = blockIdx.x; q < cnt; q += gridDim.x) {
        const int32_t r = roots[q];
        const int32_t size = min(skip[r] - r, kSubtree);
        const uint32_t cb0 = cbase[r];
        const int d0 = depth[r];
        if (t < kDepths) s_cur[t] = 0;
        // every load of this thread's nodes (local indices t + j kBlock) at
        // once: indices past the subtree read its last node, unused.  The
        // children of the subtree's nodes fill slots 1 + cbase[r] .. in the
        // grouped layout (cbase is a prefix sum over pre-order, and every
        // node of the subtree but its root is a child of one of its nodes).
        int32_t bd[kPer], kd[kPer];
        uint32_t cb[kPer];
        int dp[kPer], nc[kPer];
        double mv[kPer], xv[kPer], yv[kPer], zv[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int32_t i = t + j * kBlock;
            const int32_t p = r + min(i, size - 1);
            bd[j] = body[p];
            dp[j] = min((int)depth[p], kDepths - 1);
            nc[j] = min((int)nchild[p], 8);
            cb[j] = cbase[p];
            mv[j] = mass[p];
            xv[j] = com[p];
            yv[j] = com[cstride + p];
            zv[j] = com[2 * cstride + p];
            kd[j] = kid[1 + cb0 + (uint32_t)min(i, max(size - 2, 0))];
        }
        __syncthreads();  // s_cur cleared (and LDS free from the previous subtree)
        // the leaves' mass and COM and the child lists into LDS; internal
        // nodes counted per depth
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int32_t i = t + j * kBlock;
            if (i < size - 1) {
                const int32_t ch = kd[j] - r;
                s_kid[i] = (int16_t)(ch > 0 && ch < size ? ch : 0);
            }
            if (i < size) {
                if (bd[j] >= 0) {
                    s_m[i] = mv[j];
                    s_x[i] = xv[j];
                    s_y[i] = yv[j];
                    s_z[i] = zv[j];
                } else {
                    atomicAdd(&s_cur[dp[j]], 1u);
                }
            }
        }
        __syncthreads();
        depth_offsets(s_cur, s_off);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int32_t i = t + j * kBlock;
            if (i < size && bd[j] < 0) {
                const uint32_t c0 = cb[j] - cb0;
                s_list[atomicAdd(&s_cur[dp[j]], 1u)] =
                    (uint32_t)i | (min(c0, (uint32_t)kSubtree - 1u) << 11) | ((uint32_t)nc[j] << 22);
            }
        }
        __syncthreads();
        for (int L = kDepths - 1; L >= d0; --L) {
            const uint32_t b = s_off[L], e = s_off[L + 1];
            if (b == e) continue;  // the same for every thread (LDS values after a barrier)
            for (uint32_t j = b + t; j < e; j += kBlock) {
                const uint32_t rec = s_list[j];
                const int32_t i = (int32_t)(rec & 2047u);
                const int32_t c0 = (int32_t)((rec >> 11) & 2047u), ncj = (int32_t)(rec >> 22);
                // every child's values read at once (absent children read node i's, unused)
                int32_t cc[8];
                bool ok[8];
                double mv[8], xv[8], yv[8], zv[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int32_t ch = k < ncj ? s_kid[min(c0 + k, kSubtree - 1)] : 0;
                    ok[k] = k < ncj && ch > i;
                    cc[k] = ok[k] ? ch : i;
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    mv[k] = s_m[cc[k]];
                    xv[k] = s_x[cc[k]];
                    yv[k] = s_y[cc[k]];
                    zv[k] = s_z[cc[k]];
                }
                double ms = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
                for (int k = 7; k >= 0; --k) {  // reverse octant order (calc_com :262-266)
                    if (!ok[k]) continue;
                    ms += mv[k];
                    sx += mv[k] * xv[k];
                    sy += mv[k] * yv[k];
                    sz += mv[k] * zv[k];
                }
                s_m[i] = ms;
                s_x[i] = sx / ms;
                s_y[i] = sy / ms;
                s_z[i] = sz / ms;
            }
            __syncthreads();  // this depth's sums before the next one up reads them
        }
        if (t == 0) {
            MassCom v;
            v.m = s_m[0];
            v.x = s_x[0];
            v.y = s_y[0];
            v.z = s_z[0];
            mc[r] = v;  // the root's, for the pass over the nodes above the subtrees
            const int32_t d = rdest[q];
            if (d >= 0) tk[d] = v;
        }
        // this thread's internal nodes out, with their walk records (the
        // record inputs of all of them loaded at once; leaves read the root's)
        double sd[kPer];
        int32_t sk[kPer], fi[kPer], la[kPer];
        uint32_t at[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int32_t i = t + j * kBlock;
            const int32_t p = i < size && bd[j] < 0 ? r + i : r;
            sd[j] = side[p];
            sk[j] = skip[p];
            fi[j] = first[p];
            la[j] = last[p];
            at[j] = W.slot[p];
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int32_t i = t + j * kBlock;
            if (i < size && bd[j] < 0) {
                const int32_t p = r + i;
                mass[p] = s_m[i];
                com[p] = s_x[i];
                com[cstride + p] = s_y[i];
                com[2 * cstride + p] = s_z[i];
                if (W.on())
                    put_walk_recs(W, p, s_m[i], s_x[i], s_y[i], s_z[i], false, sd[j], sk[j], fi[j], la[j], nc[j],
                                  cb[j], at[j]);
            }
        }
        __syncthreads();  // LDS reused by the next subtree
    }
}

// The nodes above the subtrees, one workgroup.  They are sorted by depth
// into LDS together with their child lists.  Up to kTopOwn of them, each
// thread owns one and loads all its inputs before the first depth: every
// child's mass and COM from its 32-byte MassCom record (the children that
// are themselves above the subtrees excepted: their sums come from LDS, by
// way of a node -> list position table) and its own walk-record inputs.
// Each depth then costs LDS reads and a barrier, not a round trip to
// memory.  Up to kTopLds of them, each depth takes one round of independent
// loads of the children's records.  (One workgroup's loads are limited by
// the cache lines it can request, so one line per child instead of one per
// SoA array is what sets the time.)  More such nodes: every depth rescans
// the list with the child lists in memory, and the walk records follow in
// one sweep.  A node's walk records are written as soon as its sums are.
constexpr uint32_t kTopLds = 2048;
constexpr int kTopHash = 2048;  // open-addressing table of node -> list position (half full at most)
__device__ __forceinline__ uint32_t top_hash(int32_t x) { return ((uint32_t)x * 2654435761u) >> 21; }
__global__ __launch_bounds__(1024) void k_com_top(int32_t total, const int32_t *__restrict__ top,
                                                  const uint32_t *__restrict__ ntop, const int32_t *__restrict__ skip,
                                                  const int32_t *__restrict__ body, const uint8_t *__restrict__ depth,
                                                  const uint8_t *__restrict__ nchild, const uint32_t *__restrict__ cbase,
                                                  const int32_t *__restrict__ kid, double *__restrict__ mass,
                                                  double *__restrict__ com, int64_t cstride, WalkRecs W,
                                                  const double *__restrict__ side, const int32_t *__restrict__ first,
                                                  const int32_t *__restrict__ last, MassCom *__restrict__ mc,
                                                  const TopRec *__restrict__ ttop, const MassCom *__restrict__ tk) {
    __shared__ int32_t s_list[kTopLds];
    __shared__ int32_t s_kids[kTopLds][8];
    __shared__ uint32_t s_cb[kTopLds];
    __shared__ uint8_t s_nc[kTopLds];
    __shared__ uint32_t s_cur[kDepths], s_off[kDepths + 1];
    __shared__ MassCom s_mc[kTopOwn];       // results of the nodes listed, by list position
    __shared__ int32_t s_hk[kTopHash];      // node -> list position (open addressing)
    __shared__ uint16_t s_hv[kTopHash];
    __shared__ uint16_t s_orig[kTopOwn];    // index in `top` of each list position
    __shared__ uint32_t s_slot[kTopOwn];    // slot in the grouped layout of each list position
    const uint32_t cnt = *ntop;
    const int t = threadIdx.x;
    for (int i = t; i < kDepths; i += blockDim.x) s_cur[i] = 0;
    for (int i = t; i < kTopHash; i += blockDim.x) s_hk[i] = -1;
    __syncthreads();
    for (uint32_t i = t; i < cnt; i += blockDim.x)
        atomicAdd(&s_cur[min(cnt <= (uint32_t)kTopOwn ? (int)ttop[i].depth : (int)depth[top[i]], kDepths - 1)], 1u);
    __syncthreads();
    depth_offsets(s_cur, s_off);
    __syncthreads();
    if (cnt <= (uint32_t)kTopOwn && blockDim.x >= (unsigned)kTopOwn) {
        // One node per thread (list position t), every input loaded before
        // the first depth: the node's TopRec and its children's records
        // (adjacent, from tk), except children that are themselves above
        // the subtrees, whose sums come from LDS (found through a node ->
        // list position table), as do their slots (1 + the parent's cbase +
        // rank).
        if ((uint32_t)t < cnt) {
            const TopRec tr = ttop[t];
            const int32_t p = top[t];
            const uint32_t at = atomicAdd(&s_cur[min((int)tr.depth, kDepths - 1)], 1u);
            s_list[at] = p;
            s_cb[at] = cbase[p];
            s_nc[at] = tr.nc;
            s_orig[at] = (uint16_t)t;
#pragma unroll
            for (int k = 0; k < 8; ++k) s_kids[at][k] = tr.kids[k];
            uint32_t h = top_hash(p);
            while (atomicCAS(&s_hk[h], -1, p) != -1) h = (h + 1) & (kTopHash - 1);  // keys are distinct
            s_hv[h] = (uint16_t)at;
        }
        __syncthreads();
        const uint32_t j = (uint32_t)t;
        const bool mine = j < cnt;
        const int32_t p = mine ? s_list[j] : 0;
        const uint32_t orig = mine ? s_orig[j] : 0u;
        int32_t pos[8];
        bool ok[8];
        MassCom v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t c = mine ? s_kids[j][k] : -1;
            ok[k] = c > p && c < total;
            pos[k] = -1;
            if (ok[k]) {
                uint32_t h = top_hash(c);
                for (int32_t key; (key = s_hk[h]) != -1; h = (h + 1) & (kTopHash - 1))
                    if (key == c) {
                        pos[k] = s_hv[h];
                        break;
                    }
            }
            if (pos[k] >= 0) s_slot[pos[k]] = 1u + s_cb[j] + (uint32_t)k;
        }
        if (mine && p == 0) s_slot[j] = 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (ok[k] && pos[k] < 0) v[k] = tk[8 * orig + k];
        double sd = 0.0;
        int32_t sk = 0, fi = 0, la = 0;
        if (mine && W.on()) {
            const TopRec tr = ttop[orig];
            sd = tr.side;
            sk = tr.skip;
            fi = tr.first;
            la = tr.last;
        }
        __syncthreads();  // s_slot
        const uint32_t slot = mine ? s_slot[j] : 0u;
        for (int L = kDepths - 1; L >= 0; --L) {
            const uint32_t b = s_off[L], e = s_off[L + 1];
            if (b == e) continue;  // the same for every thread (LDS values after a barrier)
            if (j >= b && j < e) {
                double ms = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
                for (int k = 7; k >= 0; --k) {  // reverse octant order (calc_com :262-266)
                    if (!ok[k]) continue;
                    const MassCom u = pos[k] >= 0 ? s_mc[pos[k]] : v[k];
                    ms += u.m;
                    sx += u.m * u.x;
                    sy += u.m * u.y;
                    sz += u.m * u.z;
                }
                MassCom o;
                o.m = ms;
                o.x = sx / ms;
                o.y = sy / ms;
                o.z = sz / ms;
                s_mc[j] = o;
                mass[p] = o.m;
                com[p] = o.x;
                com[cstride + p] = o.y;
                com[2 * cstride + p] = o.z;
                mc[p] = o;
                if (W.on()) put_walk_recs(W, p, o.m, o.x, o.y, o.z, false, sd, sk, fi, la, s_nc[j], s_cb[j], slot);
            }
            __syncthreads();  // this depth's sums (LDS) before the next depth up
        }
    } else if (cnt <= kTopLds) {
        for (uint32_t i = t; i < cnt; i += blockDim.x) {
            const int32_t p = top[i];
            const uint32_t at = atomicAdd(&s_cur[min((int)depth[p], kDepths - 1)], 1u);
            const int nc = min((int)nchild[p], 8);
            const uint32_t cb = cbase[p];
            s_list[at] = p;
            s_cb[at] = cb;
            s_nc[at] = (uint8_t)nc;
#pragma unroll
            for (int k = 0; k < 8; ++k) s_kids[at][k] = k < nc ? kid[1u + cb + k] : -1;
        }
        __syncthreads();
        for (int L = kDepths - 1; L >= 0; --L) {
            const uint32_t b = s_off[L], e = s_off[L + 1];
            if (b == e) continue;  // the same for every thread (LDS values after a barrier)
            for (uint32_t j = b + t; j < e; j += blockDim.x) {
                const int32_t p = s_list[j];
                // every child's record loaded together (absent children read node p's, unused)
                int32_t cc[8];
                bool ok[8];
                MassCom v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int32_t c = s_kids[j][k];
                    ok[k] = c > p && c < total;
                    cc[k] = ok[k] ? c : p;
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = mc[cc[k]];
                double sd = 0.0;
                int32_t sk = 0, fi = 0, la = 0;
                uint32_t slot = 0;
                if (W.on()) {
                    sd = side[p];
                    sk = skip[p];
                    fi = first[p];
                    la = last[p];
                    slot = W.slot[p];
                }
                double ms = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
                for (int k = 7; k >= 0; --k) {  // reverse octant order (calc_com :262-266)
                    if (!ok[k]) continue;
                    ms += v[k].m;
                    sx += v[k].m * v[k].x;
                    sy += v[k].m * v[k].y;
                    sz += v[k].m * v[k].z;
                }
                MassCom o;
                o.m = ms;
                o.x = sx / ms;
                o.y = sy / ms;
                o.z = sz / ms;
                mass[p] = o.m;
                com[p] = o.x;
                com[cstride + p] = o.y;
                com[2 * cstride + p] = o.z;
                mc[p] = o;
                if (W.on()) put_walk_recs(W, p, o.m, o.x, o.y, o.z, false, sd, sk, fi, la, s_nc[j], s_cb[j], slot);
            }
            __syncthreads();
        }
    } else {
        for (int L = kDepths - 1; L >= 0; --L) {
            if (s_off[L] == s_off[L + 1]) continue;
            for (uint32_t i = t; i < cnt; i += blockDim.x) {
                const int32_t p = top[i];
                if (depth[p] != L) continue;
                com_node_kids(cstride, p, nchild, cbase, kid, total, mass, com);
            }
            __syncthreads();
        }
        if (W.on())
            for (uint32_t i = t; i < cnt; i += blockDim.x)
                write_walk_recs(W, top[i], cstride, mass, com, side, body, skip, first, last);
    }
}

// Positions, velocities (and accelerations when a walk wrote them) back in
// input order, through perm (slot -> input index).
__global__ void k_restore(int64_t n, const int32_t *__restrict__ perm, const double *__restrict__ x,
                          const double *__restrict__ v, const double *__restrict__ a, double *__restrict__ xo,
                          double *__restrict__ vo, double *__restrict__ ao) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int64_t b = perm[s];
    for (int k = 0; k < 3; ++k) {
        xo[k * n + b] = x[k * n + s];
        vo[k * n + b] = v[k * n + s];
        if (a) ao[k * n + b] = a[k * n + s];
    }
}

// The frontier walk's schedule: eight contiguous ranges of the group order
// (one per XCD) of about equal cost -- the previous walk's per-group counts
// plus one -- each at most `span` groups long, and every range listed
// heaviest first (a 64-bucket sort).  All eight workgroups compute the same
// cuts from the same counts; workgroup x sorts range x, workgroup 0 writes
// the bounds xr[0..8].
__global__ __launch_bounds__(1024) void k_group_order(int64_t ngroups, int64_t span, const uint32_t *__restrict__ cost,
                                                      int32_t *__restrict__ order, int64_t *__restrict__ xr) {
    __shared__ uint64_t s_part[16];
    __shared__ int64_t s_cut[9];
    __shared__ uint32_t s_max;
    __shared__ uint32_t s_hist[64], s_base[64];
    const int t = threadIdx.x;
    // prefix sums of cost + 1 over per-thread chunks
    const int64_t per = (ngroups + 1023) / 1024;
    const int64_t c0 = std::min<int64_t>(ngroups, per * t), c1 = std::min<int64_t>(ngroups, c0 + per);
    // the chunk's counts, eight independent loads at a time
    uint64_t sum = 0;
    for (int64_t g0 = c0; g0 < c1; g0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = cost[std::min<int64_t>(g0 + k, ngroups - 1)];
#pragma unroll
        for (int k = 0; k < 8; ++k) sum += g0 + k < c1 ? (uint64_t)v[k] + 1u : 0u;
    }
    if (t < 9) s_cut[t] = t == 0 ? 0 : ngroups;
    // inclusive scan of the chunk sums: within each wave by shuffles, then
    // the 16 wave totals
    const unsigned lane = __lane_id(), wv = (unsigned)t >> 6;
    uint64_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o);
        x += lane >= (unsigned)o ? y : 0ull;
    }
    if (lane == 63) s_part[wv] = x;
    __syncthreads();
    uint64_t pre = 0, W = 0;
    for (unsigned i = 0; i < blockDim.x / 64; ++i) {
        pre += i < wv ? s_part[i] : 0ull;
        W += s_part[i];
    }
    // cut x (1..7): the first group whose exclusive prefix reaches x W / 8
    uint64_t e = pre + x - sum;
    for (int64_t g0 = c0; g0 < c1; g0 += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = cost[std::min<int64_t>(g0 + k, ngroups - 1)];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (g0 + k >= c1) break;
            const uint64_t e1 = e + (uint64_t)v[k] + 1u;
            for (int x = 1; x < 8; ++x) {
                const uint64_t target = (W * (uint64_t)x) / 8u;
                if (e < target && e1 >= target) s_cut[x] = g0 + k + 1;
            }
            e = e1;
        }
    }
    __syncthreads();
    if (t == 0) {
        // no range longer than span (the grid has span groups per XCD)
        for (int x = 1; x < 8; ++x) {
            int64_t c = std::max<int64_t>(s_cut[x], s_cut[x - 1]);
            c = std::min<int64_t>(c, s_cut[x - 1] + span);
            c = std::max<int64_t>(c, ngroups - (int64_t)(8 - x) * span);
            s_cut[x] = std::min<int64_t>(c, ngroups);
        }
        s_cut[8] = ngroups;
        if (blockIdx.x == 0)
            for (int x = 0; x < 9; ++x) xr[x] = s_cut[x];
        s_max = 0;
    }
    if (t < 64) s_hist[t] = 0;
    __syncthreads();
    const int64_t lo = s_cut[blockIdx.x], hi = s_cut[blockIdx.x + 1];
    uint32_t mx = 0;
    for (int64_t g = lo + t; g < hi; g += blockDim.x) mx = max(mx, cost[g]);
    atomicMax(&s_max, mx);
    __syncthreads();
    const uint64_t M = (uint64_t)s_max + 1;
    for (int64_t g = lo + t; g < hi; g += blockDim.x)
        atomicAdd(&s_hist[63u - (uint32_t)(((uint64_t)cost[g] * 64u) / M)], 1u);
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int bb = 0; bb < 64; ++bb) {
            s_base[bb] = run;
            run += s_hist[bb];
        }
    }
    __syncthreads();
    for (int64_t g = lo + t; g < hi; g += blockDim.x) {
        const uint32_t at = atomicAdd(&s_base[63u - (uint32_t)(((uint64_t)cost[g] * 64u) / M)], 1u);
        order[lo + at] = (int32_t)g;
    }
}

__global__ void k_frontier_reset(int *__restrict__ werr, unsigned long long *__restrict__ icount2,
                                 unsigned long long *__restrict__ icount) {
    if (threadIdx.x == 0) {
        werr[0] = 0;
        icount2[0] = 0;
        icount[0] = 0;
    }
}

// ---- 6b. per-lane walk with the reference's per-body MAC, FP64 ----
__global__ __launch_bounds__(kBlock) void k_walk_exact(int64_t n, int32_t total, const uint32_t *__restrict__ idx,
                                                       const uint32_t *__restrict__ base,
                                                       const uint32_t *__restrict__ bblk, const uint8_t *__restrict__ c,
                                                       const double *__restrict__ x, const double *__restrict__ mass,
                                                       const double *__restrict__ com, int64_t cstride,
                                                       const double *__restrict__ side,
                                                       const int32_t *__restrict__ body,
                                                       const int32_t *__restrict__ skip, double G, double eps2,
                                                       double theta, double *__restrict__ accs,
                                                       unsigned long long *__restrict__ icount) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    unsigned long long cnt = 0;
    if (s < n) {
        const uint32_t b = idx[s];
        const double px = x[b], py = x[n + b], pz = x[2 * n + b];
        const int cs = c[s], cp = s > 0 ? c[s - 1] : 0;
        const int32_t own = 1 + (int32_t)(base[s] + bblk[s >> 10]) + max(0, cs - cp);
        double ax = 0.0, ay = 0.0, az = 0.0;
        int32_t node = 1;
        while (node < total) {
            if (node == own) {
                node = node + 1;
                continue;
            }
            const double dx = com[node] - px, dy = com[cstride + node] - py, dz = com[2 * cstride + node] - pz;
            bool accept = body[node] >= 0;
            if (!accept) accept = side[node] / __builtin_sqrt(dx * dx + dy * dy + dz * dz) < theta;  // is_source_cell
            if (accept) {
                double s2 = dx * dx + dy * dy + dz * dz + eps2;
                double sq = __builtin_sqrt(s2);
                double k = G * (mass[node] / s2);
                ax += k * (dx / sq);
                ay += k * (dy / sq);
                az += k * (dz / sq);
                ++cnt;
                node = skip[node];
            } else {
                node = node + 1;
            }
        }
        accs[s] = ax;
        accs[n + s] = ay;
        accs[2 * n + s] = az;
    }
    // one atomic per wave instead of one per lane: atomics on one address
    // are serialised in the L2
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
    if (__lane_id() == 0) atomicAdd(icount, cnt);
}


inline unsigned blocks(int64_t n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

void raise_bh_error(int e) {
    if (e == BH_ERR_PLACEMENT)
        throw ApiError(NB_ERR_INVALID_PLACEMENT, "Barnes-Hut: a body does not lie within the root cell (invalid_body_placement)");
    if (e == BH_ERR_INDIST)
        throw ApiError(NB_ERR_INDISTINGUISHABLE, "Barnes-Hut: two bodies share a position (indistinguishable_bodies)");
    if (e == BH_ERR_DEPTH)
        throw ApiError(NB_ERR_CAPACITY, "Barnes-Hut: two bodies share more than 42 octant levels");
}

template <typename T>
void swap_buf(DBuf<T> &a, DBuf<T> &b) {
    std::swap(a.p, b.p);
    std::swap(a.n, b.n);
}

}  // namespace

// The step's drift (and the previous step's closing kick) applied by the
// build's first pass.
struct BHDrift {
    bool drift = false;  // x += dt v
    bool kick = false;   // v += dt a first (the previous step's closing kick)
};

// Builds the octree of the current positions (after the drift, if any) into
// s->bh and leaves the bodies in the tree's order (waits once for the node
// count, to size the later launches and to surface build errors, while
// k_build_nodes runs).
static void bh_build(nb_system *s, BHDrift dr = BHDrift()) {
    const int64_t n = s->n;
    if (!s->bh) s->bh = new BHWork();
    BHWork &w = *s->bh;
    hipStream_t st = s->stream;
    if (n < 2) throw ApiError(NB_ERR_INVALID_PLACEMENT, "Barnes-Hut: the root cell of fewer than two bodies is empty (invalid_body_placement)");
    if (n >= (int64_t)1 << 31) throw ApiError(NB_ERR_CAPACITY, "Barnes-Hut: too many bodies");
    ScopedTimer tm(s->prof, "bh_build", st);
    if (w.cap_n < n) {
        w.khi.ensure(n); w.klo.ensure(n); w.klo_slot.ensure(n); w.idx.ensure(n);
        w.cnt.ensure(n); w.pos32s.ensure(n);
        w.blk.ensure((size_t)(n + 1023) / 1024 + 1);
        w.sorter.ensure(n);
        w.perm.ensure(n); w.perm_alt.ensure(n);
        w.cap_n = n;
    }
    // the other copy of every per-body array, for the move into tree order
    w.xa.ensure(3 * (size_t)n); w.va.ensure(3 * (size_t)n); w.aa.ensure(3 * (size_t)n);
    for (DBuf<double> *b : {&w.mt, &w.rt, &w.et, &w.mt_alt, &w.rt_alt, &w.et_alt}) b->ensure((size_t)n);
    w.part.ensure(6 * (size_t)kBoxParts); w.box.ensure(6); w.scal.ensure(kScalars); w.icount.ensure(1);
    if (!w.h_scal) NB_HIP(hipHostMalloc((void **)&w.h_scal, kScHost * sizeof(uint32_t), hipHostMallocDefault));
    if (!w.ev_count) NB_HIP(hipEventCreateWithFlags(&w.ev_count, hipEventDisableTiming));
    uint32_t *const total_d = w.scal.p + kScTotal, *const nroots_d = w.scal.p + kScRoots;
    int *const err_d = (int *)(w.scal.p + kScErr);

    w.khi.ensure(n); w.idx.ensure(n);  // the sort may exchange these with its own buffers
    // c and its chunk minima, padded to whole chunks of 64 for the searches' vector loads
    const int64_t n1 = (n + 63) >> 6, n2 = (n1 + 63) >> 6, n3 = (n2 + 63) >> 6;
    w.c.ensure((size_t)n + 64); w.cm1.ensure((size_t)n1 + 64); w.cm2.ensure((size_t)n2 + 64); w.cm3.ensure((size_t)n3 + 64);
    // one body per thread when the pass also moves the bodies (up to 4M of them), else a grid-stride 1024 blocks
    const int nparts = (int)std::min<int64_t>(dr.drift ? kBoxParts : 1024, (n + kBlock - 1) / kBlock);
    auto bbox = !dr.drift ? k_bbox_partial<false, false>
                          : dr.kick ? k_bbox_partial<true, true> : k_bbox_partial<false, true>;
    hipLaunchKernelGGL(bbox, dim3(nparts), dim3(kBlock), 0, st, n, s->x.p, s->v.p, s->a.p, s->dt, s->dt, w.part.p);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, st, nparts, w.part.p, w.box.p, w.scal.p);
    hipLaunchKernelGGL(k_keys, dim3(blocks(n)), dim3(kBlock), 0, st, n, s->x.p, w.box.p, w.khi.p, w.klo_slot.p,
                       w.idx.p, err_d);
    // the sort by the high key, then ties by the low key, and the node counts
    const int nb = (int)((n + 1023) / 1024);
    const int passes = s->tune_bh_sort > 0 ? s->tune_bh_sort : w.sort_passes;
    const int lo_bit = 3 * kHiLevels - RadixSorter::kDigitBits * passes;
    w.sorter.sort_swap(st, w.khi, w.idx, n, 3 * kHiLevels, lo_bit);
    hipLaunchKernelGGL(k_gather_lo, dim3(blocks(n)), dim3(kBlock), 0, st, n, w.klo_slot.p, w.idx.p, w.klo.p);
    hipLaunchKernelGGL(k_fix_ties, dim3(blocks(n)), dim3(kBlock), 0, st, n, lo_bit, w.khi.p, w.klo.p, w.idx.p);
    hipLaunchKernelGGL(k_lcp_counts, dim3(blocks(n)), dim3(kBlock), 0, st, n, w.khi.p, w.klo.p, w.idx.p, s->x.p,
                       w.c.p, w.cm1.p, w.cnt.p, err_d, w.scal.p + kScLong);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(1024), 0, st, n, w.cnt.p, w.blk.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, nb, w.blk.p, total_d, w.cm1.p, n1, w.cm2.p, n2,
                       w.cm3.p, n3);
    // Node arrays for at least 2n nodes (or the last build's count and a
    // quarter), so the nodes are written while the host waits for the count;
    // with more nodes, k_build_nodes writes nothing and runs again below.
    auto grow_nodes = [&](int64_t want) {
        if (w.cap_nodes >= want) return;
        const size_t t = (size_t)want;
        w.first.ensure(t); w.last.ensure(t); w.skip.ensure(t); w.parent.ensure(t); w.body.ensure(t);
        w.depth.ensure(t); w.nchild.ensure(t); w.octant.ensure(t); w.mass.ensure(t); w.com.ensure(3 * t);
        w.side.ensure(t); w.wnode.ensure(t); w.slot.ensure(t); w.kid.ensure(t); w.roots.ensure(t); w.top.ensure(t);
        w.wgrp.ensure(t); w.rank.ensure(t); w.cbase.ensure(t); w.cbraw.ensure(t); w.cblk.ensure(t / 1024 + 2); w.ctot.ensure(1);
        w.mc.ensure(t); w.rdest.ensure(t); w.ttop.ensure(kTopOwn); w.tk.ensure(8 * kTopOwn);
        w.cap_nodes = want;
    };
    grow_nodes(2 * n + 64);
    const RunIndex R{w.c.p, w.cm1.p, w.cm2.p, w.cm3.p, n, n1, n2, n3};
    // radii and restitutions follow the bodies only while the collision pass runs
    const bool need_re = s->coll_mode == NB_OVERLAP_COLL_CHECK;
    const bool re_from_tree = w.permuted && w.re_tree;
    const int re_mode = !need_re ? kReSkip : re_from_tree ? kReFromPrevious : kReFromInput;
    // nodes, and the bodies into tree order
    auto build_nodes = [&]() {
        hipLaunchKernelGGL(k_build_nodes, dim3(blocks(n)), dim3(kBlock), 0, st, n, R, w.cnt.p, total_d, w.blk.p,
                           w.khi.p, w.klo.p, w.box.p, w.idx.p, s->x.p, s->v.p, w.permuted ? w.mt.p : s->m.p,
                           re_from_tree ? w.rt.p : s->r.p, re_from_tree ? w.et.p : s->e.p, re_mode,
                           w.permuted ? w.perm.p : nullptr, w.xa.p, w.va.p, w.mt_alt.p, w.rt_alt.p, w.et_alt.p,
                           w.perm_alt.p, w.pos32s.p, s->G, w.first.p, w.last.p, w.skip.p, w.body.p, w.depth.p,
                           w.octant.p, w.side.p, w.mass.p, w.com.p, w.cap_nodes);
    };
    // the host waits for the count only; k_build_nodes runs meanwhile
    NB_HIP(hipMemcpyAsync(w.h_scal, w.scal.p, kScHost * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    NB_HIP(hipEventRecord(w.ev_count, st));
    build_nodes();
    NB_HIP(hipEventSynchronize(w.ev_count));
    raise_bh_error((int)w.h_scal[kScErr]);
    // the next sort: the fewest passes whose runs in this build were all at most 32 bodies
    {
        int next = 3 * kHiLevels / RadixSorter::kDigitBits;
        for (int q = kSortStats - 1; q >= 0; --q)
            if (w.h_scal[kScLong + q] == 0) next = kMinSortPasses + q;
            else break;
        w.sort_passes = next;
    }
    const int32_t total = (int32_t)w.h_scal[kScTotal];
    if (total > w.cap_nodes) {
        NB_HIP(hipStreamSynchronize(st));  // the launch that wrote nothing has ended before its arrays go
        grow_nodes((int64_t)total + total / 4);
        build_nodes();
    }
    w.nodes = total;
    const int64_t cs = w.cap_nodes;  // stride of the SoA com array
    swap_buf(s->x, w.xa); swap_buf(s->v, w.va);
    swap_buf(w.mt, w.mt_alt);
    if (need_re) {
        swap_buf(w.rt, w.rt_alt);
        swap_buf(w.et, w.et_alt);
    }
    w.re_tree = need_re;
    swap_buf(w.perm, w.perm_alt);
    w.permuted = true;
    w.acc_permuted = false;  // until a walk writes accelerations in tree order
    // children, their slots in the grouped walk layout, subtree lists
    hipLaunchKernelGGL(k_children, dim3(blocks(total)), dim3(kBlock), 0, st, total, w.skip.p, w.body.p,
                       w.nchild.p, w.rank.p, w.cbraw.p, w.parent.p, w.roots.p, nroots_d, w.top.p, nroots_d + 1,
                       w.mass.p, w.com.p, cs, w.mc.p, w.side.p, w.first.p, w.last.p, w.depth.p, w.ttop.p, w.tk.p,
                       w.rdest.p);
    const int nbn = (int)((total + 1023) / 1024);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nbn), dim3(1024), 0, st, (int64_t)total, w.cbraw.p, w.cblk.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, nbn, w.cblk.p, w.ctot.p);
    // the FP32 walks' records: the grouped layout for the frontier walk (its
    // depth-first fallback reads that too), the depth-first one for bh_walk 2
    const bool records = s->precision == NB_PREC_MIXED;
    const bool dfs = s->tune_bh_walk == 2;
    const WalkRecs W{records && dfs ? w.wnode.p : nullptr, records && !dfs ? w.wgrp.p : nullptr, w.slot.p,
                     w.cbase.p, w.nchild.p, s->G, s->theta, total};
    hipLaunchKernelGGL(k_link, dim3(blocks(total)), dim3(kBlock), 0, st, total, w.parent.p, w.rank.p, w.cbraw.p,
                       w.cblk.p, w.cbase.p, w.slot.p, w.kid.p, W, w.body.p, w.pos32s.p);
    // mass, COM and the internal nodes' walk records, bottom up
    const unsigned sub_grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, total / 128 + 1));
    hipLaunchKernelGGL(k_com_subtrees, dim3(sub_grid), dim3(kBlock), 0, st, total, w.roots.p, nroots_d, w.skip.p,
                       w.body.p, w.depth.p, w.nchild.p, w.cbase.p, w.kid.p, w.mass.p, w.com.p, cs, W, w.side.p,
                       w.first.p, w.last.p, w.mc.p, w.rdest.p, w.tk.p);
    hipLaunchKernelGGL(k_com_top, dim3(1), dim3(1024), 0, st, total, w.top.p, nroots_d + 1, w.skip.p, w.body.p,
                       w.depth.p, w.nchild.p, w.cbase.p, w.kid.p, w.mass.p, w.com.p, cs, W, w.side.p, w.first.p,
                       w.last.p, w.mc.p, w.ttop.p, w.tk.p);
    NB_HIP(hipGetLastError());
}

// The bodies back in their input order after Barnes-Hut steps (no-op when
// they already are).
void bh_restore_order(nb_system *s) {
    if (!s->bh || !s->bh->permuted) return;
    BHWork &w = *s->bh;
    const int64_t n = s->n;
    const bool acc = w.acc_permuted;
    hipLaunchKernelGGL(k_restore, dim3(blocks(n)), dim3(kBlock), 0, s->stream, n, w.perm.p, s->x.p, s->v.p,
                       acc ? s->a.p : nullptr, w.xa.p, w.va.p, w.aa.p);
    NB_HIP(hipGetLastError());
    swap_buf(s->x, w.xa); swap_buf(s->v, w.va);
    if (acc) swap_buf(s->a, w.aa);
    w.permuted = false;
    w.acc_permuted = false;
}

// Forgets the tree order (the bodies were replaced).
int bh_next_sort_passes(const nb_system *s) {
    return s->bh ? s->bh->sort_passes : 3 * kHiLevels / RadixSorter::kDigitBits;
}

void bh_forget_order(nb_system *s) {
    if (!s->bh) return;
    s->bh->permuted = false;
    s->bh->sort_passes = 3 * kHiLevels / RadixSorter::kDigitBits;  // new bodies: the full sort until measured
}

static void bh_forces_step(nb_system *s, int64_t *interactions, BHDrift dr) {
    const int64_t n = s->n;
    if (!s->bh) s->bh = new BHWork();
    BHWork &w = *s->bh;
    hipStream_t st = s->stream;
    // The frontier walk's launch: one wave per group, a multiple of 8
    // workgroups so that each XCD gets one range of `span` groups (XCD
    // ranges of equal cost hold at most 1.5x the mean count).  Its schedule
    // is computed on the second stream while the tree is built.
    const bool frontier_walk = s->precision == NB_PREC_MIXED && s->tune_bh_walk != 2 && n >= 2;
    const int64_t ngroups = (n + 63) / 64;
    const int64_t mean = (ngroups + 7) / 8;
    const int64_t wgs = std::max<int64_t>(1, (mean + mean / 2 + kFWaves - 1) / kFWaves);
    const int64_t G = 8 * wgs;
    const int64_t span = wgs * kFWaves;
    const int32_t cap = s->tune_bh_stack > 0 ? s->tune_bh_stack : kFStackDefault;
    if (frontier_walk) {
        w.fstack.ensure((size_t)G * kFWaves * (size_t)cap);
        w.werr.ensure(1); w.icount2.ensure(1); w.icount.ensure(1); w.gxr.ensure(9);
        if (!w.sched) {
            NB_HIP(hipStreamCreateWithFlags(&w.sched, hipStreamNonBlocking));
            NB_HIP(hipEventCreateWithFlags(&w.ev_sched_in, hipEventDisableTiming));
            NB_HIP(hipEventCreateWithFlags(&w.ev_sched_out, hipEventDisableTiming));
        }
        NB_HIP(hipEventRecord(w.ev_sched_in, st));
        NB_HIP(hipStreamWaitEvent(w.sched, w.ev_sched_in, 0));
        if (w.gcost_n != ngroups) {
            w.gcost.ensure((size_t)ngroups);
            w.gorder.ensure((size_t)ngroups);
            NB_HIP(hipMemsetAsync(w.gcost.p, 0, (size_t)ngroups * sizeof(uint32_t), w.sched));
            w.gcost_n = ngroups;
        }
        hipLaunchKernelGGL(k_frontier_reset, dim3(1), dim3(64), 0, w.sched, w.werr.p, w.icount2.p, w.icount.p);
        hipLaunchKernelGGL(k_group_order, dim3(8), dim3(1024), 0, w.sched, ngroups, span, w.gcost.p, w.gorder.p,
                           w.gxr.p);
        NB_HIP(hipGetLastError());
        NB_HIP(hipEventRecord(w.ev_sched_out, w.sched));
    }
    bh_build(s, dr);
    const int32_t total = (int32_t)w.nodes;
    const int64_t cs = w.cap_nodes;
    if (!frontier_walk) NB_HIP(hipMemsetAsync(w.icount.p, 0, sizeof(unsigned long long), st));  // else the schedule's reset
    bool frontier = false;
    {
        ScopedTimer tm(s->prof, "bh_walk", st);
        if (s->precision == NB_PREC_MIXED) {
            const int bucket = s->tune_bh_bucket > 0 ? s->tune_bh_bucket : w.bucket;
            const float eps2 = s->eps > 0 ? (float)(s->eps * s->eps) : 0.0f;
            auto wave = s->eps > 0 ? k_walk_wave<true> : k_walk_wave<false>;
            if (!frontier_walk) {
                hipLaunchKernelGGL(wave, dim3(blocks(n)), dim3(kBlock), 0, st, n, total, w.pos32s.p, w.wnode.p,
                                   bucket, eps2, s->a.p, w.icount.p, nullptr, nullptr, nullptr, nullptr);
            } else {
                NB_HIP(hipStreamWaitEvent(st, w.ev_sched_out, 0));
                auto front = s->eps > 0 ? k_walk_frontier<true> : k_walk_frontier<false>;
                hipLaunchKernelGGL(front, dim3((unsigned)G), dim3(kBlock), 0, st, n, total, w.pos32s.p, w.wgrp.p,
                                   bucket, eps2, s->a.p, w.icount.p, w.fstack.p, cap, w.gorder.p, w.gxr.p,
                                   w.gcost.p, w.werr.p);
                hipLaunchKernelGGL(wave, dim3(blocks(n)), dim3(kBlock), 0, st, n, total, w.pos32s.p,
                                   (const WalkNode *)nullptr, bucket, eps2, s->a.p, w.icount2.p, w.werr.p, w.wgrp.p,
                                   w.slot.p, w.skip.p);
                frontier = true;
            }
        } else {
            hipLaunchKernelGGL(k_walk_exact, dim3(blocks(n)), dim3(kBlock), 0, st, n, total, w.idx.p, w.cnt.p, w.blk.p,
                               w.c.p,
                               s->x.p, w.mass.p, w.com.p, cs, w.side.p, w.body.p, w.skip.p, s->G, s->eps * s->eps,
                               s->theta, s->a.p, w.icount.p);
        }
    }
    NB_HIP(hipGetLastError());
    w.acc_permuted = true;
    if (interactions) {
        // after a frontier overflow the gated wave walk's count is the one
        unsigned long long c = 0, c2 = 0;
        int over = 0;
        NB_HIP(hipMemcpyAsync(&c, w.icount.p, sizeof(c), hipMemcpyDeviceToHost, st));
        if (frontier) {
            NB_HIP(hipMemcpyAsync(&c2, w.icount2.p, sizeof(c2), hipMemcpyDeviceToHost, st));
            NB_HIP(hipMemcpyAsync(&over, w.werr.p, sizeof(int), hipMemcpyDeviceToHost, st));
        }
        NB_HIP(hipStreamSynchronize(st));
        *interactions = (int64_t)(over ? c2 : c);
    }
    s->acc_valid = true;
}

static void bh_collision_pass(nb_system *s) {
    BHWork &w = *s->bh;
    const int64_t n = s->n;
    const int32_t total = (int32_t)w.nodes;
    hipStream_t st = s->stream;
    ScopedTimer tm(s->prof, "bh_coll", st);
    w.nbox.ensure((size_t)kBoxStride * w.cap_nodes);
    w.leaf_of.ensure((size_t)w.cap_n);
    w.bc_cnt.ensure((size_t)w.cap_n);
    w.bc_blk.ensure((size_t)(w.cap_n + 1023) / 1024 + 1);
    w.bc_total.ensure(1);
    w.rmax.ensure(1);
    if (!w.h_bc_total) NB_HIP(hipHostMalloc((void **)&w.h_bc_total, sizeof(uint32_t), hipHostMallocDefault));
    hipLaunchKernelGGL(k_node_box, dim3(blocks(total)), dim3(kBlock), 0, st, total, w.box.p, w.first.p, w.depth.p,
                       w.khi.p, w.klo.p, w.nbox.p);
    hipLaunchKernelGGL(k_leaf_of, dim3(blocks(total)), dim3(kBlock), 0, st, total, w.body.p, w.first.p, w.leaf_of.p);
    NB_HIP(hipMemsetAsync(w.rmax.p, 0, sizeof(double), st));
    hipLaunchKernelGGL(k_rmax, dim3(std::min<unsigned>(1024u, blocks(n))), dim3(kBlock), 0, st, n, w.rt.p, w.rmax.p);
    w.c_nb.ensure((size_t)kCollCache * (size_t)w.cap_n);
    w.c_geo.ensure((size_t)4 * kCollCache * (size_t)w.cap_n);
    hipLaunchKernelGGL(k_bhc_scan<false>, dim3(blocks(n)), dim3(kBlock), 0, st, n, total, w.idx.p, w.leaf_of.p,
                       w.nbox.p, w.skip.p, w.body.p, w.parent.p, s->x.p, w.rt.p, w.rmax.p, w.bc_cnt.p, 0u, nullptr, nullptr,
                       nullptr, w.c_nb.p, w.c_geo.p);
    const int nb = (int)((n + 1023) / 1024);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(1024), 0, st, n, w.bc_cnt.p, w.bc_blk.p);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, nb, w.bc_blk.p, w.bc_total.p);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(1024), 0, st, n, w.bc_cnt.p, w.bc_blk.p);
    NB_HIP(hipMemcpyAsync(w.h_bc_total, w.bc_total.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    NB_HIP(hipStreamSynchronize(st));
    const uint32_t nent = *w.h_bc_total - 1u;  // the scan's total counts one extra (the tree root)
    if (nent == 0) return;
    if (w.e_own.n < nent) {
        const size_t cap = (size_t)nent + nent / 4 + 64;
        w.e_own.ensure(cap); w.e_nb.ensure(cap); w.e_geo.ensure(4 * cap); w.e_mom.ensure(2 * cap);
    }
    hipLaunchKernelGGL(k_bhc_scan<true>, dim3(blocks(n)), dim3(kBlock), 0, st, n, total, w.idx.p, w.leaf_of.p,
                       w.nbox.p, w.skip.p, w.body.p, w.parent.p, s->x.p, w.rt.p, w.rmax.p, w.bc_cnt.p, nent, w.e_own.p,
                       w.e_nb.p, w.e_geo.p, w.c_nb.p, w.c_geo.p);
    // components of the overlap graph, entries grouped by component in
    // (owner, visit) order, then one lane per component
    w.cc_par.ensure((size_t)w.cap_n);
    if (w.cc_keys.n < nent) {
        const size_t cap = (size_t)nent + nent / 4 + 64;
        w.cc_keys.ensure(cap); w.cc_vals.ensure(cap);
        w.g_own.ensure(cap); w.g_nb.ensure(cap); w.g_geo.ensure(4 * cap);
        w.sorter.ensure((int64_t)cap);
    }
    const unsigned eb = blocks((int64_t)nent);
    hipLaunchKernelGGL(k_cc_init, dim3(eb), dim3(kBlock), 0, st, nent, w.e_own.p, w.e_nb.p, w.cc_par.p);
    hipLaunchKernelGGL(k_cc_union, dim3(eb), dim3(kBlock), 0, st, nent, w.e_own.p, w.e_nb.p, w.cc_par.p);
    hipLaunchKernelGGL(k_cc_keys, dim3(eb), dim3(kBlock), 0, st, nent, w.e_own.p, w.cc_par.p, w.cc_keys.p,
                       w.cc_vals.p);
    w.sorter.sort(st, w.cc_keys.p, w.cc_vals.p, (int64_t)nent, 63);
    hipLaunchKernelGGL(k_cc_gather, dim3(eb), dim3(kBlock), 0, st, nent, w.cc_vals.p, w.e_own.p, w.e_nb.p, w.e_geo.p,
                       w.g_own.p, w.g_nb.p, w.g_geo.p);
    hipLaunchKernelGGL(k_bhc_replay, dim3(eb), dim3(kBlock), 0, st, n, nent, w.cc_keys.p, w.g_own.p, w.g_nb.p,
                       w.g_geo.p, w.e_mom.p, w.mt.p, w.et.p, s->v.p);
    NB_HIP(hipGetLastError());
}

void bh_forces(nb_system *s, int64_t *interactions) { bh_forces_step(s, interactions, BHDrift()); }

// evolve's KDK + Barnes-Hut loop body (gregsys.hpp:2914-2939): drift a full
// step, rebuild, walk, kick a full step (velocities stay half a step ahead),
// then bh_s_coll when collisions are on.  The drift is applied by the
// build's first pass, and so is the previous step's kick when nothing
// between the two reads the velocities (no collision pass, no record, not
// the call's last step).  The bodies stay in tree order from step to step; a
// record is taken in input order.
void bh_steps(nb_system *s, uint64_t steps) {
    bool kick_pending = false;
    for (uint64_t k = 0; k < steps; ++k) {
        BHDrift dr;
        dr.drift = true;
        dr.kick = kick_pending;
        bh_forces_step(s, nullptr, dr);
        const bool coll = s->coll_mode == NB_OVERLAP_COLL_CHECK;
        const bool rec = record_due(s, s->step_count + 1);
        kick_pending = !coll && !rec && k + 1 < steps;
        if (!kick_pending) launch_kick(s->stream, s->n, s->a.p, s->v.p, s->dt);
        if (coll) bh_collision_pass(s);
        ++s->step_count;
        if (rec) {
            bh_restore_order(s);
            record_state(s, s->step_count);
        }
    }
}

int bh_topology(nb_system *s, int64_t cap, int32_t *depth, int32_t *octant, int64_t *body, int64_t *parent,
                double *mass, double *com, double *side, int64_t *count) {
    bh_build(s);
    BHWork &w = *s->bh;
    const int64_t t = w.nodes, n = s->n, cs = w.cap_nodes;
    *count = t;
    if (cap < t) {
        bh_restore_order(s);
        throw ApiError(NB_ERR_CAPACITY, "nb_tree_topology: capacity smaller than the node count");
    }
    std::vector<uint8_t> d8((size_t)t);
    std::vector<int8_t> o8((size_t)t);
    std::vector<int32_t> b32((size_t)t), p32((size_t)t), perm((size_t)n);
    std::vector<double> cm(3 * (size_t)cs);
    hipStream_t st = s->stream;
    NB_HIP(hipMemcpyAsync(d8.data(), w.depth.p, (size_t)t, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(o8.data(), w.octant.p, (size_t)t, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(b32.data(), w.body.p, (size_t)t * 4, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(p32.data(), w.parent.p, (size_t)t * 4, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(perm.data(), w.perm.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    if (mass) NB_HIP(hipMemcpyAsync(mass, w.mass.p, (size_t)t * 8, hipMemcpyDeviceToHost, st));
    if (side) NB_HIP(hipMemcpyAsync(side, w.side.p, (size_t)t * 8, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(cm.data(), w.com.p, (size_t)cs * 24, hipMemcpyDeviceToHost, st));
    NB_HIP(hipStreamSynchronize(st));
    bh_restore_order(s);
    for (int64_t k = 0; k < t; ++k) {
        if (depth) depth[k] = d8[(size_t)k];
        if (octant) octant[k] = o8[(size_t)k];
        // leaves name the body by its input index (the tree's slots are in tree order)
        if (body) body[k] = b32[(size_t)k] >= 0 ? perm[(size_t)b32[(size_t)k]] : -1;
        if (parent) parent[k] = p32[(size_t)k];
        if (com)
            for (int q = 0; q < 3; ++q) com[3 * k + q] = cm[(size_t)q * cs + k];
    }
    return NB_OK;
}

}  // namespace nbe
