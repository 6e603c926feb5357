// nb_sort.hip -- LSD radix sort of (uint64 key, uint32 value) pairs for gfx950.
//
// 9-bit digits (a 63-bit key takes 7 passes), three kernels per pass:
// per-tile histogram, one exclusive scan over the [digit][tile] counts, and a
// stable scatter.  The scatter holds its tile's items in registers, ranks
// them round by round (kThreads items per round, in index order) with wave64
// ballots -- a lane's peers are the lanes whose digit equals its own, its
// rank is the popcount of the peers below it, per-wave digit counts in LDS
// order the waves -- and places every item at its rank in an LDS copy of the
// tile sorted by digit.  The tile then leaves LDS in that order, so
// neighbouring lanes store to neighbouring addresses of one digit's run.
// No host synchronisation: a sort can be captured.
#include "nb_internal.hpp"
#include "nb_sort.hpp"

namespace nbe {

namespace {

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = 4;
constexpr int kTileItems = kThreads * kRounds;  // 2048
constexpr int kBits = 9;
constexpr int kBins = 1 << kBits;  // 512, one per thread
static_assert(kBins == kThreads, "one digit bin per thread");

__global__ __launch_bounds__(kThreads) void k_hist(const uint64_t *__restrict__ keys, int n, int shift,
                                                   uint32_t *__restrict__ hist, int ntiles) {
    __shared__ uint32_t h[kBins];
    h[threadIdx.x] = 0;
    const int base = blockIdx.x * kTileItems;
    uint64_t k[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = base + r * kThreads + threadIdx.x;
        k[r] = i < n ? keys[i] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (base + r * kThreads + (int)threadIdx.x < n) atomicAdd(&h[(unsigned)(k[r] >> shift) & (kBins - 1)], 1u);
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Block-wide inclusive scan of one value per thread.
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t *sm) {
    const int t = threadIdx.x;
    sm[t] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const uint32_t u = t >= off ? sm[t - off] : 0;
        __syncthreads();
        sm[t] += u;
        __syncthreads();
    }
    const uint32_t r = sm[t];
    __syncthreads();
    return r;
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts
// (in place) and the digit's total.
__global__ __launch_bounds__(kThreads) void k_tile_scan(uint32_t *__restrict__ hist, int ntiles,
                                                        uint32_t *__restrict__ dtot) {
    __shared__ uint32_t sm[kThreads];
    uint32_t *row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int b = 0; b < ntiles; b += kThreads) {
        const int i = b + threadIdx.x;
        const uint32_t v = i < ntiles ? row[i] : 0;
        const uint32_t inc = block_incl_scan(v, sm);
        const uint32_t chunk = sm[kThreads - 1];
        if (i < ntiles) row[i] = carry + inc - v;
        carry += chunk;
        __syncthreads();  // sm is rewritten by the next chunk
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(kThreads) void k_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                      uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, int n,
                                                      int shift, const uint32_t *__restrict__ hist, int ntiles,
                                                      const uint32_t *__restrict__ dtot) {
    // the per-wave digit counts of the ranking rounds, then the sorted tile
    constexpr int kRankBytes = kWaves * kBins * 4;
    constexpr int kTileBytes = kTileItems * 12;
    __shared__ __align__(16) unsigned char s_raw[kRankBytes > kTileBytes ? kRankBytes : kTileBytes];
    __shared__ uint32_t s_scan[kThreads];
    __shared__ uint32_t s_run[kBins];   // next local position of each digit
    __shared__ uint32_t s_gofs[kBins];  // global position - local position, per digit
    uint32_t(*wcnt)[kBins] = reinterpret_cast<uint32_t(*)[kBins]>(s_raw);
    uint64_t *s_k = reinterpret_cast<uint64_t *>(s_raw);
    uint32_t *s_v = reinterpret_cast<uint32_t *>(s_raw + kTileItems * 8);

    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int base = blockIdx.x * kTileItems;
    const int live_n = min(kTileItems, n - base);
    uint64_t k[kRounds];
    uint32_t v[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int i = base + r * kThreads + t;
        k[r] = i < n ? kin[i] : 0;
        v[r] = i < n ? vin[i] : 0;
    }
    // the tile's own digit counts (s_run) and the digits' global starts
    s_run[t] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (r * kThreads + t < live_n) atomicAdd(&s_run[(unsigned)(k[r] >> shift) & (kBins - 1)], 1u);
    __syncthreads();
    const uint32_t own = s_run[t];
    const uint32_t lstart = block_incl_scan(own, s_scan) - own;
    const uint32_t dt = dtot[t];
    const uint32_t dstart = block_incl_scan(dt, s_scan) - dt;
    s_gofs[t] = dstart + hist[t * ntiles + blockIdx.x] - lstart;
    s_run[t] = lstart;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    // ranks: local position of each item in the digit-sorted tile
    const uint64_t lt = lanemask_lt();
    uint32_t pos[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const bool live = r * kThreads + t < live_n;
        const unsigned d = (unsigned)(k[r] >> shift) & (kBins - 1);
        uint64_t peers = __ballot(live);
#pragma unroll
        for (int b = 0; b < kBits; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const unsigned rank = __popcll(peers & lt);
        if (live && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        uint32_t off = s_run[d] + rank;
        for (int w = 0; w < wave; ++w) off += wcnt[w][d];
        pos[r] = off;
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        s_run[t] += add;
        __syncthreads();
    }
    // the tile sorted by digit in LDS (overlays the per-wave counts, which the
    // last round's barrier has retired)
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (r * kThreads + t < live_n) {
            s_k[pos[r]] = k[r];
            s_v[pos[r]] = v[r];
        }
    __syncthreads();
    for (int j = t; j < live_n; j += kThreads) {
        const uint64_t kk = s_k[j];
        const uint32_t dst = s_gofs[(unsigned)(kk >> shift) & (kBins - 1)] + (uint32_t)j;
        kout[dst] = kk;
        vout[dst] = s_v[j];
    }
}

}  // namespace

void RadixSorter::ensure(int64_t n) {
    const int nt = (int)((n + kTileItems - 1) / kTileItems);
    k2.ensure((size_t)std::max<int64_t>(n, 1));
    v2.ensure((size_t)std::max<int64_t>(n, 1));
    hist.ensure((size_t)kBins * std::max(nt, 1));
    dtot.ensure(kBins);
}

// Runs the digit passes from (keys, vals) and returns the buffers that hold
// the result: the caller's after an even number of passes, else the
// sorter's own.  Digits above `bits` are assumed zero.
std::pair<uint64_t *, uint32_t *> RadixSorter::passes(hipStream_t s, uint64_t *keys, uint32_t *vals, int64_t n,
                                                      int bits) {
    if (n <= 1) return {keys, vals};
    if (n >= ((int64_t)1 << 31) - kTileItems) throw ApiError(NB_ERR_CAPACITY, "radix sort: too many items");
    ensure(n);
    const int nt = (int)((n + kTileItems - 1) / kTileItems);
    uint64_t *ka = keys, *kb = k2.p;
    uint32_t *va = vals, *vb = v2.p;
    for (int shift = 0; shift < bits; shift += kBits) {
        hipLaunchKernelGGL(k_hist, dim3(nt), dim3(kThreads), 0, s, ka, (int)n, shift, hist.p, nt);
        hipLaunchKernelGGL(k_tile_scan, dim3(kBins), dim3(kThreads), 0, s, hist.p, nt, dtot.p);
        hipLaunchKernelGGL(k_scatter, dim3(nt), dim3(kThreads), 0, s, ka, va, kb, vb, (int)n, shift, hist.p, nt,
                           dtot.p);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    NB_HIP(hipGetLastError());
    return {ka, va};
}

// Sorts keys/vals in place (the result lands back in the caller's buffers).
void RadixSorter::sort(hipStream_t s, uint64_t *keys, uint32_t *vals, int64_t n, int bits) {
    const auto r = passes(s, keys, vals, n, bits);
    if (r.first != keys) {
        NB_HIP(hipMemcpyAsync(keys, r.first, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        NB_HIP(hipMemcpyAsync(vals, r.second, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

// Sorts and, where the result lies in the sorter's buffers, exchanges those
// with the caller's instead of copying back (both hold at least n items; the
// caller must size its buffers with ensure() before each use, as the sizes
// travel with the exchange).
void RadixSorter::sort_swap(hipStream_t s, DBuf<uint64_t> &keys, DBuf<uint32_t> &vals, int64_t n, int bits) {
    const auto r = passes(s, keys.p, vals.p, n, bits);
    if (r.first != keys.p) {
        std::swap(keys.p, k2.p);
        std::swap(keys.n, k2.n);
        std::swap(vals.p, v2.p);
        std::swap(vals.n, v2.n);
    }
}

}  // namespace nbe
