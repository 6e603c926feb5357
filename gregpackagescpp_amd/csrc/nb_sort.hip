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
static_assert(kBits == RadixSorter::kDigitBits, "the digit width the callers plan with");

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

__device__ __forceinline__ uint64_t lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Block-wide exclusive scan of one value per thread: a scan within each
// wave (shuffles), the wave totals through LDS, two barriers.  s_w holds
// kWaves entries; total receives the block's sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w, uint32_t &total) {
    const unsigned lane = __lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        x += lane >= (unsigned)o ? y : 0u;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
        const uint32_t u = s_w[i];
        pre += i < (int)w ? u : 0u;
        tot += u;
    }
    __syncthreads();  // s_w is reused by the next scan
    total = tot;
    return pre + x - v;
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts
// (in place) and the digit's total.
__global__ __launch_bounds__(kThreads) void k_tile_scan(uint32_t *__restrict__ hist, int ntiles,
                                                        uint32_t *__restrict__ dtot) {
    __shared__ uint32_t s_w[kWaves];
    uint32_t *row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int b = 0; b < ntiles; b += kThreads) {
        const int i = b + threadIdx.x;
        const uint32_t v = i < ntiles ? row[i] : 0;
        uint32_t chunk;
        const uint32_t ex = block_excl_scan(v, s_w, chunk);
        if (i < ntiles) row[i] = carry + ex;
        carry += chunk;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// The stable scatter of one tile.  Wave w holds the tile's items [w * 256,
// w * 256 + 256) in registers, 64 per round, so the (wave, round, lane)
// order is the index order.  Ranks: per wave, a running count per digit in
// LDS (this wave's row only, so the wave's own program order orders its LDS
// reads and writes) gives each item its rank among the wave's items of its
// digit; one barrier later the rows become, per digit, the tile's digit
// start plus the counts of the waves before.
__global__ __launch_bounds__(kThreads) void k_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                      uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, int n,
                                                      int shift, const uint32_t *__restrict__ hist, int ntiles,
                                                      const uint32_t *__restrict__ dtot) {
    __shared__ uint32_t wcnt[kWaves][kBins];
    __shared__ __align__(16) uint64_t s_k[kTileItems];
    __shared__ uint32_t s_v[kTileItems];
    __shared__ uint32_t s_gofs[kBins];  // global position - local position, per digit
    __shared__ uint32_t s_w[kWaves];
    constexpr int kPerWave = kTileItems / kWaves;  // 256
    const int t = threadIdx.x;
    const unsigned lane = __lane_id();
    const int wave = t >> 6;
    const int base = blockIdx.x * kTileItems;
    const int live_n = min(kTileItems, n - base);
    uint64_t k[kRounds];
    uint32_t v[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int li = wave * kPerWave + r * 64 + (int)lane;
        const int i = base + li;
        k[r] = li < live_n ? kin[i] : 0;
        v[r] = li < live_n ? vin[i] : 0;
    }
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    // ranks within the wave
    const uint64_t lt = lanemask_lt();
    uint32_t rk[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const bool live = wave * kPerWave + r * 64 + (int)lane < live_n;
        const unsigned d = (unsigned)(k[r] >> shift) & (kBins - 1);
        uint64_t peers = __ballot(live);
#pragma unroll
        for (int b = 0; b < kBits; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t before = wcnt[wave][d];
        rk[r] = before + (uint32_t)__popcll(peers & lt);
        // every lane has read the count before the leader of its digit adds
        // the round's peers (the wave's LDS accesses complete in order)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (live && (peers & lt) == 0) wcnt[wave][d] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    // per digit t: the tile's count, its start in the tile, the waves' bases
    uint32_t cnt[kWaves], tile_cnt = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        cnt[w] = wcnt[w][t];
        tile_cnt += cnt[w];
    }
    uint32_t unused;
    const uint32_t lstart = block_excl_scan(tile_cnt, s_w, unused);
    const uint32_t dt = dtot[t];
    const uint32_t dstart = block_excl_scan(dt, s_w, unused);
    s_gofs[t] = dstart + hist[t * ntiles + blockIdx.x] - lstart;
    uint32_t run = lstart;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        wcnt[w][t] = run;
        run += cnt[w];
    }
    __syncthreads();
    // the tile sorted by digit in LDS, then out in that order
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (wave * kPerWave + r * 64 + (int)lane < live_n) {
            const unsigned d = (unsigned)(k[r] >> shift) & (kBins - 1);
            const uint32_t pos = wcnt[wave][d] + rk[r];
            s_k[pos] = k[r];
            s_v[pos] = v[r];
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
// sorter's own.  Digits above `bits` are assumed zero; the passes start at
// bit lo_bit, so bits below it do not take part in the order.
std::pair<uint64_t *, uint32_t *> RadixSorter::passes(hipStream_t s, uint64_t *keys, uint32_t *vals, int64_t n,
                                                      int bits, int lo_bit) {
    if (n <= 1) return {keys, vals};
    if (n >= ((int64_t)1 << 31) - kTileItems) throw ApiError(NB_ERR_CAPACITY, "radix sort: too many items");
    ensure(n);
    const int nt = (int)((n + kTileItems - 1) / kTileItems);
    uint64_t *ka = keys, *kb = k2.p;
    uint32_t *va = vals, *vb = v2.p;
    for (int shift = lo_bit; shift < bits; shift += kBits) {
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
void RadixSorter::sort_swap(hipStream_t s, DBuf<uint64_t> &keys, DBuf<uint32_t> &vals, int64_t n, int bits,
                            int lo_bit) {
    const auto r = passes(s, keys.p, vals.p, n, bits, lo_bit);
    if (r.first != keys.p) {
        std::swap(keys.p, k2.p);
        std::swap(keys.n, k2.n);
        std::swap(vals.p, v2.p);
        std::swap(vals.n, v2.n);
    }
}

}  // namespace nbe
