// nb_sort.hip -- LSD radix sort of (uint64 key, uint32 value) pairs for gfx950.
//
// Three kernels per 8-bit digit pass: per-tile histogram, one exclusive scan
// over the [digit][tile] counts, stable scatter.  The scatter ranks a tile's
// items round by round (256 items per round, in index order) with wave64
// ballots: a lane's peers are the lanes whose 8-bit digit equals its own, its
// rank is the popcount of the peers below it, and per-wave digit counts in LDS
// order the four waves.  No host synchronisation: a sort can be captured.
#include "nb_internal.hpp"
#include "nb_sort.hpp"

namespace nbe {

namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 16;
constexpr int kTileItems = kThreads * kRounds;  // 4096
constexpr int kWaves = kThreads / 64;

__global__ __launch_bounds__(kThreads) void k_hist(const uint64_t *__restrict__ keys, int n, int shift,
                                                   uint32_t *__restrict__ hist, int ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTileItems;
    for (int r = 0; r < kRounds; ++r) {
        int i = base + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Block-wide inclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t *sm) {
    const int t = threadIdx.x;
    sm[t] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        uint32_t u = t >= off ? sm[t - off] : 0;
        __syncthreads();
        sm[t] += u;
        __syncthreads();
    }
    uint32_t r = sm[t];
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
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[kWaves][256];
    const int t = threadIdx.x;
    const int wave = t >> 6;
    // digit bases: exclusive scan of the digit totals
    const uint32_t dt = dtot[t];
    const uint32_t dinc = block_incl_scan(dt, running);
    gbase[t] = dinc - dt + hist[t * ntiles + blockIdx.x];
    running[t] = 0;
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTileItems;
    const uint64_t lt = lanemask_lt();
    for (int r = 0; r < kRounds; ++r) {
        const int i = base + r * kThreads + t;
        const bool live = i < n;
        uint64_t k = live ? kin[i] : 0;
        uint32_t v = live ? vin[i] : 0;
        const unsigned d = (unsigned)(k >> shift) & 0xFFu;
        uint64_t peers = __ballot(live);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const unsigned rank = __popcll(peers & lt);
        if (live && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (live) {
            uint32_t off = running[d] + rank;
            for (int w = 0; w < wave; ++w) off += wcnt[w][d];
            kout[gbase[d] + off] = k;
            vout[gbase[d] + off] = v;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kWaves; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        running[t] += add;
        __syncthreads();
    }
}

}  // namespace

void RadixSorter::ensure(int64_t n) {
    const int nt = (int)((n + kTileItems - 1) / kTileItems);
    k2.ensure((size_t)std::max<int64_t>(n, 1));
    v2.ensure((size_t)std::max<int64_t>(n, 1));
    hist.ensure((size_t)256 * std::max(nt, 1));
    dtot.ensure(256);
}

// Sorts keys/vals in place (the result lands back in the caller's buffers).
// Digits above `bits` are assumed zero.
void RadixSorter::sort(hipStream_t s, uint64_t *keys, uint32_t *vals, int64_t n, int bits) {
    if (n <= 1) return;
    ensure(n);
    const int nt = (int)((n + kTileItems - 1) / kTileItems);
    uint64_t *ka = keys, *kb = k2.p;
    uint32_t *va = vals, *vb = v2.p;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_hist, dim3(nt), dim3(kThreads), 0, s, ka, (int)n, shift, hist.p, nt);
        hipLaunchKernelGGL(k_tile_scan, dim3(256), dim3(kThreads), 0, s, hist.p, nt, dtot.p);
        hipLaunchKernelGGL(k_scatter, dim3(nt), dim3(kThreads), 0, s, ka, va, kb, vb, (int)n, shift, hist.p, nt,
                           dtot.p);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    NB_HIP(hipGetLastError());
    if (ka != keys) {
        NB_HIP(hipMemcpyAsync(keys, ka, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        NB_HIP(hipMemcpyAsync(vals, va, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace nbe
