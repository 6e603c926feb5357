// walk_emu.cpp -- CPU build of the FP32 Barnes-Hut walks
// (gregpackagescpp_amd/csrc/nb_walk_dev.hpp), for tests.
//
// One wave runs as 64 cooperatively scheduled contexts (ucontext stacks,
// switched with _setjmp/_longjmp; build without _FORTIFY_SOURCE).  Every wave
// intrinsic (ballot, shuffles, readfirstlane) and every fence is an exchange
// point: a context runs until it reaches one, and a round ends when all 64
// lanes wait at the same call site, whose values are then visible to every
// lane.  The scheduler stops the run when lanes wait at different sites or
// when one lane ends while others wait (on the GPU, an intrinsic reached by
// part of a wave).  Global buffers sit against PROT_NONE guard pages, so an
// access past a buffer's end names the buffer.  Workgroups run one after
// another, each wave of a workgroup in turn.  A step budget per
// launch (--max-rounds) turns a walk that does not end into an error that
// names the lane state.
//
// Input file (little endian): int64 n, int64 total, int64 bucket, float eps2,
//   float pad, float4 pos[n], WalkNode wnode[total], WalkNode wgrp[total].
// Output: per walk, the interaction count, the overflow flag and the FP32
// accelerations (double, SoA) written to OUT.
#include <math.h>
#include <setjmp.h>
#if defined(__SSE__)
#include <immintrin.h>
#endif
#include <signal.h>
#include <sys/mman.h>
#include <ucontext.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

// ---- the HIP vocabulary the device code uses, on the CPU ----
struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
static EmuDim3 threadIdx, blockIdx, gridDim;
struct float4 {
    float x, y, z, w;
};
struct int4 {
    int x, y, z, w;
};
struct int2 {
    int x, y;
};
static inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
static inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
static inline int2 make_int2(int a, int b) { return int2{a, b}; }
static inline float emu_rsqf(float x) {
#if defined(__SSE__)
    const float y = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
    return x == 0.0f ? INFINITY : y * (1.5f - 0.5f * x * y * y);
#else
    return 1.0f / std::sqrt(x);
#endif
}
static inline float __int_as_float(int v) {
    float f;
    memcpy(&f, &v, 4);
    return f;
}
static inline int __float_as_int(float f) {
    int v;
    memcpy(&v, &f, 4);
    return v;
}
#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define NB_WAVES_PER_EU(w)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define __builtin_amdgcn_rsqf emu_rsqf
#define __builtin_amdgcn_wave_barrier()
static inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
static inline int __ffsll(unsigned long long m) { return __builtin_ffsll((long long)m); }
static inline unsigned __lane_id() { return threadIdx.x & 63u; }
template <class T>
static inline T atomicAdd(T *p, T v) {
    T o = *p;
    *p = o + v;
    return o;
}
static inline int min(int a, int b) { return a < b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }

// exchange points, named by their source line
void emu_sync(int line);
uint64_t emu_ballot(bool p, int line);
int emu_xchg_i(int v, int src_lane, int line);
float emu_xchg_f(float v, int src_lane, int line);
#define __builtin_amdgcn_fence(order, scope) emu_sync(__LINE__)
#define __ballot(p) emu_ballot((p), __LINE__)
#define __builtin_amdgcn_readfirstlane(v) emu_xchg_i((int)(v), 0, __LINE__)
#define __shfl(v, src) emu_xchg_i((v), (int)(src), __LINE__)
#define __shfl_xor(v, o) emu_xchg_f((v), (int)(__lane_id() ^ (unsigned)(o)), __LINE__)
#define __shfl_up(v, o) emu_xchg_i((v), (int)__lane_id() >= (int)(o) ? (int)__lane_id() - (int)(o) : (int)__lane_id(), __LINE__)

constexpr int kBlock = 256;

namespace nbe {
#include "nb_walk_dev.hpp"
}  // namespace nbe

// ---- cooperative scheduler: one wave of 64 lanes ----
namespace {

constexpr int kL = 64;
constexpr size_t kStack = 256 * 1024;

struct Wave {
    ucontext_t main_ctx;
    ucontext_t ctx[kL];
    jmp_buf main_jb;
    jmp_buf jb[kL];
    char *stacks = nullptr;
    bool done[kL];
    int site[kL];
    int cur = 0;
    // exchange slots of the current round: the value each lane put in, and
    // the lane whose value it reads
    uint64_t put[kL];
    int src[kL];
    uint64_t got[kL];
    uint64_t snap[kL];  // the round's values, kept while lanes move on and overwrite put
    long long rounds = 0, max_rounds = 0;
    std::function<void()> body;
};
Wave *g_w = nullptr;
long long g_block = -1;
unsigned g_lane_base = 0;  // threadIdx.x of lane 0 (the wave's offset in its workgroup)

[[noreturn]] void die(const char *what) {
    fprintf(stderr, "EMU FAILURE: %s (block %lld, lane %d)\n", what, g_block, g_w ? g_w->cur : -1);
    fflush(stderr);
    _exit(3);
}

void trampoline() {
    Wave *w = g_w;
    if (!_setjmp(w->jb[w->cur])) _longjmp(w->main_jb, 1);
    g_w->body();
    g_w->done[g_w->cur] = true;
    _longjmp(g_w->main_jb, 1);
}

void run_wave(Wave &w, const std::function<void()> &body) {
    w.body = body;
    for (int t = 0; t < kL; ++t) {
        getcontext(&w.ctx[t]);
        w.ctx[t].uc_stack.ss_sp = w.stacks + (size_t)t * kStack;
        w.ctx[t].uc_stack.ss_size = kStack;
        w.ctx[t].uc_link = &w.main_ctx;
        makecontext(&w.ctx[t], trampoline, 0);
        w.done[t] = false;
        w.site[t] = 0;
    }
    g_w = &w;
    for (int t = 0; t < kL; ++t) {
        w.cur = t;
        threadIdx.x = g_lane_base + (unsigned)t;
        if (!_setjmp(w.main_jb)) swapcontext(&w.main_ctx, &w.ctx[t]);
    }
    while (true) {
        for (int t = 0; t < kL; ++t) {
            if (w.done[t]) continue;
            w.cur = t;
            threadIdx.x = g_lane_base + (unsigned)t;
            if (!_setjmp(w.main_jb)) _longjmp(w.jb[t], 1);
        }
        int waiting = 0, ended = 0, s0 = 0;
        for (int t = 0; t < kL; ++t) {
            if (w.done[t]) {
                ++ended;
                continue;
            }
            ++waiting;
            if (!s0) s0 = w.site[t];
            else if (w.site[t] != s0) {
                char msg[160];
                snprintf(msg, sizeof msg, "lanes of one wave wait at different exchange points (line %d, line %d at lane %d)",
                         s0, w.site[t], t);
                die(msg);
            }
        }
        if (!waiting) break;
        if (ended) die("a lane ended while others wait at an exchange point");
        for (int t = 0; t < kL; ++t) {
            w.snap[t] = w.put[t];
            w.got[t] = w.put[w.src[t] & 63];
        }
        if (++w.rounds > w.max_rounds) {
            char msg[160];
            snprintf(msg, sizeof msg, "more than %lld exchange rounds in one launch (line %d): the walk does not end",
                     w.max_rounds, s0);
            die(msg);
        }
    }
    g_w = nullptr;
}

struct Region {
    const char *name;
    uintptr_t lo, hi;
};
std::vector<Region> g_regions;

template <class T>
T *galloc(const char *name, size_t n) {
    const size_t page = (size_t)sysconf(_SC_PAGESIZE);
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    const size_t body = (bytes + page - 1) / page * page;
    char *base = (char *)mmap(nullptr, body + 2 * page, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (base == MAP_FAILED) die("mmap");
    mprotect(base, page, PROT_NONE);
    mprotect(base + page + body, page, PROT_NONE);
    g_regions.push_back(Region{name, (uintptr_t)base, (uintptr_t)(base + body + 2 * page)});
    char *p = base + page + body - bytes;
    memset(p, 0, bytes);
    return (T *)p;
}

void on_segv(int, siginfo_t *si, void *) {
    const uintptr_t a = (uintptr_t)si->si_addr;
    const char *which = "an address outside every buffer";
    for (const Region &r : g_regions)
        if (a >= r.lo && a < r.hi) which = r.name;
    char msg[256];
    snprintf(msg, sizeof msg, "memory access at %p, in the guard of %s", si->si_addr, which);
    die(msg);
}

template <class T>
void rd(FILE *f, T *p, size_t n) {
    if (fread(p, sizeof(T), n, f) != n) die("short input file");
}

inline void park(int line) {
    Wave *w = g_w;
    const int t = w->cur;
    w->site[t] = line;
    if (!_setjmp(w->jb[t])) _longjmp(w->main_jb, 1);
}

}  // namespace

__attribute__((noinline)) void emu_sync(int line) {
    g_w->put[g_w->cur] = 0;
    g_w->src[g_w->cur] = g_w->cur;
    park(line);
}

__attribute__((noinline)) uint64_t emu_ballot(bool p, int line) {
    Wave *w = g_w;
    w->put[w->cur] = p ? 1 : 0;
    w->src[w->cur] = w->cur;
    park(line);
    uint64_t m = 0;
    for (int t = 0; t < kL; ++t)
        if (g_w->snap[t]) m |= 1ull << t;
    return m;
}

__attribute__((noinline)) int emu_xchg_i(int v, int src_lane, int line) {
    Wave *w = g_w;
    w->put[w->cur] = (uint64_t)(uint32_t)v;
    w->src[w->cur] = src_lane;
    park(line);
    return (int)(uint32_t)g_w->got[g_w->cur];
}

__attribute__((noinline)) float emu_xchg_f(float v, int src_lane, int line) {
    return __int_as_float(emu_xchg_i(__float_as_int(v), src_lane, line));
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: walk_emu IN OUT [--cap N] [--max-rounds N]\n");
        return 2;
    }
    int32_t cap = nbe::kFStackDefault;
    long long max_rounds = 2000000000LL;
    for (int i = 3; i < argc; ++i) {
        if (!strcmp(argv[i], "--cap") && i + 1 < argc) cap = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--max-rounds") && i + 1 < argc) max_rounds = atoll(argv[++i]);
    }
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_segv;
    sa.sa_flags = SA_SIGINFO;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);

    FILE *f = fopen(argv[1], "rb");
    if (!f) die("cannot open input");
    int64_t n, total, bucket;
    float eps2, pad;
    rd(f, &n, 1);
    rd(f, &total, 1);
    rd(f, &bucket, 1);
    rd(f, &eps2, 1);
    rd(f, &pad, 1);
    float4 *pos = galloc<float4>("pos", (size_t)n);
    nbe::WalkNode *wnode = galloc<nbe::WalkNode>("wnode", (size_t)total);
    nbe::WalkNode *wgrp = galloc<nbe::WalkNode>("wgrp", (size_t)total);
    rd(f, pos, (size_t)n);
    rd(f, wnode, (size_t)total);
    rd(f, wgrp, (size_t)total);
    fclose(f);

    Wave w;
    w.stacks = (char *)malloc(kStack * kL);
    w.max_rounds = max_rounds;
    const bool soft = eps2 > 0.f;
    FILE *o = fopen(argv[2], "wb");
    if (!o) die("cannot open output");

    // depth-first wave walk: blocks of one wave, numbered as on the GPU
    {
        double *accs = galloc<double>("accs (wave walk)", 3 * (size_t)n);
        unsigned long long *icount = galloc<unsigned long long>("icount (wave walk)", 1);
        const unsigned nblk = (unsigned)((n + kBlock - 1) / kBlock);
        gridDim.x = nblk;
        for (unsigned b = 0; b < nblk; ++b)
            for (unsigned wv = 0; wv < kBlock / 64; ++wv) {
                // each wave of a 256-thread workgroup in turn, threadIdx offset by the wave
                blockIdx.x = b;
                g_block = b;
                g_lane_base = wv * 64;
                w.rounds = 0;
                run_wave(w, [&]() {
                    if (soft)
                        nbe::k_walk_wave<true>(n, (int32_t)total, pos, wnode, (int)bucket, eps2, accs, icount, nullptr,
                                               nullptr, nullptr, nullptr);
                    else
                        nbe::k_walk_wave<false>(n, (int32_t)total, pos, wnode, (int)bucket, eps2, accs, icount, nullptr,
                                                nullptr, nullptr, nullptr);
                });
            }
        printf("wave walk: interactions %llu\n", *icount);
        const int64_t cnt = (int64_t)*icount, zero = 0;
        fwrite(&cnt, 8, 1, o);
        fwrite(&zero, 8, 1, o);
        fwrite(accs, 8, 3 * (size_t)n, o);
    }
    // frontier walk: one wave per group, workgroups in a multiple of 8
    {
        // the product's grid (8 XCD ranges of at most 1.5x the mean count),
        // with uneven ranges so that every XCD maps a different span
        const int64_t ngroups = (n + 63) / 64;
        const int64_t mean = (ngroups + 7) / 8;
        const int64_t wgs = std::max<int64_t>(1, (mean + mean / 2 + nbe::kFWaves - 1) / nbe::kFWaves);
        const unsigned G = (unsigned)(8 * wgs);
        const int64_t span = wgs * nbe::kFWaves;
        int64_t *xr = galloc<int64_t>("xr", 9);
        xr[0] = 0;
        for (int x = 1; x < 8; ++x) {
            int64_t c = (int64_t)std::llround((double)ngroups * std::pow(x / 8.0, 1.3));
            c = std::max<int64_t>(c, xr[x - 1]);
            c = std::min<int64_t>(c, xr[x - 1] + span);
            c = std::max<int64_t>(c, ngroups - (int64_t)(8 - x) * span);
            xr[x] = std::min<int64_t>(c, ngroups);
        }
        xr[8] = ngroups;
        double *accs = galloc<double>("accs (frontier)", 3 * (size_t)n);
        unsigned long long *icount = galloc<unsigned long long>("icount (frontier)", 1);
        int32_t *stacks = galloc<int32_t>("stacks", (size_t)G * nbe::kFWaves * (size_t)cap);
        int32_t *order = galloc<int32_t>("order", (size_t)ngroups);
        uint32_t *cost = galloc<uint32_t>("cost", (size_t)ngroups);
        int *werr = galloc<int>("werr", 1);
        for (int64_t g = 0; g < ngroups; ++g) order[g] = (int32_t)g;
        gridDim.x = G;
        for (unsigned b = 0; b < G; ++b)
            for (unsigned wv = 0; wv < (unsigned)nbe::kFWaves; ++wv) {
                blockIdx.x = b;
                g_block = b;
                g_lane_base = wv * 64;
                w.rounds = 0;
                run_wave(w, [&]() {
                    if (soft)
                        nbe::k_walk_frontier<true>(n, (int32_t)total, pos, wgrp, (int)bucket, eps2, accs, icount,
                                                   stacks, cap, order, xr, cost, werr);
                    else
                        nbe::k_walk_frontier<false>(n, (int32_t)total, pos, wgrp, (int)bucket, eps2, accs, icount,
                                                    stacks, cap, order, xr, cost, werr);
                });
            }
        printf("frontier walk: interactions %llu overflow %d\n", *icount, *werr);
        const int64_t cnt = (int64_t)*icount, ov = *werr;
        fwrite(&cnt, 8, 1, o);
        fwrite(&ov, 8, 1, o);
        fwrite(accs, 8, 3 * (size_t)n, o);
    }
    // the depth-first walk again, reading the grouped layout (slot and skip
    // derived here: a node's children follow it in pre-order along the skip
    // chain and sit from its record's first-child slot on)
    {
        int32_t *skip = galloc<int32_t>("skip", (size_t)total);
        uint32_t *slot = galloc<uint32_t>("slot", (size_t)total);
        for (int64_t x = 0; x < total; ++x) {
            skip[x] = wnode[x].r.x;
            slot[x] = 0;
        }
        for (int64_t x = 0; x < total; ++x) {
            const int4 g = wgrp[slot[x]].r;
            if ((g.z & 15) == 0) continue;
            int32_t c = (int32_t)x + 1, k = 0;
            while (c < skip[x] && c < total) {
                slot[c] = (uint32_t)(g.x + k++);
                c = skip[c];
            }
        }
        double *accs = galloc<double>("accs (wave walk, grouped)", 3 * (size_t)n);
        unsigned long long *icount = galloc<unsigned long long>("icount (wave walk, grouped)", 1);
        const unsigned nblk = (unsigned)((n + kBlock - 1) / kBlock);
        gridDim.x = nblk;
        for (unsigned b = 0; b < nblk; ++b)
            for (unsigned wv = 0; wv < kBlock / 64; ++wv) {
                blockIdx.x = b;
                g_block = b;
                g_lane_base = wv * 64;
                w.rounds = 0;
                run_wave(w, [&]() {
                    if (soft)
                        nbe::k_walk_wave<true>(n, (int32_t)total, pos, nullptr, (int)bucket, eps2, accs, icount,
                                               nullptr, wgrp, slot, skip);
                    else
                        nbe::k_walk_wave<false>(n, (int32_t)total, pos, nullptr, (int)bucket, eps2, accs, icount,
                                                nullptr, wgrp, slot, skip);
                });
            }
        printf("wave walk (grouped layout): interactions %llu\n", *icount);
        const int64_t cnt = (int64_t)*icount, zero = 0;
        fwrite(&cnt, 8, 1, o);
        fwrite(&zero, 8, 1, o);
        fwrite(accs, 8, 3 * (size_t)n, o);
    }
    fclose(o);
    return 0;
}
