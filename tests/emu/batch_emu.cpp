// batch_emu.cpp -- CPU build of the collision pass and the batched ensemble
// kernel (gregpackagescpp_amd/csrc/nb_coll_dev.hpp), for tests.
//
// Each workgroup runs as kThreads cooperatively scheduled contexts (ucontext
// stacks, entered once with swapcontext and switched with _setjmp/_longjmp,
// which skip the signal-mask system call; build without _FORTIFY_SOURCE):
// a context runs until it reaches __syncthreads() or ends, and a barrier
// round ends when every live context has reached one.  The scheduler checks
// that all contexts of a round wait at the same barrier (call site) and that
// none has ended while others wait, which on the GPU would be a divergent
// barrier.  Global buffers sit against PROT_NONE guard pages (at their end,
// or with --guard start at their start), so an access past either edge stops
// the run with the buffer, system and thread named.  Systems are independent
// and are spread over host threads, one workgroup at a time per thread.
//
// What this cannot show: interleavings inside a barrier interval (contexts
// run one after another), and the GPU's FP32 rsqrt rounding (1/sqrtf here),
// so mixed-precision trajectories differ from the GPU's in the last bits.
//
// Input file (little endian): int64 nsys, int64 steps, int64 chunk,
//   double G, dt, eps, min_mom, int64 coll, merge, mixed,
//   int64 offsets[nsys+1], double m[T], r[T], e[T], pos[3T], vel[3T].
// Output file: int64 err, int64 counts[nsys], then per system its bodies
//   (double m, r, x, y, z, vx, vy, vz, int64 label) and int64 nlog, then the
//   merge log records (system, step, outer label, absorbed label, new label).
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
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

// ---- the HIP vocabulary the device code uses, on the CPU ----
struct EmuDim3 {
    unsigned x = 0, y = 0, z = 0;
};
static thread_local EmuDim3 threadIdx, blockIdx;
struct float4 {
    float x, y, z, w;
};
static inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
static inline float emu_rsqf(float x) {
#if defined(__SSE__)
    // the hardware estimate refined by one Newton step (the GPU's v_rsq_f32 is
    // also an approximation; neither matches 1/sqrt to the last bit)
    const float y = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
    return x == 0.0f ? INFINITY : y * (1.5f - 0.5f * x * y * y);
#else
    return 1.0f / std::sqrt(x);
#endif
}
#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...)
#define __shared__ static thread_local
#define __align__(n) __attribute__((aligned(n)))
#define __builtin_amdgcn_rsqf emu_rsqf
#define NB_OPAQUE(x) __asm__ volatile("" : "+r"(x))
// a barrier is named by its source line, which code duplication by the host
// compiler cannot change
void emu_sync(int line);
int emu_sync_or(int predicate, int line);
#define __syncthreads() emu_sync(__LINE__)
#define __syncthreads_or(p) emu_sync_or((p), __LINE__)
template <class T>
static inline T atomicAdd(T *p, T v) {
    T o = *p;
    *p = o + v;
    return o;
}
template <class T>
static inline T atomicMin(T *p, T v) {
    T o = *p;
    if (v < o) *p = v;
    return o;
}
template <class T>
static inline T atomicExch(T *p, T v) {
    T o = *p;
    *p = v;
    return o;
}
using std::max;
using std::min;

namespace nbe {
#include "nb_coll_dev.hpp"
}  // namespace nbe

// ---- cooperative scheduler ----
namespace {

constexpr int kT = nbe::kThreads;
constexpr size_t kStack = 256 * 1024;

struct Block {
    ucontext_t main_ctx;
    ucontext_t ctx[kT];
    jmp_buf main_jb;
    jmp_buf jb[kT];
    char *stacks = nullptr;
    bool done[kT];
    int site[kT];
    int cur = 0;
    int or_acc = 0, or_res = 0;
    std::function<void()> body;
};
thread_local Block *g_blk = nullptr;
thread_local int64_t g_sys = -1, g_launch = -1;

[[noreturn]] void die(const char *what) {
    fprintf(stderr, "EMU FAILURE: %s (system %lld, launch %lld, thread %d)\n", what, (long long)g_sys,
            (long long)g_launch, g_blk ? g_blk->cur : -1);
    fflush(stderr);
    _exit(3);
}

// entered once per context by swapcontext: park at once, then run the body
// when first switched to, and end by jumping back to the scheduler
void trampoline() {
    Block *b = g_blk;
    if (!_setjmp(b->jb[b->cur])) _longjmp(b->main_jb, 1);
    g_blk->body();
    g_blk->done[g_blk->cur] = true;
    _longjmp(g_blk->main_jb, 1);
}

void run_block(Block &b, const std::function<void()> &body) {
    b.body = body;
    for (int t = 0; t < kT; ++t) {
        getcontext(&b.ctx[t]);
        b.ctx[t].uc_stack.ss_sp = b.stacks + (size_t)t * kStack;
        b.ctx[t].uc_stack.ss_size = kStack;
        b.ctx[t].uc_link = &b.main_ctx;
        makecontext(&b.ctx[t], trampoline, 0);
        b.done[t] = false;
        b.site[t] = 0;
    }
    b.or_acc = b.or_res = 0;
    g_blk = &b;
    for (int t = 0; t < kT; ++t) {  // park every context at its start
        b.cur = t;
        if (!_setjmp(b.main_jb)) swapcontext(&b.main_ctx, &b.ctx[t]);
    }
    while (true) {
        for (int t = 0; t < kT; ++t) {
            if (b.done[t]) continue;
            b.cur = t;
            threadIdx.x = (unsigned)t;
            if (!_setjmp(b.main_jb)) _longjmp(b.jb[t], 1);
        }
        int waiting = 0, ended = 0, s0 = 0;
        for (int t = 0; t < kT; ++t) {
            if (b.done[t]) {
                ++ended;
                continue;
            }
            ++waiting;
            if (!s0) s0 = b.site[t];
            else if (b.site[t] != s0) {
                char msg[160];
                snprintf(msg, sizeof msg, "threads of one workgroup wait at different barriers (line %d: thread 0.., "
                         "line %d: thread %d)", s0, b.site[t], t);
                die(msg);
            }
        }
        if (!waiting) break;
        if (ended) die("a thread ended while others wait at a barrier");
        b.or_res = b.or_acc;
        b.or_acc = 0;
    }
    g_blk = nullptr;
}

// ---- guarded global buffers ----
struct Region {
    const char *name;
    uintptr_t lo, hi;  // mapped span including guards
};
std::vector<Region> g_regions;
bool g_guard_start = false;

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
    char *p = g_guard_start ? base + page : base + page + body - bytes;
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

}  // namespace

__attribute__((noinline)) void emu_sync(int line) {
    Block *b = g_blk;
    const int t = b->cur;
    b->site[t] = line;
    if (!_setjmp(b->jb[t])) _longjmp(b->main_jb, 1);
}

__attribute__((noinline)) int emu_sync_or(int predicate, int line) {
    Block *b = g_blk;
    b->or_acc |= predicate != 0;
    const int t = b->cur;
    b->site[t] = line;
    if (!_setjmp(b->jb[t])) _longjmp(b->main_jb, 1);
    return g_blk->or_res;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: batch_emu IN OUT [--guard end|start] [--threads N] [--progress EVERY]\n");
        return 2;
    }
    int nthreads = 8;
    int64_t progress = 0;
    for (int i = 3; i < argc; ++i) {
        if (!strcmp(argv[i], "--guard") && i + 1 < argc) g_guard_start = !strcmp(argv[++i], "start");
        else if (!strcmp(argv[i], "--threads") && i + 1 < argc) nthreads = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--progress") && i + 1 < argc) progress = atoll(argv[++i]);
    }
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_segv;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);

    FILE *f = fopen(argv[1], "rb");
    if (!f) die("cannot open input");
    int64_t nsys, steps, chunk, coll, merge, mixed;
    double G, dt, eps, min_mom;
    rd(f, &nsys, 1); rd(f, &steps, 1); rd(f, &chunk, 1);
    rd(f, &G, 1); rd(f, &dt, 1); rd(f, &eps, 1); rd(f, &min_mom, 1);
    rd(f, &coll, 1); rd(f, &merge, 1); rd(f, &mixed, 1);
    std::vector<int64_t> offsets((size_t)nsys + 1);
    rd(f, offsets.data(), offsets.size());
    const int64_t T = offsets[(size_t)nsys];
    std::vector<double> m(T), r(T), e(T), pos(3 * T), vel(3 * T);
    rd(f, m.data(), T); rd(f, r.data(), T); rd(f, e.data(), T); rd(f, pos.data(), 3 * T); rd(f, vel.data(), 3 * T);
    fclose(f);

    // device buffers, laid out as nb_batch_create lays them out
    double *x = galloc<double>("x", 3 * T), *v = galloc<double>("v", 3 * T), *a = galloc<double>("a", 3 * T);
    double *dm = galloc<double>("m", T), *dr = galloc<double>("r", T), *de = galloc<double>("e", T);
    int64_t *lab = galloc<int64_t>("labels", T), *log = galloc<int64_t>("log", 4 * T);
    int64_t *off = galloc<int64_t>("offsets", nsys + 1), *next = galloc<int64_t>("next_labels", nsys);
    int64_t *loglen = galloc<int64_t>("log_lens", nsys);
    int *cnt = galloc<int>("counts", nsys), *ord = galloc<int>("order", nsys), *err = galloc<int>("err", 1);
    for (int64_t i = 0; i < T; ++i) {
        for (int c = 0; c < 3; ++c) {
            x[c * T + i] = pos[3 * i + c];
            v[c * T + i] = vel[3 * i + c];
        }
        dm[i] = m[i]; dr[i] = r[i]; de[i] = e[i];
    }
    for (int64_t s = 0; s < nsys; ++s) {
        off[s] = offsets[s];
        const int64_t k = offsets[s + 1] - offsets[s];
        if (k > nbe::kBatchMaxN) die("a system holds more than kBatchMaxN bodies");
        cnt[s] = (int)k;
        next[s] = k;
        ord[s] = (int)s;
        for (int64_t q = 0; q < k; ++q) lab[offsets[s] + q] = q;
    }
    off[nsys] = offsets[nsys];
    const nbe::BodyView all{x, x + T, x + 2 * T, v, v + T, v + 2 * T, a, a + T, a + 2 * T, dm, dr, de, lab};
    nbe::BatchParams P0{G, dt, dt / 2, eps * eps, (int)coll, (int)merge, min_mom, (int)mixed};
    nbe::BatchParams Pf = P0;
    Pf.dt = 0.0;
    Pf.half_dt = 0.0;
    Pf.coll = 0;

    // nb_batch_evolve: a force-only zero-dt launch, then launches of `chunk`
    // steps; systems are taken largest first, as the GPU dispatch order has them
    std::vector<int64_t> by_size((size_t)nsys);
    for (int64_t s = 0; s < nsys; ++s) by_size[(size_t)s] = s;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int64_t p, int64_t q) { return cnt[p] > cnt[q]; });
    // the instance nb_batch_evolve picks: precision, and 64-body blocks per
    // wave for the largest system
    int max_n = 0;
    for (int64_t s = 0; s < nsys; ++s) max_n = std::max(max_n, cnt[s]);
    const bool small = max_n <= nbe::batch_capacity<3>();
    auto launch = [&](const nbe::BatchParams &P, int64_t step0, int64_t k) {
        if (mixed) {
            if (small) nbe::k_batch_steps<true, 3>(all, ord, off, cnt, next, log, loglen, err, P, step0, k);
            else nbe::k_batch_steps<true, 4>(all, ord, off, cnt, next, log, loglen, err, P, step0, k);
        } else {
            if (small) nbe::k_batch_steps<false, 3>(all, ord, off, cnt, next, log, loglen, err, P, step0, k);
            else nbe::k_batch_steps<false, 4>(all, ord, off, cnt, next, log, loglen, err, P, step0, k);
        }
    };
    std::atomic<int64_t> next_sys{0}, finished{0};
    auto worker = [&]() {
        Block b;
        b.stacks = (char *)malloc((size_t)kT * kStack);
        if (!b.stacks) die("stack allocation");
        for (int64_t q; (q = next_sys.fetch_add(1)) < nsys;) {
            const int64_t s = by_size[(size_t)q];
            g_sys = s;
            blockIdx.x = (unsigned)s;
            g_launch = 0;
            run_block(b, [&]() { launch(Pf, 0, 1); });
            for (int64_t done = 0; done < steps;) {
                const int64_t k = std::min(chunk, steps - done);
                ++g_launch;
                run_block(b, [&]() { launch(P0, done, k); });
                done += k;
            }
            const int64_t f = finished.fetch_add(1) + 1;
            if (progress && f % progress == 0) {
                fprintf(stderr, "emu: %lld of %lld systems done\n", (long long)f, (long long)nsys);
                fflush(stderr);
            }
        }
        free(b.stacks);
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < std::max(1, nthreads); ++i) pool.emplace_back(worker);
    for (auto &th : pool) th.join();

    FILE *o = fopen(argv[2], "wb");
    if (!o) die("cannot open output");
    const int64_t errv = *err;
    fwrite(&errv, 8, 1, o);
    for (int64_t s = 0; s < nsys; ++s) {
        const int64_t c = cnt[s];
        fwrite(&c, 8, 1, o);
    }
    for (int64_t s = 0; s < nsys; ++s) {
        for (int64_t q = 0; q < cnt[s]; ++q) {
            const int64_t i = offsets[s] + q;
            const double rec[8] = {dm[i], dr[i], x[i], x[T + i], x[2 * T + i], v[i], v[T + i], v[2 * T + i]};
            fwrite(rec, 8, 8, o);
            fwrite(&lab[i], 8, 1, o);
        }
    }
    int64_t nlog = 0;
    for (int64_t s = 0; s < nsys; ++s) nlog += std::min<int64_t>(loglen[s], offsets[s + 1] - offsets[s]);
    fwrite(&nlog, 8, 1, o);
    for (int64_t s = 0; s < nsys; ++s) {
        const int64_t k = std::min<int64_t>(loglen[s], offsets[s + 1] - offsets[s]);
        for (int64_t q = 0; q < k; ++q) {
            const int64_t *L = log + 4 * (offsets[s] + q);
            const int64_t rec[5] = {s, L[0], L[1], L[2], L[3]};
            fwrite(rec, 8, 5, o);
        }
    }
    fclose(o);
    return 0;
}
