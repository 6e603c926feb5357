// nb_internal.hpp -- shared host/device declarations of the gfx950 N-body engine.
#pragma once
#include <hip/hip_runtime.h>
#include <array>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/gregnbody.h"

namespace nbe {

// Error carried from a failing HIP call up to the C ABI.
struct Status {
    int code = NB_OK;
    std::string msg;
};

#define NB_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            throw ::nbe::DeviceError(std::string(hipGetErrorName(_e)) + " (" + hipGetErrorString(_e) + ") in " + \
                                     #expr);                                                           \
        }                                                                                              \
    } while (0)

struct DeviceError {
    std::string what;
    explicit DeviceError(std::string w) : what(std::move(w)) {}
};
struct ApiError {
    int code;
    std::string what;
    ApiError(int c, std::string w) : code(c), what(std::move(w)) {}
};

// Blocking copies and fills go on the context's own stream, never on the
// legacy null stream: HIP refuses a null-stream copy while any stream of the
// process is being captured into a graph, and invalidates that capture, so
// one context's copy would break another thread's step graph.
inline void copy_wait(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
    NB_HIP(hipMemcpyAsync(dst, src, bytes, kind, st));
    NB_HIP(hipStreamSynchronize(st));
}
inline void zero_wait(void *dst, size_t bytes, hipStream_t st) {
    NB_HIP(hipMemsetAsync(dst, 0, bytes, st));
    NB_HIP(hipStreamSynchronize(st));
}

// Device buffer owned by the library.
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t count) {
        if (count <= n && p) return;
        release();
        if (count == 0) return;
        NB_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
};

// Per-kernel-class event timing (enabled by nb_profile_enable).
struct Profiler {
    bool on = false;
    struct Rec {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
        double total_ms = 0;
        int64_t launches = 0;
    };
    std::map<std::string, Rec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t get();
    void begin(const char *name, hipStream_t s, hipEvent_t *out);
    void end(const char *name, hipStream_t s, hipEvent_t start);
    void collect();  // after a stream sync
    void clear();
    ~Profiler();
};

struct ScopedTimer {
    Profiler &p;
    const char *name;
    hipStream_t s;
    hipEvent_t start = nullptr;
    ScopedTimer(Profiler &pp, const char *n, hipStream_t st) : p(pp), name(n), s(st) {
        if (p.on) p.begin(name, s, &start);
    }
    ~ScopedTimer() {
        if (p.on && start) p.end(name, s, start);
    }
};

// A sequence of launches captured from one stream and replayed with one
// graph launch (used for step loops of small systems, where launch latency
// dominates).  The captured code must not record events or memsets.
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    StepGraph() = default;
    StepGraph(const StepGraph &) = delete;
    StepGraph &operator=(const StepGraph &) = delete;
    ~StepGraph() { reset(); }
    bool valid() const { return exec != nullptr; }
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    // Returns false, with nothing launched and no graph kept, when the capture
    // was invalidated from outside (another thread's call that HIP refuses
    // during a capture); the caller then issues the same launches eagerly.
    template <typename F>
    bool capture(hipStream_t st, F &&f) {
        reset();
        NB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            f();
        } catch (...) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            (void)hipStreamIsCapturing(st, &cs);
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            if (cs == hipStreamCaptureStatusInvalidated) {
                (void)hipGetLastError();
                return false;
            }
            throw;
        }
        const hipError_t ee = hipStreamEndCapture(st, &g);
        if (ee == hipErrorStreamCaptureInvalidated) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return false;
        }
        NB_HIP(ee);
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            exec = nullptr;
            throw DeviceError(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        return true;
    }
    void launch(hipStream_t st) { NB_HIP(hipGraphLaunch(exec, st)); }
};

// Direct step loops of systems up to kGraphMaxBodies bodies replay a graph of
// kGraphSteps steps.
constexpr int kGraphSteps = 32;
constexpr int64_t kGraphMaxBodies = 16384;

struct GraphKey {
    int64_t n = -1;
    int precision = -1;
    std::array<const void *, 8> ptrs{};
    std::array<int, 5> shape{};
    std::array<double, 4> scalars{};
    bool operator==(const GraphKey &o) const {
        return n == o.n && precision == o.precision && ptrs == o.ptrs && shape == o.shape && scalars == o.scalars;
    }
};

// ---- direct summation (nb_direct.hip) ----
struct DirectWork {
    DBuf<float4> pos32;    // x, y, z, G*m in FP32
    DBuf<float> partial;   // [split][6][n]: Kahan sum and compensation per component
    int nsplit = 1;
    int ipl = 2;           // i-bodies per lane
    int variant = 1;       // 0: j stream through scalar loads, 1: j tiles through LDS
    DBuf<double> partial64;  // [split][3][n] for the FP64 split kernels
    DBuf<double> terms;      // [j block][3][n] pair terms of the FP64 reference-order path
    DBuf<double> px, pv, pa; // predicted state of modified Euler / midpoint (SoA, 3n)
    DBuf<unsigned long long> ovl_best;  // first overlapping pair of the construction check
    int nsplit64 = 1;
};

// Launchers (stream-ordered, no sync).
void launch_pack(hipStream_t s, int64_t n, const double *x, const double *m, double G, float4 *pos32);
void launch_force_mixed(hipStream_t s, int64_t n, const float4 *pos32, float eps2, bool soft, float *partial,
                        int nsplit, int ipl, int variant);
// Reduce the split partials to FP64 accelerations, then the closing half kick
// (v += half_dt*a); when `next` the next step's opening half kick and drift
// (v += half_dt*a; x += dt*v) and the FP32 repack follow in the same pass.
void launch_update_mixed(hipStream_t s, int64_t n, const float *partial, int nsplit, double *a, double *v, double *x,
                         const double *m, double G, double half_dt, double dt, int mode, float4 *pos32);
// FP64 reference-order forces.  For n <= kTermsMaxBodies (and a terms buffer of
// fp64_terms_block(n) * 3n doubles) the pair terms are evaluated in parallel
// and summed in order by a second kernel; otherwise one lane per i.
constexpr int64_t kTermsMaxBodies = 16384;
constexpr int64_t kTermsBudget = int64_t(32) << 20;  // doubles (256 MB)
int64_t fp64_terms_block(int64_t n);  // j per term block, 0 = one-lane-per-i kernel
void launch_force_fp64(hipStream_t s, int64_t n, const double *x, const double *m, double G, double eps2, int form,
                       double *a, double *terms);
void launch_kick(hipStream_t s, int64_t n, const double *a, double *v, double coef);
void launch_kick_drift(hipStream_t s, int64_t n, const double *a, double *v, double *x, double half_dt, double dt);
void launch_drift(hipStream_t s, int64_t n, const double *v, double *x, double dt);
// the other integrators' update kernels (reference operation order)
void launch_euler(hipStream_t s, int64_t n, const double *a, double *v, double *x, double dt);
void launch_predict(hipStream_t s, int64_t n, const double *x, const double *v, const double *a, double h,
                    double *px, double *pv);
void launch_modeuler(hipStream_t s, int64_t n, const double *a, const double *pv, const double *pa, double *v,
                     double *x, double half_dt);
void launch_midpoint(hipStream_t s, int64_t n, const double *pv, const double *pa, double *v, double *x, double dt);
void launch_pe_fp64(hipStream_t s, int64_t n, const double *x, const double *m, double G, double eps2, int form,
                    double *pe);
int choose_nsplit(int64_t n, int ipl, int waves_per_simd);
// FP64 j-split forces (deterministic split order, not the reference order)
int choose_nsplit64(int64_t n);
void launch_force_fp64_split(hipStream_t s, int64_t n, const double *x, const double *m, double eps2, bool soft,
                             double *partial, int nsplit);
void launch_update_fp64_split(hipStream_t s, int64_t n, const double *partial, int nsplit, double G, double *a,
                              double *v, double *x, double half_dt, double dt, int mode);
// As above with the drift reading xin and writing xout, and (when snap_x is
// set) the synchronised state x_q, v_q written to snap_x / snap_v on the way.
void launch_update_fp64_split_ex(hipStream_t s, int64_t n, const double *partial, int nsplit, double G, double *a,
                                 double *v, const double *xin, double *xout, double half_dt, double dt, int mode,
                                 double *snap_x, double *snap_v);

// update modes for launch_update_mixed
enum { UPD_STORE = 0, UPD_CLOSE = 1, UPD_CLOSE_OPEN = 2 };

}  // namespace nbe
