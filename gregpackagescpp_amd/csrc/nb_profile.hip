// nb_profile.hip -- per-kernel-class HIP event timing for nb_profile_*.
#include "nb_internal.hpp"

namespace nbe {

hipEvent_t Profiler::get() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    NB_HIP(hipEventCreate(&e));
    return e;
}

void Profiler::begin(const char *name, hipStream_t s, hipEvent_t *out) {
    (void)name;
    hipEvent_t e = get();
    NB_HIP(hipEventRecord(e, s));
    *out = e;
}

void Profiler::end(const char *name, hipStream_t s, hipEvent_t start) {
    hipEvent_t e = get();
    NB_HIP(hipEventRecord(e, s));
    recs[name].pending.emplace_back(start, e);
    // bound the number of outstanding event pairs
    if (recs[name].pending.size() > 4096) {
        NB_HIP(hipStreamSynchronize(s));
        collect();
    }
}

void Profiler::collect() {
    for (auto &kv : recs) {
        for (auto &pr : kv.second.pending) {
            float ms = 0.f;
            NB_HIP(hipEventSynchronize(pr.second));
            NB_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            kv.second.total_ms += ms;
            kv.second.launches += 1;
            pool.push_back(pr.first);
            pool.push_back(pr.second);
        }
        kv.second.pending.clear();
    }
}

void Profiler::clear() {
    collect();
    for (auto &kv : recs) {
        kv.second.total_ms = 0;
        kv.second.launches = 0;
    }
}

Profiler::~Profiler() {
    for (auto &kv : recs)
        for (auto &pr : kv.second.pending) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

}  // namespace nbe
