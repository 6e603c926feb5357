// nb_tune.hip -- tuning knobs and the project autotuner (engine extension).
//
// Candidates are timed on the context's own bodies with HIP events on its
// stream; the fastest setting is kept in the context.  An optional text cache
// holds lines "kernel n precision knob value".
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "nb_system.hpp"

using namespace nbe;

namespace {

int *knob(nb_system *s, const char *key) {
    if (!std::strcmp(key, "direct_ipl")) return &s->tune_direct_ipl;
    if (!std::strcmp(key, "direct_waves")) return &s->tune_direct_waves;
    if (!std::strcmp(key, "direct_kernel")) return &s->tune_direct_kernel;
    if (!std::strcmp(key, "step_graphs")) return &s->tune_step_graphs;
    if (!std::strcmp(key, "bh_bucket")) return &s->tune_bh_bucket;
    if (!std::strcmp(key, "bh_walk")) return &s->tune_bh_walk;
    if (!std::strcmp(key, "bh_stack")) return &s->tune_bh_stack;
    if (!std::strcmp(key, "bh_sort_passes")) return &s->tune_bh_sort;
    return nullptr;
}

// Shortest of `reps` timed launches of fn on the stream (ms), after one
// untimed warm-up (code object load, allocations); the minimum is the least
// noisy estimate when candidates differ by a few percent.
template <typename F>
double time_ms(hipStream_t st, int reps, F &&fn) {
    hipEvent_t a, b;
    NB_HIP(hipEventCreate(&a));
    NB_HIP(hipEventCreate(&b));
    fn();
    double best = 1e30;
    for (int r = 0; r < reps; ++r) {
        NB_HIP(hipEventRecord(a, st));
        fn();
        NB_HIP(hipEventRecord(b, st));
        NB_HIP(hipEventSynchronize(b));
        float ms = 0.f;
        NB_HIP(hipEventElapsedTime(&ms, a, b));
        best = ms < best ? ms : best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return best;
}

bool cache_lookup(const char *path, const std::string &kernel, int64_t n, int prec, const char *k, int *out) {
    if (!path) return false;
    std::ifstream f(path);
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string kk, name;
        int64_t nn;
        int pp, v;
        if (is >> kk >> nn >> pp >> name >> v)
            if (kk == kernel && nn == n && pp == prec && name == k) {
                *out = v;
                found = true;
            }
    }
    return found;
}

void cache_store(const char *path, const std::string &kernel, int64_t n, int prec, const char *k, int v) {
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << kernel << ' ' << n << ' ' << prec << ' ' << k << ' ' << v << '\n';
}

}  // namespace

namespace nbe {
void bh_forces(nb_system *s, int64_t *interactions);
}

extern "C" {

int nb_set_tuning(nb_system *s, const char *key, int64_t value) {
    if (!s || !key) return NB_ERR_INVALID_ARGUMENT;
    int *k = knob(s, key);
    if (!k) {
        s->err = std::string("unknown tuning knob ") + key;
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (!std::strcmp(key, "direct_ipl") && value != 0 && value != 1 && value != 2 && value != 4) {
        s->err = "direct_ipl must be 1, 2 or 4";
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (!std::strcmp(key, "direct_kernel") && value != 0 && value != 1 && value != 2) {
        s->err = "direct_kernel must be 1 (scalar-load j stream) or 2 (LDS j tiles)";
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (!std::strcmp(key, "step_graphs") && value != 0 && value != 1 && value != 2) {
        s->err = "step_graphs must be 1 (on) or 2 (off)";
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (!std::strcmp(key, "bh_walk") && value != 0 && value != 1 && value != 2) {
        s->err = "bh_walk must be 1 (frontier walk) or 2 (per-wave DFS walk)";
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (!std::strcmp(key, "bh_sort_passes") && value != 0 && (value < 4 || value > 7)) {
        s->err = "bh_sort_passes must be between 4 and 7";
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (!std::strcmp(key, "bh_stack") && value != 0 && (value < 64 || value > 4096)) {
        s->err = "bh_stack must be between 64 and 4096";
        return NB_ERR_INVALID_ARGUMENT;
    }
    if (value < 0 || value > 4096) {
        s->err = "tuning value out of range";
        return NB_ERR_INVALID_ARGUMENT;
    }
    *k = (int)value;
    s->acc_valid = s->acc_valid && std::strcmp(key, "direct_ipl") && std::strcmp(key, "direct_waves") &&
                   std::strcmp(key, "direct_kernel");
    return NB_OK;
}

int nb_get_tuning(const nb_system *cs, const char *key, int64_t *value) {
    nb_system *s = const_cast<nb_system *>(cs);
    if (!s || !key || !value) return NB_ERR_INVALID_ARGUMENT;
    int *k = knob(s, key);
    if (!k) return NB_ERR_INVALID_ARGUMENT;
    *value = *k;
    // the adaptive sort reports the passes its next build takes
    if (!std::strcmp(key, "bh_sort_passes") && *k == 0) *value = bh_next_sort_passes(s);
    return NB_OK;
}

int nb_autotune(nb_system *s, int method, const char *cache_path) {
    if (!s) return NB_ERR_INVALID_ARGUMENT;
    try {
        NB_HIP(hipSetDevice(s->device));
        const int64_t n = s->n;
        if (n < 2) return NB_OK;
        hipStream_t st = s->stream;
        const bool prof = s->prof.on;
        s->prof.on = false;
        if ((method & NB_BARNES_HUT) == NB_BARNES_HUT) {
            if (s->precision == NB_PREC_MIXED) {
                int v;
                if (cache_lookup(cache_path, "bh_walk", n, s->precision, "bh_bucket", &v)) {
                    s->tune_bh_bucket = v;
                } else {
                    double best = 1e30;
                    int arg = 16;
                    for (int b : {1, 4, 8, 12, 16, 32, 64}) {
                        s->tune_bh_bucket = b;
                        double ms = time_ms(st, 5, [&]() { bh_forces(s, nullptr); });
                        if (ms < best) {
                            best = ms;
                            arg = b;
                        }
                    }
                    s->tune_bh_bucket = arg;
                    cache_store(cache_path, "bh_walk", n, s->precision, "bh_bucket", arg);
                }
            }
        } else if (s->precision == NB_PREC_MIXED) {
            int vk, vi, vw;
            if (cache_lookup(cache_path, "direct", n, s->precision, "direct_kernel", &vk) &&
                cache_lookup(cache_path, "direct", n, s->precision, "direct_ipl", &vi) &&
                cache_lookup(cache_path, "direct", n, s->precision, "direct_waves", &vw)) {
                s->tune_direct_kernel = vk;
                s->tune_direct_ipl = vi;
                s->tune_direct_waves = vw;
            } else {
                double best = 1e30;
                int bk = 2, bi = 2, bw = 8;
                const int reps = n >= 32768 ? 4 : 20;
                for (int kern : {1, 2})
                    for (int ipl : {1, 2, 4})
                        for (int waves : {4, 8, 16}) {
                            s->tune_direct_kernel = kern;
                            s->tune_direct_ipl = ipl;
                            s->tune_direct_waves = waves;
                            double ms = time_ms(st, reps, [&]() { direct_forces(s); });
                            if (ms < best) {
                                best = ms;
                                bk = kern;
                                bi = ipl;
                                bw = waves;
                            }
                        }
                s->tune_direct_kernel = bk;
                s->tune_direct_ipl = bi;
                s->tune_direct_waves = bw;
                cache_store(cache_path, "direct", n, s->precision, "direct_kernel", bk);
                cache_store(cache_path, "direct", n, s->precision, "direct_ipl", bi);
                cache_store(cache_path, "direct", n, s->precision, "direct_waves", bw);
            }
        }
        s->prof.on = prof;
        s->acc_valid = false;  // accelerations were recomputed for timing only
        bh_restore_order(s);   // the timed builds left the bodies in tree order
        NB_HIP(hipStreamSynchronize(st));
        return NB_OK;
    } catch (const ApiError &e) {
        s->err = e.what;
        restore_order_quietly(s);
        return e.code;
    } catch (const DeviceError &e) {
        s->err = e.what;
        restore_order_quietly(s);
        return NB_ERR_DEVICE;
    }
}

}  // extern "C"
