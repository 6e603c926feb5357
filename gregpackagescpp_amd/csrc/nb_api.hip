// nb_api.hip -- C ABI (include/gregnbody.h) over the gfx950 kernels.
//
// Mirrors gtd::system's public surface (gregsys.hpp:1478-1760, 2593-2607):
// construction from bodies, setters with the reference's validation rules,
// evolve(method) with a fresh a(x0) per call, state read-back.  Every entry
// point converts library exceptions into status codes.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

#include "nb_system.hpp"

using namespace nbe;

nb_system::~nb_system() {
    if (bh) bh_work_free(bh);
    if (coll) coll_work_free(coll);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

thread_local std::string g_static_err;

// Device scratch of nb_advance_noalloc: the split partials of the direct
// force pass live in this file-scope array instead of context buffers, so
// that entry point allocates nothing.  Each device holds its own copy until
// the process ends.  In host code `scratch` is only the symbol's host-side
// stand-in: its device address comes from hipGetSymbolAddress, resolved once
// per device at nb_create (with that device current) and kept in
// g_scratch_ptr; host code passes that pointer, never the array's name.
__device__ __align__(256) unsigned char scratch[NB_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device index

void resolve_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[(size_t)device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[(size_t)device] = p;
}

void *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device] : nullptr;
}

}  // namespace

namespace nbe {
// message for nb_last_error(NULL) from host-only entry points in other files
void set_static_error(const std::string &msg) { g_static_err = msg; }
}  // namespace nbe

namespace {

int fail(nb_system *s, int code, const std::string &msg) {
    if (s) s->err = msg;
    else g_static_err = msg;
    return code;
}

// Runs an entry point's body, turning library exceptions into status codes.
// Barnes-Hut steps keep the bodies in the tree's order between steps; every
// entry point returns with them back in input order.
template <typename F>
int guarded(nb_system *s, F &&f) {
    try {
        if (s) NB_HIP(hipSetDevice(s->device));
        const int rc = f();
        if (s) bh_restore_order(s);
        return rc;
    } catch (const ApiError &e) {
        if (s) restore_order_quietly(s);
        return fail(s, e.code, e.what);
    } catch (const DeviceError &e) {
        if (s) restore_order_quietly(s);
        return fail(s, NB_ERR_DEVICE, e.what);
    } catch (const std::bad_alloc &) {
        if (s) restore_order_quietly(s);
        return fail(s, NB_ERR_CAPACITY, "host allocation failed");
    }
}

// parse_units_format (gregsys.hpp:245-282): "Ma:b,Dc:x,Ty:z" with positive
// integers of 1-18 digits containing a non-zero digit; G = 66743 *
// (m_num d_num^3 t_num^2) / (m_den d_den^3 t_den^2) / 1e15, where the integer
// products are unsigned 64-bit and the arithmetic is long double.
bool parse_units(const char *str, long double *G, uint64_t vals[6]) {
    if (!str) return false;
    const char *p = str;
    auto skip = [&]() { while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v') ++p; };
    auto num = [&](uint64_t *out) -> bool {
        skip();
        const char *b = p;
        uint64_t v = 0;
        bool nz = false;
        while (*p >= '0' && *p <= '9') {
            v = v * 10 + (uint64_t)(*p - '0');
            if (*p != '0') nz = true;
            ++p;
        }
        size_t len = (size_t)(p - b);
        if (len < 1 || len > 18 || !nz) return false;
        *out = v;
        skip();
        return true;
    };
    auto lit = [&](char a, char b2) -> bool {
        skip();
        if (*p != a && *p != b2) return false;
        ++p;
        return true;
    };
    uint64_t a, b, c, x, y, z;
    if (!lit('m', 'M') || !num(&a) || !lit(':', ':') || !num(&b) || !lit(',', ',')) return false;
    if (!lit('d', 'D') || !num(&c) || !lit(':', ':') || !num(&x) || !lit(',', ',')) return false;
    if (!lit('t', 'T') || !num(&y) || !lit(':', ':') || !num(&z)) return false;
    skip();
    if (*p != '\0') return false;
    uint64_t m_den = a, m_num = b, d_num = c, d_den = x, t_den = y, t_num = z;
    long double g = 66743;
    g *= m_num * d_num * d_num * d_num * t_num * t_num;
    g /= m_den * d_den * d_den * d_den * t_den * t_den;
    g /= 1000000000000000ULL;
    *G = g;
    vals[0] = m_num; vals[1] = m_den; vals[2] = d_num; vals[3] = d_den; vals[4] = t_num; vals[5] = t_den;
    return true;
}

// check_option (gregsys.hpp:1406-1411)
bool check_option(int option) {
    uint16_t lo = (uint16_t)(option & 0x0000ffff);
    uint16_t hi = (uint16_t)((unsigned)option >> 16);
    return !(lo & (lo - 1)) || !lo || lo > NB_RK4 || !(hi & (hi - 1)) || hi > 2;
}

void to_soa(int64_t n, const double *aos, std::vector<double> &soa) {
    soa.resize((size_t)(3 * n));
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) soa[(size_t)(c * n + i)] = aos[3 * i + c];
}
void from_soa(int64_t n, const std::vector<double> &soa, double *aos) {
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) aos[3 * i + c] = soa[(size_t)(c * n + i)];
}

}  // namespace

namespace nbe {

int form_of(const nb_system *s) { return s->eps > 0 ? 1 : 0; }

void run_hook(nb_system *s, uint64_t step, int where) {
    if (!s->hook_fn || !(s->hook_mask & where)) return;
    NB_HIP(hipStreamSynchronize(s->stream));
    s->in_hook = true;
    int rc;
    try {
        rc = s->hook_fn(s->hook_user, s, step, where);
    } catch (...) {
        rc = -1;  // a hook must not throw across the C ABI; treat it as a failure
    }
    s->in_hook = false;
    if (rc != 0) throw ApiError(NB_ERR_CALLBACK, "step hook returned " + std::to_string(rc) + " at step " +
                                                     std::to_string(step));
}

void direct_forces(nb_system *s) {
    const int64_t n = s->n;
    if (s->precision == NB_PREC_MIXED) {
        s->dw.pos32.ensure((size_t)n);
        s->dw.ipl = s->tune_direct_ipl > 0 ? s->tune_direct_ipl : 2;
        s->dw.variant = s->tune_direct_kernel == 1 ? 0 : 1;
        s->dw.nsplit = choose_nsplit(n, s->dw.ipl, s->tune_direct_waves);
        s->dw.partial.ensure((size_t)s->dw.nsplit * 6 * (size_t)n);
        launch_pack(s->stream, n, s->x.p, s->m.p, s->G, s->dw.pos32.p);
        {
            ScopedTimer t(s->prof, "force", s->stream);
            launch_force_mixed(s->stream, n, s->dw.pos32.p, (float)(s->eps * s->eps), form_of(s) == 1,
                               s->dw.partial.p, s->dw.nsplit, s->dw.ipl, s->dw.variant);
        }
        ScopedTimer t(s->prof, "update", s->stream);
        launch_update_mixed(s->stream, n, s->dw.partial.p, s->dw.nsplit, s->a.p, s->v.p, s->x.p, s->m.p, s->G,
                            s->half_dt, s->dt, UPD_STORE, s->dw.pos32.p);
    } else if (s->precision == NB_PREC_FP64_FAST) {
        s->dw.nsplit64 = choose_nsplit64(n);
        s->dw.partial64.ensure((size_t)s->dw.nsplit64 * 3 * (size_t)n);
        {
            ScopedTimer t(s->prof, "force", s->stream);
            launch_force_fp64_split(s->stream, n, s->x.p, s->m.p, s->eps * s->eps, form_of(s) == 1,
                                    s->dw.partial64.p, s->dw.nsplit64);
        }
        launch_update_fp64_split(s->stream, n, s->dw.partial64.p, s->dw.nsplit64, s->G, s->a.p, s->v.p, s->x.p,
                                 s->half_dt, s->dt, UPD_STORE);
    } else {
        s->dw.terms.ensure((size_t)(fp64_terms_block(n) * 3 * n));
        ScopedTimer t(s->prof, "force", s->stream);
        launch_force_fp64(s->stream, n, s->x.p, s->m.p, s->G, s->eps * s->eps, form_of(s), s->a.p, s->dw.terms.p);
    }
    s->acc_valid = true;
}

// KDK steps from valid accelerations (evolve's KDK branch :2949-2968 with
// leapfrog_kdk_acc_e_and_step :514-526 and s_coll when collisions are on).
// a(x) for arbitrary positions (SoA, 3n) into `a`, in the context's precision
// (the predicted-state forces of modified Euler and midpoint).
void forces_at(nb_system *s, const double *x, double *a) {
    const int64_t n = s->n;
    ScopedTimer t(s->prof, "force", s->stream);
    if (s->precision == NB_PREC_MIXED) {
        s->dw.pos32.ensure((size_t)n);
        s->dw.ipl = s->tune_direct_ipl > 0 ? s->tune_direct_ipl : 2;
        s->dw.variant = s->tune_direct_kernel == 1 ? 0 : 1;
        s->dw.nsplit = choose_nsplit(n, s->dw.ipl, s->tune_direct_waves);
        s->dw.partial.ensure((size_t)s->dw.nsplit * 6 * (size_t)n);
        launch_pack(s->stream, n, x, s->m.p, s->G, s->dw.pos32.p);
        launch_force_mixed(s->stream, n, s->dw.pos32.p, (float)(s->eps * s->eps), form_of(s) == 1, s->dw.partial.p,
                           s->dw.nsplit, s->dw.ipl, s->dw.variant);
        launch_update_mixed(s->stream, n, s->dw.partial.p, s->dw.nsplit, a, s->v.p, s->x.p, s->m.p, s->G,
                            s->half_dt, s->dt, UPD_STORE, s->dw.pos32.p);
    } else if (s->precision == NB_PREC_FP64_FAST) {
        s->dw.nsplit64 = choose_nsplit64(n);
        s->dw.partial64.ensure((size_t)s->dw.nsplit64 * 3 * (size_t)n);
        launch_force_fp64_split(s->stream, n, x, s->m.p, s->eps * s->eps, form_of(s) == 1, s->dw.partial64.p,
                                s->dw.nsplit64);
        launch_update_fp64_split(s->stream, n, s->dw.partial64.p, s->dw.nsplit64, s->G, a, s->v.p, s->x.p,
                                 s->half_dt, s->dt, UPD_STORE);
    } else {
        s->dw.terms.ensure((size_t)(fp64_terms_block(n) * 3 * n));
        launch_force_fp64(s->stream, n, x, s->m.p, s->G, s->eps * s->eps, form_of(s), a, s->dw.terms.p);
    }
}

// Steps of evolve's other direct-sum integrators (gregsys.hpp:2666-2891 for
// Euler, modified Euler and midpoint, :2993-3039 for leapfrog DKD), each step
// followed by s_coll when collisions are on.  Every step evaluates its own
// forces, so no cached accelerations carry over between calls.
void direct_other_steps(nb_system *s, int method, uint64_t steps) {
    const bool coll = s->coll_mode == NB_OVERLAP_COLL_CHECK;
    for (uint64_t st = 1; st <= steps; ++st) {
        const int64_t n = s->n;
        if (n == 0) break;
        const size_t n3 = 3 * (size_t)n;
        if (method == NB_EULER) {
            forces_at(s, s->x.p, s->a.p);
            launch_euler(s->stream, n, s->a.p, s->v.p, s->x.p, s->dt);
        } else if (method == NB_MODIFIED_EULER || method == NB_MIDPOINT) {
            s->dw.px.ensure(n3);
            s->dw.pv.ensure(n3);
            s->dw.pa.ensure(n3);
            forces_at(s, s->x.p, s->a.p);
            const bool mid = method == NB_MIDPOINT;
            launch_predict(s->stream, n, s->x.p, s->v.p, s->a.p, mid ? s->half_dt : s->dt, s->dw.px.p, s->dw.pv.p);
            forces_at(s, s->dw.px.p, s->dw.pa.p);
            if (mid)
                launch_midpoint(s->stream, n, s->dw.pv.p, s->dw.pa.p, s->v.p, s->x.p, s->dt);
            else
                launch_modeuler(s->stream, n, s->a.p, s->dw.pv.p, s->dw.pa.p, s->v.p, s->x.p, s->half_dt);
        } else {  // NB_LEAPFROG_DKD: drift half, forces, kick full, drift half
            launch_drift(s->stream, n, s->v.p, s->x.p, s->half_dt);
            forces_at(s, s->x.p, s->a.p);
            launch_kick_drift(s->stream, n, s->a.p, s->v.p, s->x.p, s->dt, s->half_dt);
        }
        ++s->step_count;
        run_hook(s, s->step_count, NB_HOOK_STEP);
        if (coll) collision_pass(s, s->step_count);
        run_hook(s, s->step_count, NB_HOOK_BODIES);
        if (record_due(s, s->step_count)) record_state(s, s->step_count);
    }
}

// Everything a captured step graph depends on: buffer addresses, launch
// shapes and the scalars baked into its kernel arguments.
GraphKey graph_key(const nb_system *s) {
    GraphKey k;
    k.n = s->n;
    k.precision = s->precision;
    k.ptrs = {s->x.p, s->v.p, s->a.p, s->m.p, s->dw.pos32.p, s->dw.partial.p, s->dw.partial64.p, s->dw.terms.p};
    k.shape = {s->dw.nsplit, s->dw.ipl, s->dw.variant, s->dw.nsplit64, form_of(s)};
    k.scalars = {s->G, s->dt, s->half_dt, s->eps};
    return k;
}

// One direct KDK step's launches from valid accelerations.  `open`: the
// opening half kick and drift are still due; `last`: end with the closing
// half kick instead of fusing it with the next step's opening kick and drift.
// `partial` / `partial64` receive the split partials of the mixed / FP64-fast
// force pass.  Returns whether the next step must open.
bool kdk_step(nb_system *s, bool open, bool last, float *partial, double *partial64) {
    const bool mixed = s->precision == NB_PREC_MIXED;
    const bool fast64 = s->precision == NB_PREC_FP64_FAST;
    const int64_t n = s->n;
    if (open) {
        launch_kick_drift(s->stream, n, s->a.p, s->v.p, s->x.p, s->half_dt, s->dt);
        if (mixed) launch_pack(s->stream, n, s->x.p, s->m.p, s->G, s->dw.pos32.p);
    }
    if (fast64) {
        {
            ScopedTimer t(s->prof, "force", s->stream);
            launch_force_fp64_split(s->stream, n, s->x.p, s->m.p, s->eps * s->eps, form_of(s) == 1, partial64,
                                    s->dw.nsplit64);
        }
        launch_update_fp64_split(s->stream, n, partial64, s->dw.nsplit64, s->G, s->a.p, s->v.p, s->x.p, s->half_dt,
                                 s->dt, last ? UPD_CLOSE : UPD_CLOSE_OPEN);
        return last;
    }
    if (mixed) {
        {
            ScopedTimer t(s->prof, "force", s->stream);
            launch_force_mixed(s->stream, n, s->dw.pos32.p, (float)(s->eps * s->eps), form_of(s) == 1, partial,
                               s->dw.nsplit, s->dw.ipl, s->dw.variant);
        }
        ScopedTimer t(s->prof, "update", s->stream);
        launch_update_mixed(s->stream, n, partial, s->dw.nsplit, s->a.p, s->v.p, s->x.p, s->m.p, s->G, s->half_dt,
                            s->dt, last ? UPD_CLOSE : UPD_CLOSE_OPEN, s->dw.pos32.p);
        return last;
    }
    {
        ScopedTimer t(s->prof, "force", s->stream);
        launch_force_fp64(s->stream, n, s->x.p, s->m.p, s->G, s->eps * s->eps, form_of(s), s->a.p,
                          s->dw.terms.n >= (size_t)(fp64_terms_block(n) * 3 * n) ? s->dw.terms.p : nullptr);
    }
    launch_kick(s->stream, n, s->a.p, s->v.p, s->half_dt);
    return true;
}

void direct_kdk_steps(nb_system *s, uint64_t steps) {
    const bool mixed = s->precision == NB_PREC_MIXED;
    const bool fast64 = s->precision == NB_PREC_FP64_FAST;
    const bool coll = s->coll_mode == NB_OVERLAP_COLL_CHECK;
    auto issue_step = [&](bool open, bool last) -> bool {
        return kdk_step(s, open, last, s->dw.partial.p, s->dw.partial64.p);
    };
    // Interior steps of small systems replay a captured graph of kGraphSteps
    // steps (no collisions, no per-launch timing).
    const bool hooks = hooks_on(s);
    const bool graphs = !coll && !hooks && !s->prof.on && !s->rec_every &&
                        (s->tune_step_graphs == 1 || (s->tune_step_graphs == 0 && s->n <= kGraphMaxBodies));
    bool need_open = true;
    for (uint64_t st = 1; st <= steps;) {
        const int64_t n = s->n;
        if (n == 0) break;
        if (graphs && st > 1 && steps - st >= (uint64_t)kGraphSteps) {
            const GraphKey key = graph_key(s);
            auto chunk = [&]() {
                bool open = need_open;
                for (int q = 0; q < kGraphSteps; ++q) open = issue_step(open, false);
            };
            bool have = s->step_graph.valid() && s->step_graph_key == key;
            if (!have) {
                have = s->step_graph.capture(s->stream, chunk);
                s->step_graph_key = key;
            }
            if (have)
                s->step_graph.launch(s->stream);
            else
                chunk();  // capture invalidated from outside: the same launches, eagerly
            st += kGraphSteps;
            s->step_count += kGraphSteps;
            continue;
        }
        if (fast64 && s->dw.partial64.n < (size_t)s->dw.nsplit64 * 3 * (size_t)n) direct_forces(s);
        if (mixed && s->dw.pos32.n < (size_t)n) direct_forces(s);  // defensive: sizes follow n
        // a step that records, runs collisions or calls hooks ends with its closing kick
        need_open = issue_step(need_open, st == steps || coll || hooks || record_due(s, s->step_count + 1));
        ++st;
        ++s->step_count;
        run_hook(s, s->step_count, NB_HOOK_STEP);
        if (coll) {
            collision_pass(s, s->step_count);
            if (mixed && s->n != n && s->n > 0) {
                // body count changed: resize split partials for the new n
                s->dw.nsplit = choose_nsplit(s->n, s->dw.ipl, s->tune_direct_waves);
                s->dw.partial.ensure((size_t)s->dw.nsplit * 6 * (size_t)s->n);
            }
        }
        run_hook(s, s->step_count, NB_HOOK_BODIES);
        if (record_due(s, s->step_count)) record_state(s, s->step_count);
    }
}

}  // namespace nbe

extern "C" {

int nb_create(int device, nb_system **out) {
    if (!out) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_create: out is NULL");
    *out = nullptr;
    return guarded(nullptr, [&]() {
        int count = 0;
        NB_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count)
            throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_create: device index out of range");
        NB_HIP(hipSetDevice(device));
        nb_system *s = new nb_system();
        s->device = device;
        hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete s;
            throw DeviceError(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
        }
        try {
            resolve_scratch(device);
        } catch (...) {
            delete s;
            throw;
        }
        long double G;
        parse_units("M1:1,D1:1,T1:1", &G, s->G_vals);
        s->G = (double)G;
        *out = s;
        return NB_OK;
    });
}

int nb_destroy(nb_system *s) {
    if (!s) return NB_OK;
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    delete s;
    return NB_OK;
}

const char *nb_last_error(const nb_system *s) { return s ? s->err.c_str() : g_static_err.c_str(); }

int nb_set_bodies(nb_system *s, int64_t n, const double *m, const double *r, const double *pos, const double *vel,
                  const double *rest_c, int check_overlap) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_set_bodies: NULL system");
    return guarded(s, [&]() {
        if (n < 0 || (n > 0 && (!m || !pos || !vel)))
            throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_set_bodies: bad arguments");
        if (n > (int64_t)std::numeric_limits<int>::max() / 2)
            throw ApiError(NB_ERR_CAPACITY, "nb_set_bodies: too many bodies for one context");
        for (int64_t i = 0; i < n; ++i) {
            if (m[i] < 0) throw ApiError(NB_ERR_NEGATIVE_MASS, "body " + std::to_string(i) + " has negative mass");
            if (r && r[i] < 0)
                throw ApiError(NB_ERR_NEGATIVE_RADIUS, "body " + std::to_string(i) + " has negative radius");
        }
        if (s->in_hook) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_set_bodies cannot be called from a step hook");
        NB_HIP(hipStreamSynchronize(s->stream));
        bh_forget_order(s);
        s->n = n;
        size_t nn = (size_t)std::max<int64_t>(n, 1);
        s->m.ensure(nn); s->r.ensure(nn); s->e.ensure(nn); s->pe.ensure(nn);
        s->x.ensure(3 * nn); s->v.ensure(3 * nn); s->a.ensure(3 * nn);
        if (n > 0) {
            std::vector<double> tmp;
            copy_wait(s->m.p, m, (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
            tmp.assign((size_t)n, 1.0);
            copy_wait(s->r.p, r ? r : tmp.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
            copy_wait(s->e.p, rest_c ? rest_c : tmp.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
            to_soa(n, pos, tmp);
            copy_wait(s->x.p, tmp.data(), (size_t)(3 * n) * 8, hipMemcpyHostToDevice, s->stream);
            to_soa(n, vel, tmp);
            copy_wait(s->v.p, tmp.data(), (size_t)(3 * n) * 8, hipMemcpyHostToDevice, s->stream);
            zero_wait(s->a.p, (size_t)(3 * n) * 8, s->stream);
        }
        s->labels.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) s->labels[(size_t)i] = i;
        s->next_label = n;
        s->evolved = false;
        s->acc_valid = false;
        s->step_count = 0;
        s->merge_log.clear();
        s->rec_total = 0;  // records of the previous bodies are dropped
        s->rec_n0 = std::max<int64_t>(s->rec_n0, n);
        s->time_elapsed = 0;
        int64_t oi = 0, oj = 0;
        if (check_overlap && n > 1 && first_overlap(s, 0, &oi, &oj)) {
            // system::check_overlap (gregsys.hpp:307-342) without merging:
            // the constructor throws, so no bodies are kept
            s->n = 0;
            s->labels.clear();
            s->next_label = 0;
            throw ApiError(NB_ERR_OVERLAPPING_BODIES, "The bodies with id=" + std::to_string(oi) + " and id=" +
                                                          std::to_string(oj) +
                                                          " that were added to this system object overlap.");
        }
        return NB_OK;
    });
}

int nb_merge_overlapping(nb_system *s, int64_t *removed) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_merge_overlapping: NULL system");
    if (removed) *removed = 0;
    return guarded(s, [&]() {
        if (s->in_hook) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_merge_overlapping cannot be called from a step hook");
        const int64_t k = merge_overlapping(s);
        if (removed) *removed = k;
        return NB_OK;
    });
}

int nb_update_state(nb_system *s, int64_t n, const double *m, const double *r, const double *pos, const double *vel,
                    const double *acc, const double *rest_c) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_update_state: NULL system");
    return guarded(s, [&]() {
        if (n != s->n) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_update_state: n differs from the body count");
        for (int64_t i = 0; i < n; ++i) {
            if (m && m[i] < 0) throw ApiError(NB_ERR_NEGATIVE_MASS, "body " + std::to_string(i) + " has negative mass");
            if (r && r[i] < 0)
                throw ApiError(NB_ERR_NEGATIVE_RADIUS, "body " + std::to_string(i) + " has negative radius");
        }
        NB_HIP(hipStreamSynchronize(s->stream));
        if (n == 0) return NB_OK;
        if (m) copy_wait(s->m.p, m, (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
        if (r) copy_wait(s->r.p, r, (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
        if (rest_c) copy_wait(s->e.p, rest_c, (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
        std::vector<double> tmp;
        const double *srcs[3] = {pos, vel, acc};
        double *dsts[3] = {s->x.p, s->v.p, s->a.p};
        for (int q = 0; q < 3; ++q) {
            if (!srcs[q]) continue;
            to_soa(n, srcs[q], tmp);
            copy_wait(dsts[q], tmp.data(), (size_t)(3 * n) * 8, hipMemcpyHostToDevice, s->stream);
        }
        return NB_OK;
    });
}

int nb_set_step_hook(nb_system *s, nb_step_hook fn, void *user, int where_mask) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_set_step_hook: NULL system");
    if (where_mask & ~(NB_HOOK_STEP | NB_HOOK_BODIES))
        return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_set_step_hook: unknown bits in where_mask");
    if (s->in_hook) return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_set_step_hook cannot be called from a step hook");
    s->hook_fn = fn;
    s->hook_user = fn ? user : nullptr;
    s->hook_mask = fn ? where_mask : 0;
    return NB_OK;
}

int nb_get_state(const nb_system *cs, int64_t cap, double *m, double *r, double *pos, double *vel, double *acc,
                 double *rest_c) {
    nb_system *s = const_cast<nb_system *>(cs);
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_get_state: NULL system");
    return guarded(s, [&]() {
        const int64_t n = s->n;
        if (cap < n) throw ApiError(NB_ERR_CAPACITY, "nb_get_state: capacity smaller than the body count");
        NB_HIP(hipStreamSynchronize(s->stream));
        if (n == 0) return NB_OK;
        if (m) copy_wait(m, s->m.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
        if (r) copy_wait(r, s->r.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
        if (rest_c) copy_wait(rest_c, s->e.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
        std::vector<double> tmp((size_t)(3 * n));
        const double *srcs[3] = {s->x.p, s->v.p, s->a.p};
        double *dsts[3] = {pos, vel, acc};
        for (int q = 0; q < 3; ++q) {
            if (!dsts[q]) continue;
            copy_wait(tmp.data(), srcs[q], (size_t)(3 * n) * 8, hipMemcpyDeviceToHost, s->stream);
            from_soa(n, tmp, dsts[q]);
        }
        return NB_OK;
    });
}

int nb_num_bodies(const nb_system *s, int64_t *n) {
    if (!s || !n) return fail(const_cast<nb_system *>(s), NB_ERR_INVALID_ARGUMENT, "nb_num_bodies: NULL argument");
    *n = s->n;
    return NB_OK;
}

int nb_body_ids(const nb_system *cs, int64_t cap, int64_t *ids) {
    nb_system *s = const_cast<nb_system *>(cs);
    if (!s || !ids) return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_body_ids: NULL argument");
    if (cap < s->n) return fail(s, NB_ERR_CAPACITY, "nb_body_ids: capacity smaller than the body count");
    std::copy(s->labels.begin(), s->labels.end(), ids);
    return NB_OK;
}

int nb_set_units(nb_system *s, const char *units) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    long double G;
    uint64_t vals[6];
    if (!parse_units(units, &G, vals))
        return fail(s, NB_ERR_INVALID_ARGUMENT,
                    std::string("The units_format string passed, \"") + (units ? units : "(null)") +
                        "\", does not match the format required:\n\"Ma:b,Dc:x,Ty:z\"");
    s->G = (double)G;
    std::copy(vals, vals + 6, s->G_vals);
    s->acc_valid = false;
    return NB_OK;
}

int nb_set_G(nb_system *s, double G) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (!(G > 0) || !std::isfinite(G)) return fail(s, NB_ERR_INVALID_ARGUMENT, "G must be positive and finite");
    s->G = G;
    s->acc_valid = false;
    return NB_OK;
}

int nb_get_G(const nb_system *s, double *G) {
    if (!s || !G) return fail(const_cast<nb_system *>(s), NB_ERR_INVALID_ARGUMENT, "NULL argument");
    *G = s->G;
    return NB_OK;
}

int nb_set_timestep(nb_system *s, double dt) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (!(dt > 0)) return fail(s, NB_ERR_INVALID_ARGUMENT, "Time-step cannot be zero or negative.");
    s->dt = dt;
    s->half_dt = dt / 2;
    return NB_OK;
}

int nb_set_iterations(nb_system *s, uint64_t it) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (!it) return fail(s, NB_ERR_INVALID_ARGUMENT, "cannot have zero iterations");
    s->iterations = it;
    return NB_OK;
}

int nb_set_softening(nb_system *s, double eps) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (eps != 0.0 && !(eps >= std::sqrt(std::numeric_limits<double>::min())))
        return fail(s, NB_ERR_INVALID_ARGUMENT, "softening below sqrt(DBL_MIN)");
    if (!std::isfinite(eps)) return fail(s, NB_ERR_INVALID_ARGUMENT, "softening must be finite");
    s->eps = eps;
    s->acc_valid = false;
    return NB_OK;
}

int nb_set_bh_opening_angle(nb_system *s, double theta) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (theta < 0 || theta >= 3.14159265358979323846)
        return fail(s, NB_ERR_INVALID_ARGUMENT, "opening angle must lie in [0, pi)");
    s->theta = theta;
    s->acc_valid = false;
    return NB_OK;
}

int nb_set_collisions(nb_system *s, int mode, int mergers, double min_mom) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (mode != NB_NO_COLL_CHECK && mode != NB_OVERLAP_COLL_CHECK)
        return fail(s, NB_ERR_NOT_SUPPORTED, "only collisions = 0 or 3 (overlap check) are supported");
    if (min_mom < 0) return fail(s, NB_ERR_INVALID_ARGUMENT, "min_tot_com_mom cannot be negative");
    s->coll_mode = mode;
    s->mergers = mergers != 0;
    s->min_tot_com_mom = min_mom;
    return NB_OK;
}

int nb_set_precision(nb_system *s, int p) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (p != NB_PREC_FP64 && p != NB_PREC_MIXED && p != NB_PREC_FP64_FAST)
        return fail(s, NB_ERR_INVALID_ARGUMENT, "unknown precision");
    s->precision = p;
    s->acc_valid = false;
    return NB_OK;
}

int nb_evolve(nb_system *s, int method, int64_t *elapsed_ns) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (elapsed_ns) *elapsed_ns = 0;
    return guarded(s, [&]() {
        if (s->in_hook) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_evolve cannot be called from a step hook");
        if (!s->n || !check_option(method)) return NB_OK;  // evolve returns 0 ns (:2611-2612)
        auto t0 = std::chrono::high_resolution_clock::now();
        if ((method & NB_TWO_BODY) == NB_TWO_BODY) {
            if (s->n != 2) throw ApiError(NB_ERR_TWO_BODY, "A two-body integration can only be performed if there are two bodies present.");
            throw ApiError(NB_ERR_NOT_SUPPORTED, "two_body integration is empty in the reference");
        }
        // the reference's dispatch order (evolve :2626-3055): the first matching flag wins
        const int direct_other = (method & NB_EULER) == NB_EULER                   ? NB_EULER
                                 : (method & NB_MODIFIED_EULER) == NB_MODIFIED_EULER ? NB_MODIFIED_EULER
                                 : (method & NB_MIDPOINT) == NB_MIDPOINT             ? NB_MIDPOINT
                                 : (method & NB_LEAPFROG_KDK) == NB_LEAPFROG_KDK     ? 0
                                 : (method & NB_LEAPFROG_DKD) == NB_LEAPFROG_DKD     ? NB_LEAPFROG_DKD
                                                                                     : -1;
        if (direct_other < 0)
            throw ApiError(NB_ERR_NOT_SUPPORTED, "rk4 / rk3_8 are empty in the reference");
        if (direct_other > 0) {
            if ((method & NB_BARNES_HUT) == NB_BARNES_HUT)
                throw ApiError(NB_ERR_NOT_SUPPORTED,
                               "Barnes-Hut is built with leapfrog_kdk only (the reference's other BH branches are "
                               "empty or do not terminate)");
            s->step_count = 0;
            if (s->rec_every) record_state(s, 0);
            direct_other_steps(s, direct_other, s->iterations);
            s->acc_method = direct_other;
            s->acc_valid = false;
            NB_HIP(hipStreamSynchronize(s->stream));
            s->prof.collect();
            s->evolved = true;
            s->time_elapsed = (long double)s->iterations * (long double)s->dt;  // evolve :3069
            auto t1 = std::chrono::high_resolution_clock::now();
            if (elapsed_ns)
                *elapsed_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
            return NB_OK;
        }
        s->step_count = 0;
        if ((method & NB_BARNES_HUT) == NB_BARNES_HUT) {
            // with collisions each step ends with bh_s_coll (bounces only, also
            // in the mergers build: gregsys.hpp:2937-2939)
            if (s->rec_every) record_state(s, 0);
            bh_forces(s, nullptr);
            launch_kick(s->stream, s->n, s->a.p, s->v.p, s->half_dt);  // initial half kick (:2910-2913)
            bh_steps(s, s->iterations);
            s->acc_method = NB_BARNES_HUT;
        } else {
            direct_forces(s);
            if (s->rec_every) record_state(s, 0);
            direct_kdk_steps(s, s->iterations);
            s->acc_method = NB_LEAPFROG_KDK;
        }
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.collect();
        s->evolved = true;
        s->time_elapsed = (long double)s->iterations * (long double)s->dt;  // evolve :3069
        auto t1 = std::chrono::high_resolution_clock::now();
        if (elapsed_ns) *elapsed_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        return NB_OK;
    });
}

int nb_advance(nb_system *s, uint64_t steps, int64_t *elapsed_ns) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (elapsed_ns) *elapsed_ns = 0;
    return guarded(s, [&]() {
        if (s->in_hook) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_advance cannot be called from a step hook");
        const bool other = s->evolved && (s->acc_method == NB_EULER || s->acc_method == NB_MODIFIED_EULER ||
                                           s->acc_method == NB_MIDPOINT || s->acc_method == NB_LEAPFROG_DKD);
        if (!s->evolved || (!other && !s->acc_valid))
            throw ApiError(NB_ERR_NO_EVOLUTION, "nb_advance needs a previous nb_evolve with unchanged parameters");
        auto t0 = std::chrono::high_resolution_clock::now();
        if (other) direct_other_steps(s, s->acc_method, steps);
        else if (s->acc_method == NB_BARNES_HUT) bh_steps(s, steps);
        else direct_kdk_steps(s, steps);
        s->time_elapsed += (long double)steps * (long double)s->dt;
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.collect();
        auto t1 = std::chrono::high_resolution_clock::now();
        if (elapsed_ns) *elapsed_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        return NB_OK;
    });
}

int nb_advance_noalloc(nb_system *s, uint64_t steps, int64_t *elapsed_ns) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (elapsed_ns) *elapsed_ns = 0;
    return guarded(s, [&]() {
        if (s->in_hook) throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_advance_noalloc cannot be called from a step hook");
        if (!s->evolved || !s->acc_valid || s->acc_method != NB_LEAPFROG_KDK)
            throw ApiError(NB_ERR_NO_EVOLUTION,
                           "nb_advance_noalloc needs a previous direct-sum NB_LEAPFROG_KDK nb_evolve with unchanged "
                           "parameters");
        const bool mixed = s->precision == NB_PREC_MIXED;
        if (!mixed && s->precision != NB_PREC_FP64_FAST)
            throw ApiError(NB_ERR_NOT_SUPPORTED, "nb_advance_noalloc: precision must be NB_PREC_MIXED or "
                                                 "NB_PREC_FP64_FAST");
        if (s->coll_mode != NB_NO_COLL_CHECK || s->rec_every)
            throw ApiError(NB_ERR_NOT_SUPPORTED, "nb_advance_noalloc: collisions and recording must be off");
        const int64_t n = s->n;
        const size_t need = mixed ? (size_t)s->dw.nsplit * 6 * (size_t)n * sizeof(float)
                                  : (size_t)s->dw.nsplit64 * 3 * (size_t)n * sizeof(double);
        if (need > (size_t)NB_NOALLOC_SCRATCH_BYTES)
            throw ApiError(NB_ERR_CAPACITY, "nb_advance_noalloc: the split partials need " + std::to_string(need) +
                                                " bytes, more than the " +
                                                std::to_string((size_t)NB_NOALLOC_SCRATCH_BYTES) +
                                                "-byte static scratch (use nb_advance)");
        if (mixed && s->dw.pos32.n < (size_t)n)
            throw ApiError(NB_ERR_NO_EVOLUTION, "nb_advance_noalloc: no packed positions for this body count");
        void *p = scratch_ptr(s->device);
        if (!p) throw DeviceError("nb_advance_noalloc: device scratch address not resolved");
        auto t0 = std::chrono::high_resolution_clock::now();
        bool open = true;
        for (uint64_t st = 1; st <= steps && n > 0; ++st)
            open = kdk_step(s, open, st == steps, mixed ? (float *)p : nullptr, mixed ? nullptr : (double *)p);
        s->step_count += steps;
        s->time_elapsed += (long double)steps * (long double)s->dt;
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.collect();
        auto t1 = std::chrono::high_resolution_clock::now();
        if (elapsed_ns) *elapsed_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        return NB_OK;
    });
}

int nb_energy(nb_system *s, double *ke, double *pe) {
    if (!s || !ke || !pe) return fail(s, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return guarded(s, [&]() {
        const int64_t n = s->n;
        *ke = 0;
        *pe = 0;
        if (n == 0) return NB_OK;
        {
            ScopedTimer t(s->prof, "energy", s->stream);
            launch_pe_fp64(s->stream, n, s->x.p, s->m.p, s->G, s->eps * s->eps, form_of(s), s->pe.p);
        }
        std::vector<double> hpe((size_t)n), hm((size_t)n), hv((size_t)(3 * n));
        NB_HIP(hipMemcpyAsync(hpe.data(), s->pe.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream));
        NB_HIP(hipMemcpyAsync(hm.data(), s->m.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream));
        NB_HIP(hipMemcpyAsync(hv.data(), s->v.p, (size_t)(3 * n) * 8, hipMemcpyDeviceToHost, s->stream));
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.collect();
        // body-order totals (calc_energy :620-625); KE = 0.5*m*|v|*|v| (set_ke)
        volatile double tke = 0, tpe = 0;
        for (int64_t i = 0; i < n; ++i) {
            double vx = hv[(size_t)i], vy = hv[(size_t)(n + i)], vz = hv[(size_t)(2 * n + i)];
            double vm2 = vx * vx;
            vm2 = vm2 + vy * vy;
            vm2 = vm2 + vz * vz;
            double vm = std::sqrt(vm2);
            double k = 0.5 * hm[(size_t)i];
            k = k * vm;
            k = k * vm;
            tpe = tpe + hpe[(size_t)i];
            tke = tke + k;
        }
        *ke = tke;
        *pe = tpe;
        return NB_OK;
    });
}

int nb_set_recording(nb_system *s, uint64_t every, uint64_t capacity, int what) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (every && !capacity) return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_set_recording: capacity must be positive");
    if (every && (what & ~(NB_RECORD_STATE | NB_RECORD_ENERGY) || !what))
        return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_set_recording: what is NB_RECORD_STATE and/or NB_RECORD_ENERGY");
    return guarded(s, [&]() {
        s->rec_what = every ? what : 0;
        s->rec_every = every;
        s->rec_cap = every ? capacity : 0;
        s->rec_total = 0;
        s->rec_n0 = s->n;
        s->rec_step.clear();
        s->rec_n.clear();
        s->rec_ids.clear();
        if (!every) {
            s->rec_x.release();
            s->rec_v.release();
            s->rec_e.release();
        }
        return NB_OK;
    });
}

int nb_record_count(const nb_system *s, uint64_t *taken, uint64_t *stored) {
    if (!s || !taken || !stored) return fail(const_cast<nb_system *>(s), NB_ERR_INVALID_ARGUMENT, "NULL argument");
    *taken = s->rec_total;
    *stored = std::min<uint64_t>(s->rec_total, s->rec_cap);
    return NB_OK;
}

int nb_get_record(nb_system *s, uint64_t k, int64_t *step, int64_t *n, double *pos, double *vel, int64_t *ids,
                  double *ke, double *pe) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    const uint64_t stored = std::min<uint64_t>(s->rec_total, s->rec_cap);
    if (k >= stored) return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_get_record: no such record");
    return guarded(s, [&]() {
        const uint64_t slot = (s->rec_total - stored + k) % s->rec_cap;  // k = 0: the oldest kept
        const int64_t nb = s->rec_n[slot];
        if (step) *step = s->rec_step[slot];
        if (n) *n = nb;
        const size_t stride = (size_t)3 * (size_t)s->rec_n0;
        double e2[2] = {0, 0};
        NB_HIP(hipMemcpyAsync(e2, s->rec_e.p + 2 * slot, 2 * 8, hipMemcpyDeviceToHost, s->stream));
        NB_HIP(hipStreamSynchronize(s->stream));
        if (ke) *ke = e2[0];
        if (pe) *pe = e2[1];
        std::vector<double> buf((size_t)3 * (size_t)nb);
        auto get = [&](const DBuf<double> &src, double *dst) {
            if (!dst || nb == 0) return;
            if (!(s->rec_what & NB_RECORD_STATE) || !src.p) {  // an energy-only ring keeps no state
                std::fill(dst, dst + 3 * nb, 0.0);
                return;
            }
            copy_wait(buf.data(), src.p + slot * stride, buf.size() * 8, hipMemcpyDeviceToHost, s->stream);
            for (int64_t i = 0; i < nb; ++i)
                for (int c = 0; c < 3; ++c) dst[3 * i + c] = buf[(size_t)c * nb + i];
        };
        get(s->rec_x, pos);
        get(s->rec_v, vel);
        if (ids) std::copy(s->rec_ids[slot].begin(), s->rec_ids[slot].end(), ids);
        return NB_OK;
    });
}

int nb_track(nb_system *s, const int64_t *idx, int64_t k, const double *point, double *mass, double *com,
             double *com_vel, double *mean_dist, double *mean_sep) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    if (k < 0 || (k > 0 && !idx)) return fail(s, NB_ERR_INVALID_ARGUMENT, "nb_track: bad body list");
    return guarded(s, [&]() {
        if (k == 0 && mean_sep)  // mean_sep() on an empty tracker throws nbody_error (:3564-3565)
            throw ApiError(NB_ERR_INVALID_ARGUMENT, "Error: no bodies being tracked.\n");
        double out[9];
        track_bodies(s, idx, k, point, out);
        if (mass) *mass = out[0];
        for (int c = 0; c < 3; ++c) {
            if (com) com[c] = out[1 + c];
            if (com_vel) com_vel[c] = out[4 + c];
        }
        if (mean_dist) *mean_dist = out[7];
        if (mean_sep) *mean_sep = out[8];
        return NB_OK;
    });
}

int nb_merge_log(const nb_system *cs, int64_t cap, int64_t *records, int64_t *count) {
    nb_system *s = const_cast<nb_system *>(cs);
    if (!s || !count) return fail(s, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    int64_t nrec = (int64_t)(s->merge_log.size() / 4);
    *count = nrec;
    if (records) {
        if (cap < nrec) return fail(s, NB_ERR_CAPACITY, "nb_merge_log: capacity too small");
        std::copy(s->merge_log.begin(), s->merge_log.end(), records);
    }
    return NB_OK;
}

int nb_tree_topology(nb_system *s, int64_t cap, int32_t *depth, int32_t *octant, int64_t *body, int64_t *parent,
                     double *mass, double *com, double *side, int64_t *count) {
    if (!s || !count) return fail(s, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return guarded(s, [&]() { return bh_topology(s, cap, depth, octant, body, parent, mass, com, side, count); });
}

int nb_bh_accelerations(nb_system *s, double *acc, int64_t *interactions) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    return guarded(s, [&]() {
        if (s->n == 0) return NB_OK;
        bh_forces(s, interactions);
        bh_restore_order(s);   // accelerations back in input order before they are read
        s->acc_valid = false;  // cached a no longer matches an evolve in progress
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.collect();
        if (acc) {
            std::vector<double> tmp((size_t)(3 * s->n));
            copy_wait(tmp.data(), s->a.p, tmp.size() * 8, hipMemcpyDeviceToHost, s->stream);
            from_soa(s->n, tmp, acc);
        }
        return NB_OK;
    });
}

int nb_adjoint(nb_system *s, uint64_t steps, uint64_t ckpt, const double *gxT, const double *gvT, double *gx0,
               double *gv0, double *gm, int64_t *fwd_ns, int64_t *bwd_ns) {
    if (!s || !gxT || !gvT || !gx0 || !gv0 || !gm) return fail(s, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return guarded(s, [&]() {
        if (s->n == 0) throw ApiError(NB_ERR_EMPTY_SYSTEM, "This operation is not permitted on empty system objects (no bodies).");
        if (steps == 0 || ckpt == 0) throw ApiError(NB_ERR_INVALID_ARGUMENT, "steps and ckpt_every must be positive");
        if (s->coll_mode != NB_NO_COLL_CHECK)
            throw ApiError(NB_ERR_NOT_SUPPORTED, "the adjoint covers collision-free KDK only");
        adjoint_run(s, steps, ckpt, gxT, gvT, gx0, gv0, gm, fwd_ns, bwd_ns);
        return NB_OK;
    });
}

int nb_profile_enable(nb_system *s, int on) {
    if (!s) return fail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL system");
    return guarded(s, [&]() {
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.clear();
        s->prof.on = on != 0;
        return NB_OK;
    });
}

int nb_profile_read(nb_system *s, const char *name, double *total_ms, int64_t *launches) {
    if (!s || !name || !total_ms || !launches) return fail(s, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return guarded(s, [&]() {
        NB_HIP(hipStreamSynchronize(s->stream));
        s->prof.collect();
        auto it = s->prof.recs.find(name);
        *total_ms = it == s->prof.recs.end() ? 0.0 : it->second.total_ms;
        *launches = it == s->prof.recs.end() ? 0 : it->second.launches;
        return NB_OK;
    });
}

}  // extern "C"
