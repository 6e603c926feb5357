// nb_coll.hip -- body-body collision detect / bounce / merge (gfx950), for one
// system and for batched ensembles of independent small systems.
//
// Reference path: system::s_coll (gregsys.hpp:690-914) with and without
// -DGREGSYS_MERGERS, the merge constructor and operator+= of gtd::body
// (gregbod.hpp:933-940, 1065-1082, 1094-1150), and the KDK driver that calls
// s_coll after each step's closing kick (FUNC_TEMPL_SELECT :16-57).
//
// s_coll is serial and order dependent (ordered map keyed by distance with
// equal keys dropped, axis momenta captured at scan time, one max_mom per call,
// a full rescan after every merge), so it is replayed exactly, by one
// workgroup per system:
//  * every pair is first tested in parallel, with no barrier between outer
//    bodies (the reference's own FP64 test r_i + r_j > |p_j - p_i|); outer
//    bodies with no overlapping j > i cannot change anything and are skipped
//    (positions do not move during the pass until a merge, after which every
//    later outer body is scanned); the workgroup finds the next flagged
//    outer body together, 8 flags per thread;
//  * a scanned outer body's j range is split over the threads; the merge
//    candidate is the LAST j reaching the largest curr_mom (the sequential
//    "curr >= max_mom" record rule), found by a block reduction; a merge is
//    computed one body field per thread;
//  * bounce entries are sorted by (distance, j) -- equal distances keep the
//    smallest j, as the map's emplace does -- and applied in order by one
//    thread, since each entry updates the outer body's velocity.
// The batched kernel runs the pass only on steps where its force loop's
// overlap prefilter flags a body (nb_coll_dev.hpp).
// Arithmetic is FP64 with contraction off; the merged radius uses a fixed
// Newton cube root (bit-reproducible, within 1 ulp of cbrt).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "nb_system.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {

#include "nb_coll_dev.hpp"

inline BodyView view_of(nb_system *s, int64_t *lab) {
    const int64_t n = s->n;
    return BodyView{s->x.p, s->x.p + n, s->x.p + 2 * n, s->v.p, s->v.p + n, s->v.p + 2 * n,
                    s->a.p, s->a.p + n, s->a.p + 2 * n, s->m.p, s->r.p, s->e.p, lab};
}

}  // namespace

struct CollWork {
    DBuf<int64_t> lab;
    DBuf<unsigned char> flags;
    DBuf<int64_t> log, loglen, next;
    DBuf<int> n, err;
    int64_t log_cap = 0;
    int64_t synced_log = 0;
};

void coll_work_free(CollWork *w) { delete w; }

// s_coll after a step's closing kick.  Labels live on the device while
// collisions are on; the body count and new merge records come back here.
void collision_pass(nb_system *s, uint64_t step) {
    if (!s->coll) s->coll = new CollWork();
    CollWork &w = *s->coll;
    const int64_t n = s->n;
    hipStream_t st = s->stream;
    // the pass addresses bodies by 32-bit byte offsets (el() in nb_coll_dev.hpp)
    if (n >= ((int64_t)1 << 28)) throw ApiError(NB_ERR_CAPACITY, "collision pass: at most 2^28 bodies");
    // labels and the next label travel with every pass (the host copies are authoritative)
    w.lab.ensure((size_t)std::max<int64_t>(n, 1));
    w.loglen.ensure(1);
    w.next.ensure(1);
    if (n) NB_HIP(hipMemcpyAsync(w.lab.p, s->labels.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    NB_HIP(hipMemcpyAsync(w.next.p, &s->next_label, 8, hipMemcpyHostToDevice, st));
    w.flags.ensure((size_t)std::max<int64_t>(n, 1));
    w.log_cap = std::max<int64_t>(n, 1);
    w.log.ensure(4 * (size_t)w.log_cap);
    w.n.ensure(1);
    w.err.ensure(1);
    int nn = (int)n;
    int zero_i = 0;
    int64_t zero = 0;
    NB_HIP(hipMemcpyAsync(w.n.p, &nn, sizeof(int), hipMemcpyHostToDevice, st));
    NB_HIP(hipMemcpyAsync(w.err.p, &zero_i, sizeof(int), hipMemcpyHostToDevice, st));
    NB_HIP(hipMemcpyAsync(w.loglen.p, &zero, 8, hipMemcpyHostToDevice, st));
    BodyView B = view_of(s, w.lab.p);
    {
        ScopedTimer tm(s->prof, "collide", st);
        hipLaunchKernelGGL(k_coll_single, dim3(1), dim3(kThreads), 0, st, B, w.n.p, s->mergers ? 1 : 0,
                           s->min_tot_com_mom, (int64_t)step, w.next.p, w.flags.p, w.log.p, w.log_cap, w.loglen.p,
                           w.err.p);
    }
    NB_HIP(hipGetLastError());
    int n_out = 0, errv = 0;
    int64_t nlog = 0, next = 0;
    NB_HIP(hipMemcpyAsync(&n_out, w.n.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(&errv, w.err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(&nlog, w.loglen.p, 8, hipMemcpyDeviceToHost, st));
    NB_HIP(hipMemcpyAsync(&next, w.next.p, 8, hipMemcpyDeviceToHost, st));
    NB_HIP(hipStreamSynchronize(st));
    if (errv) throw ApiError(NB_ERR_CAPACITY, "collision pass: more than 256 overlapping partners for one body");
    if (nlog > 0) {
        std::vector<int64_t> rec(4 * (size_t)nlog);
        copy_wait(rec.data(), w.log.p, rec.size() * 8, hipMemcpyDeviceToHost, st);
        s->merge_log.insert(s->merge_log.end(), rec.begin(), rec.end());
        w.synced_log = (int64_t)(s->merge_log.size() / 4);
    }
    s->next_label = next;
    if (n_out != n) {
        // the SoA blocks are laid out with stride n: repack to the new count
        const int64_t m2 = n_out;
        std::vector<double> buf;
        for (DBuf<double> *d : {&s->x, &s->v, &s->a}) {
            buf.resize(3 * (size_t)n);
            copy_wait(buf.data(), d->p, buf.size() * 8, hipMemcpyDeviceToHost, st);
            std::vector<double> packed(3 * (size_t)m2);
            for (int c = 0; c < 3; ++c)
                std::copy(buf.begin() + (size_t)c * n, buf.begin() + (size_t)c * n + m2,
                          packed.begin() + (size_t)c * m2);
            copy_wait(d->p, packed.data(), packed.size() * 8, hipMemcpyHostToDevice, st);
        }
        s->n = m2;
    }
    s->labels.resize((size_t)s->n);
    if (s->n) copy_wait(s->labels.data(), w.lab.p, (size_t)s->n * 8, hipMemcpyDeviceToHost, st);
}

}  // namespace nbe

// ---------------- batched ensembles (C ABI) ----------------
using namespace nbe;

struct nb_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int64_t nsys = 0, total = 0;
    std::vector<int64_t> offsets;
    std::vector<int> order;
    DBuf<double> x, v, a, m, r, e;
    DBuf<int64_t> lab, off, next, log, loglen;
    DBuf<int> cnt, ord, errf;
    int max_n = 0;  // largest system at creation (merges only shrink systems)
    BatchParams P{6.6743e-11, 1.0, 0.5, 0.0, 3, 1, 1e10, 1};
    int64_t step = 0;
    // HIP events around the stepping launches of the last nb_batch_evolve
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double kernel_ms = 0.0;
    int64_t kernel_launches = 0;
    ~nb_batch() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
thread_local std::string g_batch_err;

// k_batch_steps instance for the batch: precision, and 64-body blocks per
// wave for its largest system (3 covers 768 bodies, 4 the 1024 maximum)
void launch_batch_steps(nb_batch *b, const BodyView &all, const BatchParams &P, int64_t step0, int64_t steps) {
    const dim3 grid((unsigned)b->nsys), block(kThreads);
    const bool small = b->max_n <= batch_capacity<3>();
    if (P.mixed) {
        if (small)
            hipLaunchKernelGGL((k_batch_steps<true, 3>), grid, block, 0, b->stream, all, b->ord.p, b->off.p, b->cnt.p,
                               b->next.p, b->log.p, b->loglen.p, b->errf.p, P, step0, steps);
        else
            hipLaunchKernelGGL((k_batch_steps<true, 4>), grid, block, 0, b->stream, all, b->ord.p, b->off.p, b->cnt.p,
                               b->next.p, b->log.p, b->loglen.p, b->errf.p, P, step0, steps);
    } else {
        if (small)
            hipLaunchKernelGGL((k_batch_steps<false, 3>), grid, block, 0, b->stream, all, b->ord.p, b->off.p,
                               b->cnt.p, b->next.p, b->log.p, b->loglen.p, b->errf.p, P, step0, steps);
        else
            hipLaunchKernelGGL((k_batch_steps<false, 4>), grid, block, 0, b->stream, all, b->ord.p, b->off.p,
                               b->cnt.p, b->next.p, b->log.p, b->loglen.p, b->errf.p, P, step0, steps);
    }
    NB_HIP(hipGetLastError());
}

int bfail(nb_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    else g_batch_err = msg;
    return code;
}

template <typename F>
int bguard(nb_batch *b, F &&f) {
    try {
        if (b) NB_HIP(hipSetDevice(b->device));
        return f();
    } catch (const ApiError &e) {
        return bfail(b, e.code, e.what);
    } catch (const DeviceError &e) {
        return bfail(b, NB_ERR_DEVICE, e.what);
    } catch (const std::bad_alloc &) {
        return bfail(b, NB_ERR_CAPACITY, "host allocation failed");
    }
}
}  // namespace

extern "C" {

int nb_batch_create(int device, int64_t nsys, const int64_t *offsets, const double *m, const double *r,
                    const double *pos, const double *vel, const double *rest_c, nb_batch **out) {
    if (!out) return bfail(nullptr, NB_ERR_INVALID_ARGUMENT, "nb_batch_create: out is NULL");
    *out = nullptr;
    return bguard(nullptr, [&]() {
        if (nsys <= 0 || !offsets || !m || !pos || !vel)
            throw ApiError(NB_ERR_INVALID_ARGUMENT, "nb_batch_create: bad arguments");
        int count = 0;
        NB_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw ApiError(NB_ERR_INVALID_ARGUMENT, "device index out of range");
        NB_HIP(hipSetDevice(device));
        const int64_t total = offsets[nsys];
        if (offsets[0] != 0 || total < 0) throw ApiError(NB_ERR_INVALID_ARGUMENT, "offsets must start at 0");
        for (int64_t s = 0; s < nsys; ++s) {
            int64_t k = offsets[s + 1] - offsets[s];
            if (k < 0) throw ApiError(NB_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
            if (k > kBatchMaxN) throw ApiError(NB_ERR_CAPACITY, "a batched system holds at most 1024 bodies");
        }
        for (int64_t i = 0; i < total; ++i) {
            if (m[i] < 0) throw ApiError(NB_ERR_NEGATIVE_MASS, "negative mass");
            if (r && r[i] < 0) throw ApiError(NB_ERR_NEGATIVE_RADIUS, "negative radius");
        }
        nb_batch *b = new nb_batch();
        b->device = device;
        hipError_t he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (he != hipSuccess) {
            delete b;
            throw DeviceError(std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(he));
        }
        b->nsys = nsys;
        b->total = total;
        b->offsets.assign(offsets, offsets + nsys + 1);
        const size_t T = (size_t)std::max<int64_t>(total, 1);
        b->x.ensure(3 * T); b->v.ensure(3 * T); b->a.ensure(3 * T);
        b->m.ensure(T); b->r.ensure(T); b->e.ensure(T); b->lab.ensure(T); b->log.ensure(4 * T);
        b->off.ensure((size_t)nsys + 1); b->next.ensure((size_t)nsys); b->loglen.ensure((size_t)nsys);
        b->cnt.ensure((size_t)nsys); b->ord.ensure((size_t)nsys); b->errf.ensure(1);
        std::vector<double> tmp(3 * T, 0.0);
        auto up3 = [&](DBuf<double> &d, const double *src) {
            for (int64_t i = 0; i < total; ++i)
                for (int c = 0; c < 3; ++c) tmp[(size_t)c * T + (size_t)i] = src ? src[3 * i + c] : 0.0;
            copy_wait(d.p, tmp.data(), 3 * T * 8, hipMemcpyHostToDevice, b->stream);
        };
        up3(b->x, pos);
        up3(b->v, vel);
        up3(b->a, nullptr);
        std::vector<double> one(T, 1.0);
        copy_wait(b->m.p, m, (size_t)total * 8, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->r.p, r ? r : one.data(), (size_t)total * 8, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->e.p, rest_c ? rest_c : one.data(), (size_t)total * 8, hipMemcpyHostToDevice, b->stream);
        std::vector<int64_t> lab(T), next((size_t)nsys), zeros((size_t)nsys, 0);
        std::vector<int> cnt((size_t)nsys);
        for (int64_t s = 0; s < nsys; ++s) {
            int64_t k = offsets[s + 1] - offsets[s];
            cnt[(size_t)s] = (int)k;
            next[(size_t)s] = k;
            for (int64_t q = 0; q < k; ++q) lab[(size_t)(offsets[s] + q)] = q;
        }
        for (int c : cnt) b->max_n = std::max(b->max_n, c);
        // longest systems first (n^2 work): the dispatcher hands out workgroups in order
        b->order.resize((size_t)nsys);
        std::iota(b->order.begin(), b->order.end(), 0);
        std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return cnt[(size_t)a] > cnt[(size_t)c]; });
        copy_wait(b->lab.p, lab.data(), (size_t)total * 8, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->off.p, offsets, (size_t)(nsys + 1) * 8, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->next.p, next.data(), (size_t)nsys * 8, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->loglen.p, zeros.data(), (size_t)nsys * 8, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->cnt.p, cnt.data(), (size_t)nsys * 4, hipMemcpyHostToDevice, b->stream);
        copy_wait(b->ord.p, b->order.data(), (size_t)nsys * 4, hipMemcpyHostToDevice, b->stream);
        zero_wait(b->errf.p, sizeof(int), b->stream);
        *out = b;
        return NB_OK;
    });
}

int nb_batch_destroy(nb_batch *b) {
    if (!b) return NB_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return NB_OK;
}

const char *nb_batch_last_error(const nb_batch *b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int nb_batch_set_params(nb_batch *b, double G, double dt, double eps, int mode, int mergers, double min_mom,
                        int precision) {
    if (!b) return bfail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL batch");
    if (!(G > 0) || !(dt > 0) || eps < 0 || min_mom < 0)
        return bfail(b, NB_ERR_INVALID_ARGUMENT, "G and dt must be positive, eps and min_tot_com_mom non-negative");
    if (mode != NB_NO_COLL_CHECK && mode != NB_OVERLAP_COLL_CHECK)
        return bfail(b, NB_ERR_NOT_SUPPORTED, "only collisions = 0 or 3 (overlap check) are supported");
    if (precision != NB_PREC_FP64 && precision != NB_PREC_MIXED)
        return bfail(b, NB_ERR_INVALID_ARGUMENT, "batched systems run NB_PREC_FP64 or NB_PREC_MIXED");
    b->P = BatchParams{G, dt, dt / 2, eps * eps, mode == NB_OVERLAP_COLL_CHECK ? 1 : 0, mergers ? 1 : 0, min_mom,
                       precision == NB_PREC_MIXED ? 1 : 0};
    return NB_OK;
}

int nb_batch_evolve(nb_batch *b, uint64_t steps, int64_t *elapsed_ns) {
    if (!b) return bfail(nullptr, NB_ERR_INVALID_ARGUMENT, "NULL batch");
    if (elapsed_ns) *elapsed_ns = 0;
    return bguard(b, [&]() {
        auto t0 = std::chrono::high_resolution_clock::now();
        // initial accelerations: one launch with zero steps would not compute
        // them, so the first step's opening kick uses a from a force-only pass
        const BodyView all{b->x.p, b->x.p + b->total, b->x.p + 2 * b->total, b->v.p, b->v.p + b->total,
                           b->v.p + 2 * b->total, b->a.p, b->a.p + b->total, b->a.p + 2 * b->total,
                           b->m.p, b->r.p, b->e.p, b->lab.p};
        // evolve semantics: a(x0) first (calc_acc_and_e), then `steps` steps
        BatchParams P0 = b->P;
        {
            // a(x0): run the force part only by a zero-dt step without collisions
            BatchParams Pf = P0;
            Pf.dt = 0.0;
            Pf.half_dt = 0.0;
            Pf.coll = 0;
            launch_batch_steps(b, all, Pf, b->step, 1);
        }
        const int64_t chunk = 1000;  // bound the duration of one launch
        if (!b->ev0) NB_HIP(hipEventCreate(&b->ev0));
        if (!b->ev1) NB_HIP(hipEventCreate(&b->ev1));
        NB_HIP(hipEventRecord(b->ev0, b->stream));
        int64_t launches = 0;
        for (uint64_t done = 0; done < steps;) {
            const int64_t k = (int64_t)std::min<uint64_t>(chunk, steps - done);
            launch_batch_steps(b, all, P0, b->step, k);
            NB_HIP(hipGetLastError());
            b->step += k;
            done += (uint64_t)k;
            ++launches;
        }
        NB_HIP(hipEventRecord(b->ev1, b->stream));
        int errv = 0;
        NB_HIP(hipMemcpyAsync(&errv, b->errf.p, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        NB_HIP(hipStreamSynchronize(b->stream));
        float ms = 0.f;
        NB_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->kernel_ms = ms;
        b->kernel_launches = launches;
        if (errv) throw ApiError(NB_ERR_CAPACITY, "collision pass: more than 256 overlapping partners for one body");
        auto t1 = std::chrono::high_resolution_clock::now();
        if (elapsed_ns) *elapsed_ns = (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
        return NB_OK;
    });
}

int nb_batch_kernel_time(const nb_batch *b, double *total_ms, int64_t *launches) {
    if (!b || !total_ms || !launches) return bfail(const_cast<nb_batch *>(b), NB_ERR_INVALID_ARGUMENT, "NULL argument");
    *total_ms = b->kernel_ms;
    *launches = b->kernel_launches;
    return NB_OK;
}

int nb_batch_get_state(const nb_batch *cb, int64_t cap, int64_t *counts, double *m, double *r, double *pos,
                       double *vel, int64_t *ids) {
    nb_batch *b = const_cast<nb_batch *>(cb);
    if (!b || !counts) return bfail(b, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return bguard(b, [&]() {
        NB_HIP(hipStreamSynchronize(b->stream));
        const int64_t T = b->total;
        std::vector<int> cnt((size_t)b->nsys);
        copy_wait(cnt.data(), b->cnt.p, cnt.size() * 4, hipMemcpyDeviceToHost, b->stream);
        int64_t live = 0;
        for (int c : cnt) live += c;
        if (cap < live) throw ApiError(NB_ERR_CAPACITY, "nb_batch_get_state: capacity too small");
        std::vector<double> hx(3 * (size_t)T), hv(3 * (size_t)T), hm((size_t)T), hr((size_t)T);
        std::vector<int64_t> hl((size_t)T);
        copy_wait(hx.data(), b->x.p, hx.size() * 8, hipMemcpyDeviceToHost, b->stream);
        copy_wait(hv.data(), b->v.p, hv.size() * 8, hipMemcpyDeviceToHost, b->stream);
        copy_wait(hm.data(), b->m.p, hm.size() * 8, hipMemcpyDeviceToHost, b->stream);
        copy_wait(hr.data(), b->r.p, hr.size() * 8, hipMemcpyDeviceToHost, b->stream);
        copy_wait(hl.data(), b->lab.p, hl.size() * 8, hipMemcpyDeviceToHost, b->stream);
        int64_t k = 0;
        for (int64_t s = 0; s < b->nsys; ++s) {
            counts[s] = cnt[(size_t)s];
            for (int q = 0; q < cnt[(size_t)s]; ++q, ++k) {
                const size_t i = (size_t)(b->offsets[(size_t)s] + q);
                if (m) m[k] = hm[i];
                if (r) r[k] = hr[i];
                for (int c = 0; c < 3; ++c) {
                    if (pos) pos[3 * k + c] = hx[(size_t)c * T + i];
                    if (vel) vel[3 * k + c] = hv[(size_t)c * T + i];
                }
                if (ids) ids[k] = hl[i];
            }
        }
        return NB_OK;
    });
}

int nb_batch_merge_log(const nb_batch *cb, int64_t cap, int64_t *records, int64_t *count) {
    nb_batch *b = const_cast<nb_batch *>(cb);
    if (!b || !count) return bfail(b, NB_ERR_INVALID_ARGUMENT, "NULL argument");
    return bguard(b, [&]() {
        NB_HIP(hipStreamSynchronize(b->stream));
        std::vector<int64_t> lens((size_t)b->nsys), lg(4 * (size_t)std::max<int64_t>(b->total, 1));
        copy_wait(lens.data(), b->loglen.p, lens.size() * 8, hipMemcpyDeviceToHost, b->stream);
        copy_wait(lg.data(), b->log.p, lg.size() * 8, hipMemcpyDeviceToHost, b->stream);
        int64_t tot = 0;
        for (auto l : lens) tot += l;
        *count = tot;
        if (!records) return NB_OK;
        if (cap < tot) throw ApiError(NB_ERR_CAPACITY, "nb_batch_merge_log: capacity too small");
        int64_t k = 0;
        for (int64_t s = 0; s < b->nsys; ++s)
            for (int64_t q = 0; q < lens[(size_t)s]; ++q, ++k) {
                const size_t base = 4 * (size_t)(b->offsets[(size_t)s] + q);
                records[5 * k + 0] = s;
                for (int c = 0; c < 4; ++c) records[5 * k + 1 + c] = lg[base + c];
            }
        return NB_OK;
    });
}

}  // extern "C"
