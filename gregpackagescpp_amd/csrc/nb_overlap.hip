// nb_overlap.hip -- the construction-time overlap check of gtd::system
// (system::check_overlap, gregsys.hpp:307-342), with and without merging.
//
// Without merging the reference throws at the first overlapping pair in
// (outer, inner > outer) order; that pair is found here by one GPU pass over
// all pairs (each workgroup tests 256 outer bodies against a chunk of inner
// bodies staged through LDS; a 64-bit atomic minimum of (outer << 32 | inner)
// keeps the first pair), instead of the reference's O(N^2) host loop.
//
// With mergeOverlappingBodies the reference merges an overlapping pair
// (body::operator+=, gregbod.hpp:1065-1082), erases the inner body and
// restarts the inner loop.  Its restart leaves the inner pointer one element
// behind the inner index (the loop's `continue` advances only the index), so
// after a first merge into outer o it tests o against bods[0 .. o-2] (the
// index o skips bods[o-1]) and then against itself: distance 0 < 2 r_o, a
// self-addition that body::allow_self_addition (false) turns into a no-op,
// followed by the erasure of the body after it.  Rounds repeat until no
// body follows o, unless a test against bods[q], q <= o-2, merges first (that
// erases bods[q+1] and shifts o down).  The engine reproduces that loop:
// the GPU pass finds the first outer with an overlap (before any merge the
// loop is the plain lexicographic scan), and the host then follows the loop
// from that merge on, on the downloaded state, with the shortcut that a
// round whose only event is the self-addition ends in the truncation the
// repeated rounds would reach.  A zero-radius outer (distance 0 < 0 fails)
// keeps scanning literally; outers after it go back to the GPU pass.
//
// The overlap test is the reference's in FP64: d = sqrt((dx*dx + dy*dy) +
// dz*dz) with d = p_inner - p_outer, overlap iff d < r_outer + r_inner.  The
// GPU pass first compares d^2 with (r_o + r_i)^2 widened by 2^-30, far above
// the rounding of either side, and takes the square root only inside it.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "nb_system.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {

constexpr int kOvBlock = 256;
constexpr int64_t kOvChunk = 8192;  // inner bodies per workgroup
constexpr unsigned long long kNone = ~0ull;

__global__ __launch_bounds__(kOvBlock) void k_first_overlap(int64_t n, int64_t from, const double *__restrict__ x,
                                                           const double *__restrict__ r,
                                                           unsigned long long *__restrict__ best) {
    __shared__ double tx[kOvBlock], ty[kOvBlock], tz[kOvBlock], tr[kOvBlock];
    const int64_t i0 = from + (int64_t)blockIdx.x * kOvBlock;
    const int64_t i = i0 + threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.y * kOvChunk;
    const int64_t c1 = min(n, c0 + kOvChunk);
    if (c1 <= i0 + 1) return;  // every inner index of the chunk is <= every outer index of the block
    // a pair found earlier with a smaller outer index cannot be beaten here
    if ((int64_t)(__hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) < i0) return;
    const bool live = i < n;
    const double xi = live ? x[i] : 0.0, yi = live ? x[n + i] : 0.0, zi = live ? x[2 * n + i] : 0.0;
    const double ri = live ? r[i] : 0.0;
    bool found = !live;
    for (int64_t jt = max(c0, i0 + 1); jt < c1; jt += kOvBlock) {
        const int64_t cnt = min((int64_t)kOvBlock, c1 - jt);
        __syncthreads();
        if ((int64_t)threadIdx.x < cnt) {
            const int64_t j = jt + threadIdx.x;
            tx[threadIdx.x] = x[j];
            ty[threadIdx.x] = x[n + j];
            tz[threadIdx.x] = x[2 * n + j];
            tr[threadIdx.x] = r[j];
        }
        __syncthreads();
        if (found) continue;
        for (int k = 0; k < (int)cnt; ++k) {
            const int64_t j = jt + k;
            if (j <= i) continue;
            const double dx = tx[k] - xi, dy = ty[k] - yi, dz = tz[k] - zi;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double s = ri + tr[k];
            if (d2 > (s * s) * (1.0 + 0x1.0p-30)) continue;
            if (__builtin_sqrt(d2) < s) {
                atomicMin(best, ((unsigned long long)i << 32) | (unsigned long long)j);
                found = true;
                break;
            }
        }
    }
}

// Host body for the merge loop (SoA state downloaded once).
struct HBody {
    int64_t lab;
    double m, r, e;
    double p[3], v[3], a[3];
};

double host_newton_cbrt(double a) {
    if (!(a > 0) || std::isinf(a)) return std::cbrt(a);
    int e;
    double f = std::frexp(a, &e);
    int r = ((e % 3) + 3) % 3;
    f = std::ldexp(f, r);
    e -= r;
    double y = 1.0;
    for (int k = 0; k < 7; ++k) {
        double y2 = y * y;
        y = y - (y2 * y - f) / (3.0 * y2);
    }
    return std::ldexp(y, e / 3);
}

// body::operator+= (gregbod.hpp:1065-1082): position, velocity and
// acceleration as mass-weighted means with the old masses, the mean
// restitution, then mass and radius (the same arithmetic as the collision
// pass's merge, so merged radii use the same Newton cube root).
void host_merge_into(HBody &s, const HBody &t) {
    double q[9];
    for (int k = 0; k < 3; ++k) q[k] = (s.m * s.p[k] + t.m * t.p[k]) / (s.m + t.m);
    for (int k = 0; k < 3; ++k) q[3 + k] = (s.m * s.v[k] + t.m * t.v[k]) / (s.m + t.m);
    for (int k = 0; k < 3; ++k) q[6 + k] = (s.m * s.a[k] + t.m * t.a[k]) / (s.m + t.m);
    for (int k = 0; k < 3; ++k) {
        s.p[k] = q[k];
        s.v[k] = q[3 + k];
        s.a[k] = q[6 + k];
    }
    s.e = (s.e + t.e) / 2;
    s.m += t.m;
    s.r = host_newton_cbrt(s.r * s.r * s.r + t.r * t.r * t.r);
}

bool host_overlap(const HBody &o, const HBody &b) {
    const double dx = b.p[0] - o.p[0], dy = b.p[1] - o.p[1], dz = b.p[2] - o.p[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz) < o.r + b.r;
}

void download(nb_system *s, std::vector<HBody> &h) {
    const int64_t n = s->n;
    std::vector<double> m(n), r(n), e(n), x(3 * n), v(3 * n), a(3 * n);
    copy_wait(m.data(), s->m.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
    copy_wait(r.data(), s->r.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
    copy_wait(e.data(), s->e.p, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream);
    copy_wait(x.data(), s->x.p, (size_t)(3 * n) * 8, hipMemcpyDeviceToHost, s->stream);
    copy_wait(v.data(), s->v.p, (size_t)(3 * n) * 8, hipMemcpyDeviceToHost, s->stream);
    copy_wait(a.data(), s->a.p, (size_t)(3 * n) * 8, hipMemcpyDeviceToHost, s->stream);
    h.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        HBody &b = h[(size_t)i];
        b.lab = s->labels[(size_t)i];
        b.m = m[i]; b.r = r[i]; b.e = e[i];
        for (int k = 0; k < 3; ++k) {
            b.p[k] = x[k * n + i];
            b.v[k] = v[k * n + i];
            b.a[k] = a[k * n + i];
        }
    }
}

void upload(nb_system *s, const std::vector<HBody> &h) {
    const int64_t n = (int64_t)h.size();
    std::vector<double> m(n), r(n), e(n), x(3 * n), v(3 * n), a(3 * n);
    for (int64_t i = 0; i < n; ++i) {
        const HBody &b = h[(size_t)i];
        m[i] = b.m; r[i] = b.r; e[i] = b.e;
        for (int k = 0; k < 3; ++k) {
            x[k * n + i] = b.p[k];
            v[k * n + i] = b.v[k];
            a[k * n + i] = b.a[k];
        }
    }
    s->n = n;
    if (n) {
        copy_wait(s->m.p, m.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
        copy_wait(s->r.p, r.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
        copy_wait(s->e.p, e.data(), (size_t)n * 8, hipMemcpyHostToDevice, s->stream);
        copy_wait(s->x.p, x.data(), (size_t)(3 * n) * 8, hipMemcpyHostToDevice, s->stream);
        copy_wait(s->v.p, v.data(), (size_t)(3 * n) * 8, hipMemcpyHostToDevice, s->stream);
        copy_wait(s->a.p, a.data(), (size_t)(3 * n) * 8, hipMemcpyHostToDevice, s->stream);
    }
    s->labels.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) s->labels[(size_t)i] = h[(size_t)i].lab;
}

// The inner loop of check_overlap for `outer`, entered at a merge with
// `inner` (the pointers in step), followed exactly from there.  Returns the
// outer index when the inner loop ends.
int64_t follow_merges(std::vector<HBody> &b, int64_t outer, int64_t inner) {
    int64_t num = (int64_t)b.size();
    int64_t ob = outer, ib = inner;
    for (; inner < num; ++inner) {
        if (inner == outer) {
            ++ib;
            continue;
        }
        if (host_overlap(b[(size_t)ob], b[(size_t)ib])) {
            if (ib == ob) {
                // self-addition: a no-op, then bods[inner] is erased.  Every
                // later round repeats the same tests on an unchanged prefix
                // and outer, so the rounds erase every body after the outer.
                b.resize((size_t)outer + 1);
                return outer;
            }
            host_merge_into(b[(size_t)ob], b[(size_t)ib]);
            b.erase(b.begin() + inner);
            if (inner < outer) {
                --outer;
                --ob;
            }
            inner = 0;  // the loop's ++inner makes it 1; the pointer stays at element 0
            ib = 0;
            --num;
            continue;
        }
        ++ib;
    }
    return outer;
}

}  // namespace

// First (outer, inner > outer) overlapping pair with outer >= from, or false.
bool first_overlap(nb_system *s, int64_t from, int64_t *outer, int64_t *inner) {
    const int64_t n = s->n;
    if (n - from < 2) return false;
    s->dw.ovl_best.ensure(1);
    const unsigned long long none = kNone;
    NB_HIP(hipMemcpyAsync(s->dw.ovl_best.p, &none, sizeof(none), hipMemcpyHostToDevice, s->stream));
    const unsigned gx = (unsigned)((n - from + kOvBlock - 1) / kOvBlock);
    const unsigned gy = (unsigned)((n + kOvChunk - 1) / kOvChunk);
    hipLaunchKernelGGL(k_first_overlap, dim3(gx, gy), dim3(kOvBlock), 0, s->stream, n, from, s->x.p, s->r.p,
                       s->dw.ovl_best.p);
    NB_HIP(hipGetLastError());
    unsigned long long best = kNone;
    copy_wait(&best, s->dw.ovl_best.p, sizeof(best), hipMemcpyDeviceToHost, s->stream);
    if (best == kNone) return false;
    *outer = (int64_t)(best >> 32);
    *inner = (int64_t)(best & 0xffffffffull);
    return true;
}

// check_overlap with mergeOverlappingBodies on the context's bodies; returns
// the number of bodies removed.
int64_t merge_overlapping(nb_system *s) {
    const int64_t n0 = s->n;
    int64_t from = 0, o = 0, j = 0;
    bool changed = false;
    std::vector<HBody> h;
    while (first_overlap(s, from, &o, &j)) {
        download(s, h);
        const int64_t last = follow_merges(h, o, j);
        upload(s, h);
        changed = true;
        from = last + 1;
    }
    if (changed) {
        s->acc_valid = false;
        s->evolved = false;
    }
    return n0 - s->n;
}

}  // namespace nbe
