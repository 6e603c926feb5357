/*
 * nbody_oracle.cpp -- CPU restatement of the reference's force-and-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the parity reference that
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare the
 * HIP library against.  Nothing in the product library links, loads or calls
 * it.  It is written fresh from the reference's algorithm (GregW10/GregPackagesCPP,
 * header-only C++20), with every function citing the reference lines whose
 * arithmetic order it follows.  Two instantiations are exported:
 *   *_ld : x87 long double, the reference's own type (bit-exact target vs the
 *          compiled reference, see oracle/ref_harness.cpp);
 *   *_d  : IEEE double with the same operation order (bit-exact target for the
 *          library's FP64 parity kernels).
 * Built with  g++ -O2 -ffp-contract=off  so no FMA is formed.
 *
 * Pinning (what checks each function beyond retyping its formula):
 *   direct_acc / direct_kdk (form 0): bit-exact vs the compiled reference's
 *     evolve(leapfrog_kdk) fixtures (tests/golden/ref_kdk_*); Kepler two-body
 *     closed form; Newton's third law; KDK time reversibility.
 *   direct (form 1, softened): eps->0 limit equals form 0 to rounding;
 *     energy conservation O(dt^2); brute-force per-pair sum at N<=8.
 *   energy: closed-form two-body energy; reference cumulative_pe formula.
 *   bh_build: node count / leaf membership / depths bit-exact vs the compiled
 *     reference's bh_tree<double> dump (tests/golden/ref_tree_*); node count =
 *     N + #distinct prefixes shared by >=2 bodies (brute force).
 *   bh_acc: theta = 0 equals the softened direct sum; per-body accepted-cell
 *     lists obey the reference cell_iterator DFS (brute-force recursion).
 *   s_coll (bounce and merge): per-step state bit-exact vs the compiled
 *     reference with -DGREGSYS_MERGERS (tests/golden/ref_coll_*); momentum and
 *     mass conservation; equal-mass elastic swap closed form.
 *   adjoint: central finite differences on tiny systems (tests/test_oracle.py).
 *   integrate (Euler, modified Euler, midpoint, DKD, with and without s_coll):
 *     bit-exact vs the compiled reference's evolve (tests/golden/ref_integ_*,
 *     ref_icoll_*).
 *   bh_s_coll / nn_bodies: bit-exact vs the compiled reference's own pass on
 *     its own tree (tests/golden/ref_bhcoll_*); neighbour lists equal a
 *     brute-force enumeration of touching leaf boxes (tests/test_bh_coll.py).
 *   track (body_tracker): bit-exact vs the compiled reference's tracker
 *     (tests/golden/ref_track_*); one- and two-body closed forms.
 */
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>
#include <algorithm>

namespace {

template <typename T> inline T t_sqrt(T v);
template <> inline long double t_sqrt<long double>(long double v) { return sqrtl(v); }
template <> inline double t_sqrt<double>(double v) { return std::sqrt(v); }
/* Cube root of the merged radius (gregbod.hpp:936, cbrtl).  The long-double
 * instantiation calls cbrtl as the reference does.  The double instantiation
 * uses a fixed Newton iteration built from IEEE basic operations only
 * (a = f 2^e with e a multiple of 3, y <- y - (y^3 - f)/(3 y^2) seven times
 * from y = 1), so that it is reproducible bit for bit on any IEEE machine:
 * the GPU's FP64 collision pass uses the same iteration. */
inline double newton_cbrt(double a) {
    if (!(a > 0) || !std::isfinite(a)) return std::cbrt(a);
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
template <typename T> inline T t_cbrt(T v);
template <> inline long double t_cbrt<long double>(long double v) { return cbrtl(v); }
template <> inline double t_cbrt<double>(double v) { return newton_cbrt(v); }

enum { FORM_NEWTON = 0, FORM_SOFT = 1 };

/* Pair update.  form 0 follows system::cumulative_acc (gregsys.hpp:352-357):
 *   d = p_j - p_i ; s = (dx*dx + dy*dy) + dz*dz ; w = s*d ;
 *   r3 = sqrt((wx*wx + wy*wy) + wz*wz) ; a_i += (G*m_j)/r3 * d ; a_j -= (G*m_i)/r3 * d
 * form 1 is the reference's softened kernel bh_acc (gregsys.hpp:371-375) applied
 * pairwise:  s = d.d + eps^2 ; a_i += G*(m_j/s) * (d/sqrt(s)).                      */
/* integration methods (the reference's bit flags, gregsys.hpp:156-161) */
enum { METHOD_EULER = 2, METHOD_MODIFIED_EULER = 4, METHOD_MIDPOINT = 8, METHOD_KDK = 16, METHOD_DKD = 32 };

template <typename T>
inline void pair_acc(int form, T G, T eps2, const T *pi, const T *pj, T mi, T mj, T *ai, T *aj) {
    T dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
    if (form == FORM_NEWTON) {
        T s = dx * dx + dy * dy + dz * dz;
        T wx = s * dx, wy = s * dy, wz = s * dz;
        T r3 = t_sqrt<T>(wx * wx + wy * wy + wz * wz);
        T ki = (G * mj) / r3, kj = (G * mi) / r3;
        ai[0] += ki * dx; ai[1] += ki * dy; ai[2] += ki * dz;
        aj[0] -= kj * dx; aj[1] -= kj * dy; aj[2] -= kj * dz;
    } else {
        T s = dx * dx + dy * dy + dz * dz + eps2;
        T sq = t_sqrt<T>(s);
        T ux = dx / sq, uy = dy / sq, uz = dz / sq;
        T ki = G * (mj / s), kj = G * (mi / s);
        ai[0] += ki * ux; ai[1] += ki * uy; ai[2] += ki * uz;
        aj[0] -= kj * ux; aj[1] -= kj * uy; aj[2] -= kj * uz;
    }
}

/* calc_acc_and_e / MEM_LOOP (gregsys.hpp:107-122, 439-447): zero, then i<j. */
template <typename T>
void direct_acc(int64_t n, int form, T G, T eps, const T *m, const T *pos, T *acc) {
    T eps2 = eps * eps;
    for (int64_t i = 0; i < 3 * n; ++i) acc[i] = T(0);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j)
            pair_acc<T>(form, G, eps2, pos + 3 * i, pos + 3 * j, m[i], m[j], acc + 3 * i, acc + 3 * j);
}

/* The same sums as direct_acc, one body per thread: body k takes its terms in
 * the serial loop's order (partners j < k ascending, as the inner body of pair
 * (j, k), then j > k ascending, as the outer body), so the result is bit for
 * bit direct_acc's (tests/test_oracle.py checks it).  Used where the oracle
 * runs long FP64 trajectories (the adjoint). */
template <typename T>
void direct_acc_rows(int64_t n, int form, T G, T eps, const T *m, const T *pos, T *acc) {
    T eps2 = eps * eps;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        T a[3] = {T(0), T(0), T(0)}, other[3];
        for (int64_t j = 0; j < n; ++j) {
            if (j == k) continue;
            other[0] = other[1] = other[2] = T(0);
            if (j < k)
                pair_acc<T>(form, G, eps2, pos + 3 * j, pos + 3 * k, m[j], m[k], other, a);
            else
                pair_acc<T>(form, G, eps2, pos + 3 * k, pos + 3 * j, m[k], m[j], a, other);
        }
        acc[3 * k] = a[0];
        acc[3 * k + 1] = a[1];
        acc[3 * k + 2] = a[2];
    }
}

/* Energy: set_ke (gregbod.hpp:917-920) KE = 0.5*m*|v|*|v|; cumulative_pe
 * (gregsys.hpp:358-363) / bh_pe (:376-379) pair PE added to both bodies, then
 * per-body halving and body-order totals (calc_energy :608-630). */
template <typename T>
void energy(int64_t n, int form, T G, T eps, const T *m, const T *pos, const T *vel, T *ke_out, T *pe_out) {
    T eps2 = eps * eps;
    std::vector<T> pe(n, T(0));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            T dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1], dz = pos[3 * j + 2] - pos[3 * i + 2];
            T r = form == FORM_NEWTON ? t_sqrt<T>(dx * dx + dy * dy + dz * dz)
                                      : t_sqrt<T>(dx * dx + dy * dy + dz * dz + eps2);
            T p = -(G * m[i] * m[j]) / r;
            pe[i] += p;
            pe[j] += p;
        }
    T tke = 0, tpe = 0;
    for (int64_t i = 0; i < n; ++i) {
        pe[i] /= T(2);
        T vm = t_sqrt<T>(vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] + vel[3 * i + 2] * vel[3 * i + 2]);
        T ke = T(0.5) * m[i] * vm * vm;
        tpe += pe[i];
        tke += ke;
    }
    *ke_out = tke;
    *pe_out = tpe;
}

/* ---------------- collisions: system::s_coll (gregsys.hpp:690-914) ---------------- */
template <typename T>
struct CBody {
    int64_t label;
    T m, r, e;
    T p[3], v[3], a[3];
};

/* merge ctor body(one, two) (gregbod.hpp:933-940) and com/vel_com/acc_com
 * (gregbod.hpp:1094-1150): mass-weighted means with the old masses. */
template <typename T>
inline void mw_mean(T m1, const T *x1, T m2, const T *x2, T *out) {
    T msum = m1 + m2;
    T o[3];
    for (int k = 0; k < 3; ++k) o[k] = (m1 * x1[k] + m2 * x2[k]) / msum;
    for (int k = 0; k < 3; ++k) out[k] = o[k];
}
template <typename T>
CBody<T> merge_new(const CBody<T> &one, const CBody<T> &two, int64_t label) {
    CBody<T> b;
    b.label = label;
    b.m = one.m + two.m;
    b.r = t_cbrt<T>(one.r * one.r * one.r + two.r * two.r * two.r);
    mw_mean<T>(one.m, one.p, two.m, two.p, b.p);
    mw_mean<T>(one.m, one.v, two.m, two.v, b.v);
    mw_mean<T>(one.m, one.a, two.m, two.a, b.a);
    b.e = (one.e + two.e) / 2;
    return b;
}
/* body::operator+= (gregbod.hpp:1065-1082): position/velocity/acceleration from
 * the old masses, restitution mean, then mass and radius. */
template <typename T>
void merge_into(CBody<T> &self, const CBody<T> &other) {
    mw_mean<T>(self.m, self.p, other.m, other.p, self.p);
    mw_mean<T>(self.m, self.v, other.m, other.v, self.v);
    mw_mean<T>(self.m, self.a, other.m, other.a, self.a);
    self.e = (self.e + other.e) / 2;
    self.m += other.m;
    self.r = t_cbrt<T>(self.r * self.r * self.r + other.r * other.r * other.r);
}

struct MergeRec { int64_t step, outer_label, absorbed_label, new_label; };

/* ---------------- construction-time overlap check ----------------
 * system::check_overlap (gregsys.hpp:307-342), restated loop for loop,
 * including how it walks after a merge: `continue` advances the inner index
 * but not the inner pointer, so after the first reset the body tested at
 * inner index k is bods[k - 1]; the skip at inner == outer then leaves
 * bods[outer - 1] untested and the next test meets the outer body itself
 * (distance 0, merged with itself -- a no-op while body::allow_self_addition
 * is false, gregbod.hpp:1068-1070 -- and the body at that inner index is
 * erased); a prefix merge erases the body after the merged one.  The overlap
 * test is distance(p_o, p_i) < r_o + r_i with distance = sqrt of the ordered
 * sum of squares of (p_i - p_o) (vec_ops::distance, gregvec.hpp:4701-4706).
 * merge = false returns at the first overlap: *first_outer / *first_inner
 * receive its indices (-1 when none). */
template <typename T>
void check_overlap(std::vector<CBody<T>> &bods, bool merge, int64_t *first_outer, int64_t *first_inner) {
    *first_outer = *first_inner = -1;
    int64_t num = (int64_t)bods.size();
    for (int64_t outer = 0; outer < num; ++outer) {
        int64_t ob = outer;          // outer_b as an index
        int64_t ib = outer + 1;      // inner_b as an index
        for (int64_t inner = outer + 1; inner < num; ++inner) {
            if (merge && inner == outer) {
                ++ib;
                continue;
            }
            const CBody<T> &o = bods[ob], &b = bods[ib];
            T dx = b.p[0] - o.p[0], dy = b.p[1] - o.p[1], dz = b.p[2] - o.p[2];
            if (t_sqrt<T>(dx * dx + dy * dy + dz * dz) < o.r + b.r) {
                if (!merge) {
                    *first_outer = outer;
                    *first_inner = inner;
                    return;
                }
                if (ib != ob) merge_into<T>(bods[ob], bods[ib]);  // self-addition is a no-op
                bods.erase(bods.begin() + inner);
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
    }
}


template <typename T>
struct Overlap { int64_t idx; T mo, mi; T n[3]; };

template <typename T>
inline T dot3(const T *a, const T *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Bounce pass over the overlaps of one outer body in distance order
 * (gregsys.hpp:810-837; the same arithmetic in bh_s_coll :660-684): axis
 * momenta were captured when the overlap was found. */
template <typename T, typename Key>
void bounce_all(CBody<T> *bo, std::map<Key, Overlap<T>> &overlapping, std::vector<CBody<T>> &bods) {
    for (auto &kv : overlapping) {
        Overlap<T> &ov = kv.second;
        CBody<T> &bi = bods[ov.idx];
        T o_vel = ov.mo / bo->m;
        T i_vel = ov.mi / bi.m;
        T o_minus_i = o_vel - i_vel;
        T avg_rest = (bo->e + bi.e) / 2;
        if (o_vel > 0 || i_vel < 0) {
            if (o_vel <= 0 && i_vel < 0)
                if (o_vel < i_vel) continue;
            if (o_vel > 0 && i_vel >= 0)
                if (o_vel < i_vel) continue;
            T tot = ov.mo + ov.mi;
            T msum = bo->m + bi.m;
            T new_o = (tot - bi.m * avg_rest * o_minus_i) / msum;
            T new_i = (tot + bo->m * avg_rest * o_minus_i) / msum;
            T fo = new_o - o_vel, fi = new_i - i_vel;
            for (int k = 0; k < 3; ++k) bo->v[k] += fo * ov.n[k];
            for (int k = 0; k < 3; ++k) bi.v[k] += fi * ov.n[k];
        }
    }
}

/* One s_coll pass.  merge=false is the default build; merge=true is the
 * -DGREGSYS_MERGERS build.  Follows gregsys.hpp:690-914 including: overlaps
 * keyed by distance in an ordered map (equal distances dropped, :760-761);
 * axis momenta captured at scan time; max_mom initialised once per call (:703);
 * after a merge the j scan restarts over every index except the outer one
 * (:803-807, :726-728); the merged body replaces the outer slot (:845-846). */
template <typename T>
void s_coll(std::vector<CBody<T>> &bods, bool merge, T min_tot_com_mom, int64_t step, int64_t &next_label,
            std::vector<MergeRec> &log) {
    std::map<T, Overlap<T>> overlapping;
    size_t num = bods.size();
    T max_mom = min_tot_com_mom;
    size_t outer = 0;
    while (outer < num) {
        bool first_merged = false;
        CBody<T> merged{};
        CBody<T> *bo = &bods[outer];
        int64_t outer_label = bo->label;
        overlapping.clear();
        bool rescan = false;
        while (true) {
            bool have_cand = false;
            size_t cand = 0;
            size_t start = rescan ? 0 : outer + 1;
            for (size_t inner = start; inner < num; ++inner) {
                if (inner == outer) continue;
                CBody<T> &bi = bods[inner];
                T rad_dist = bo->r + bi.r;
                T d[3] = {bi.p[0] - bo->p[0], bi.p[1] - bo->p[1], bi.p[2] - bo->p[2]};
                T dist = t_sqrt<T>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                d[0] /= dist; d[1] /= dist; d[2] /= dist;
                if (rad_dist > dist) {
                    T mom_o[3] = {bo->m * bo->v[0], bo->m * bo->v[1], bo->m * bo->v[2]};
                    T mom_i[3] = {bi.m * bi.v[0], bi.m * bi.v[1], bi.m * bi.v[2]};
                    T amo = dot3<T>(mom_o, d);
                    T ami = dot3<T>(mom_i, d);
                    if (merge) {
                        T msum = bo->m + bi.m;
                        T vc[3];
                        for (int k = 0; k < 3; ++k) vc[k] = (mom_o[k] + mom_i[k]) / msum;
                        T curr = amo - ami - (bo->m - bi.m) * dot3<T>(vc, d);
                        if (curr >= max_mom) {
                            max_mom = curr;
                            have_cand = true;
                            cand = inner;
                            continue;
                        }
                        if (!have_cand) {
                            Overlap<T> ov{(int64_t)inner, amo, ami, {d[0], d[1], d[2]}};
                            overlapping.emplace(dist, ov);
                        }
                    } else {
                        Overlap<T> ov{(int64_t)inner, amo, ami, {d[0], d[1], d[2]}};
                        overlapping.emplace(dist, ov);
                    }
                }
            }
            if (!have_cand) break;
            /* merge outer with candidate (gregsys.hpp:766-808) */
            int64_t absorbed = bods[cand].label;
            if (!first_merged) {
                merged = merge_new<T>(*bo, bods[cand], next_label++);
                first_merged = true;
            } else {
                merge_into<T>(merged, bods[cand]);
            }
            bo = &merged;
            log.push_back(MergeRec{step, outer_label, absorbed, merged.label});
            bods.erase(bods.begin() + (std::ptrdiff_t)cand);
            if (cand < outer) --outer;
            --num;
            overlapping.clear();
            rescan = true;
        }
        bounce_all<T>(bo, overlapping, bods);
        if (first_merged) bods[outer] = merged;
        ++outer;
    }
}

/* Optional per-step trace of every body after the step (and its s_coll):
 * (step, label) + (m, r, x, y, z, vx, vy, vz) -- the same record the reference
 * harness dumps from evolve's per-body callback. */
bool g_trace_on = false;
std::vector<int64_t> g_trace_hdr;
std::vector<long double> g_trace_vals;

/* Leapfrog KDK, direct summation (gregsys.hpp:2946-2968 with loop_kdk :477-513
 * and leapfrog_kdk_acc_e_and_step :514-526): initial a(x0); per step
 * v += half_dt*a ; x += dt*v ; recompute a ; v += half_dt*a ; then s_coll when
 * collisions are on (FUNC_TEMPL_SELECT :16-57). */
template <typename T>
int integrate(int method, int64_t *n_io, int form, T G, T dt, T eps, int64_t steps, int coll, T min_mom, T *m, T *r,
              T *e, T *pos, T *vel, T *acc, int64_t *labels, int64_t *log_out, int64_t log_cap, int64_t *log_len) {
    int64_t n = *n_io;
    T half_dt = dt / 2;
    std::vector<CBody<T>> bods(n);
    for (int64_t i = 0; i < n; ++i) {
        CBody<T> &b = bods[i];
        b.label = labels ? labels[i] : i;
        b.m = m[i]; b.r = r ? r[i] : T(0); b.e = e ? e[i] : T(1);
        for (int k = 0; k < 3; ++k) { b.p[k] = pos[3 * i + k]; b.v[k] = vel[3 * i + k]; b.a[k] = T(0); }
    }
    int64_t next_label = 0;
    for (auto &b : bods) next_label = std::max(next_label, b.label + 1);
    std::vector<MergeRec> log;
    T eps2 = eps * eps;
    auto force = [&]() {
        size_t nb = bods.size();
        for (auto &b : bods) b.a[0] = b.a[1] = b.a[2] = T(0);
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = i + 1; j < nb; ++j)
                pair_acc<T>(form, G, eps2, bods[i].p, bods[j].p, bods[i].m, bods[j].m, bods[i].a, bods[j].a);
    };
    // predicted positions, velocities and accelerations of modified Euler and
    // midpoint (the reference's `predicted` tuples)
    std::vector<T> pp, pv, pa;
    auto predict = [&](T h) {
        size_t nb = bods.size();
        pp.assign(3 * nb, T(0));
        pv.assign(3 * nb, T(0));
        pa.assign(3 * nb, T(0));
        for (size_t i = 0; i < nb; ++i)
            for (int k = 0; k < 3; ++k) {
                pp[3 * i + k] = bods[i].p[k] + h * bods[i].v[k];
                pv[3 * i + k] = bods[i].v[k] + h * bods[i].a[k];
            }
        for (size_t i = 0; i < nb; ++i)
            for (size_t j = i + 1; j < nb; ++j)
                pair_acc<T>(form, G, eps2, &pp[3 * i], &pp[3 * j], bods[i].m, bods[j].m, &pa[3 * i], &pa[3 * j]);
    };
    if (method == METHOD_KDK) force();
    for (int64_t s = 1; s <= steps; ++s) {
        if (method == METHOD_KDK) {
            for (auto &b : bods)
                for (int k = 0; k < 3; ++k) {
                    b.v[k] += half_dt * b.a[k];
                    b.p[k] += dt * b.v[k];
                }
            force();
            for (auto &b : bods)
                for (int k = 0; k < 3; ++k) b.v[k] += half_dt * b.a[k];
        } else if (method == METHOD_EULER) {
            // calc_acc_and_e, then take_euler_step (gregsys.hpp:380-391, 2666-2712)
            force();
            for (auto &b : bods)
                for (int k = 0; k < 3; ++k) {
                    b.p[k] += b.v[k] * dt;
                    b.v[k] += b.a[k] * dt;
                }
        } else if (method == METHOD_MODIFIED_EULER) {
            // predictor with dt, corrector with the mean slopes (:392-405, 2714-2803)
            force();
            predict(dt);
            for (size_t i = 0; i < bods.size(); ++i)
                for (int k = 0; k < 3; ++k) {
                    bods[i].p[k] += half_dt * (bods[i].v[k] + pv[3 * i + k]);
                    bods[i].v[k] += half_dt * (bods[i].a[k] + pa[3 * i + k]);
                }
        } else if (method == METHOD_MIDPOINT) {
            // slopes at the half-step prediction (:406-419, 2804-2891)
            force();
            predict(half_dt);
            for (size_t i = 0; i < bods.size(); ++i)
                for (int k = 0; k < 3; ++k) {
                    bods[i].p[k] += dt * pv[3 * i + k];
                    bods[i].v[k] += dt * pa[3 * i + k];
                }
        } else {
            // leapfrog DKD: drift half, forces, kick full, drift half (:540-607, 2993-3039)
            for (auto &b : bods)
                for (int k = 0; k < 3; ++k) b.p[k] += half_dt * b.v[k];
            force();
            for (auto &b : bods)
                for (int k = 0; k < 3; ++k) {
                    b.v[k] += dt * b.a[k];
                    b.p[k] += half_dt * b.v[k];
                }
        }
        if (coll) s_coll<T>(bods, coll == 2, min_mom, s, next_label, log);
        if (g_trace_on)
            for (auto &b : bods) {
                g_trace_hdr.push_back(s);
                g_trace_hdr.push_back(b.label);
                long double v[8] = {(long double)b.m, (long double)b.r, (long double)b.p[0], (long double)b.p[1],
                                    (long double)b.p[2], (long double)b.v[0], (long double)b.v[1], (long double)b.v[2]};
                g_trace_vals.insert(g_trace_vals.end(), v, v + 8);
            }
    }
    n = (int64_t)bods.size();
    *n_io = n;
    for (int64_t i = 0; i < n; ++i) {
        const CBody<T> &b = bods[i];
        m[i] = b.m;
        if (r) r[i] = b.r;
        if (e) e[i] = b.e;
        if (labels) labels[i] = b.label;
        for (int k = 0; k < 3; ++k) { pos[3 * i + k] = b.p[k]; vel[3 * i + k] = b.v[k]; if (acc) acc[3 * i + k] = b.a[k]; }
    }
    if (log_len) {
        *log_len = (int64_t)log.size();
        for (size_t q = 0; q < log.size() && (int64_t)q < log_cap; ++q) {
            log_out[4 * q + 0] = log[q].step;
            log_out[4 * q + 1] = log[q].outer_label;
            log_out[4 * q + 2] = log[q].absorbed_label;
            log_out[4 * q + 3] = log[q].new_label;
        }
    }
    return 0;
}

/* ---------------- Barnes-Hut octree: greg8tree.hpp (T = double) ---------------- */
struct Cell {
    double btm[3], top[3];
    double cb[8][3], ct[8][3]; /* child corners, set_corners :81-131 */
    double mass;
    double com[3];
    int64_t body;     /* -1 when internal */
    int64_t parent;
    int64_t sub[8];
    int depth;
};

struct Tree {
    std::vector<Cell> cells;
    double btm[3], top[3];
};

/* bh_cube::set_corners (greg8tree.hpp:81-131): mid = (btm + top)/2.0l, which for
 * T = double is the double sum divided exactly by two. */
static void set_corners(Cell &c) {
    double h[3];
    for (int k = 0; k < 3; ++k) h[k] = (double)((long double)(c.btm[k] + c.top[k]) / 2.0L);
    for (int o = 0; o < 8; ++o) {
        int b[3] = {o & 1, (o >> 1) & 1, (o >> 2) & 1};
        for (int k = 0; k < 3; ++k) {
            c.cb[o][k] = b[k] ? h[k] : c.btm[k];
            c.ct[o][k] = b[k] ? c.top[k] : h[k];
        }
    }
}
/* vector3D::between (gregvec.hpp:2115-2120): half-open box test. */
static inline bool between(const double *p, const double *lo, const double *hi) {
    return lo[0] <= p[0] && p[0] < hi[0] && lo[1] <= p[1] && p[1] < hi[1] && lo[2] <= p[2] && p[2] < hi[2];
}

enum { BH_OK = 0, BH_INVALID_PLACEMENT = -20, BH_INDISTINGUISHABLE = -21 };

static int64_t new_cell(Tree &t, int64_t parent, const double *btm, const double *top, int64_t body, const double *m,
                        const double *pos, int *err) {
    Cell c;
    std::memset(&c, 0, sizeof(c));
    for (int k = 0; k < 3; ++k) { c.btm[k] = btm[k]; c.top[k] = top[k]; }
    c.parent = parent;
    c.body = body;
    for (int o = 0; o < 8; ++o) c.sub[o] = -1;
    c.depth = parent < 0 ? 0 : t.cells[parent].depth + 1;
    set_corners(c);
    /* check_first_body (greg8tree.hpp:132-145) */
    if (!between(pos + 3 * body, c.btm, c.top)) *err = BH_INVALID_PLACEMENT;
    c.mass = m[body];
    t.cells.push_back(c);
    return (int64_t)t.cells.size() - 1;
}

/* bh_tree::find_min_max (greg8tree.hpp:1012-1047). pad = longest/20.0l is a
 * long-double division converted back to double. */
static void find_min_max(int64_t n, const double *pos, double *btm, double *top) {
    double mn[3] = {pos[0], pos[1], pos[2]}, mx[3] = {pos[0], pos[1], pos[2]};
    for (int64_t i = 1; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (pos[3 * i + k] < mn[k]) mn[k] = pos[3 * i + k];
            if (pos[3 * i + k] > mx[k]) mx[k] = pos[3 * i + k];
        }
    double xr = mx[0] - mn[0], yr = mx[1] - mn[1], zr = mx[2] - mn[2];
    double le = xr > yr ? (xr > zr ? xr : zr) : (yr > zr ? yr : zr);
    double pad = (double)((long double)le / 20.0L);
    for (int k = 0; k < 3; ++k) btm[k] = mn[k] - pad;
    le += 2 * pad;
    for (int k = 0; k < 3; ++k) top[k] = btm[k] + le;
}

/* bh_cube::dispatch (greg8tree.hpp:146-210) and add_body (:878-909). */
static int tree_insert(Tree &t, int64_t b, const double *m, const double *pos) {
    int err = 0;
    Cell *root = &t.cells[0];
    if (!between(pos + 3 * b, root->btm, root->top)) return BH_INVALID_PLACEMENT;
    root->mass += m[b];
    int64_t c = 0;
    int64_t body = b;
    while (true) {
        int o;
        for (o = 0; o < 8; ++o)
            if (between(pos + 3 * body, t.cells[c].cb[o], t.cells[c].ct[o])) break;
        if (o == 8) return BH_INVALID_PLACEMENT;
        if (t.cells[c].sub[o] < 0) {
            double cb[3], ct[3];
            for (int k = 0; k < 3; ++k) { cb[k] = t.cells[c].cb[o][k]; ct[k] = t.cells[c].ct[o][k]; }
            int64_t nc = new_cell(t, c, cb, ct, body, m, pos, &err);
            if (err) return err;
            t.cells[c].sub[o] = nc;
            if (t.cells[c].body >= 0) {
                int64_t old = t.cells[c].body;
                const double *p1 = pos + 3 * body, *p2 = pos + 3 * old;
                if (p1[0] == p2[0] && p1[1] == p2[1] && p1[2] == p2[2]) return BH_INDISTINGUISHABLE;
                body = old;
                t.cells[c].body = -1;
                continue;
            }
            return 0;
        }
        c = t.cells[c].sub[o];
        t.cells[c].mass += m[body];
    }
}

/* bh_cube::calc_com non-RUNNING_MR_SUM form (greg8tree.hpp:211-282): leaf com =
 * position; node com = (sum over children k = 7..0 of m_k*com_k) / node mass.
 * (The reference's extra read of _sub[-1] adds parent->_mass * 0 and changes
 * nothing on a fresh tree.) */
static void calc_com(Tree &t, int64_t c, const double *pos) {
    Cell &cell = t.cells[c];
    if (cell.body >= 0) {
        for (int k = 0; k < 3; ++k) cell.com[k] = pos[3 * cell.body + k];
        return;
    }
    for (int o = 0; o < 8; ++o)
        if (cell.sub[o] >= 0) calc_com(t, cell.sub[o], pos);
    double s[3] = {0, 0, 0};
    for (int o = 7; o >= 0; --o)
        if (cell.sub[o] >= 0) {
            const Cell &ch = t.cells[cell.sub[o]];
            for (int k = 0; k < 3; ++k) s[k] += ch.mass * ch.com[k];
        }
    for (int k = 0; k < 3; ++k) t.cells[c].com[k] = s[k] / t.cells[c].mass;
}

/* bh_tree::add_bodies (greg8tree.hpp:1066-1078): root from body 0, the rest in
 * vector order. */
static int build_tree(Tree &t, int64_t n, const double *m, const double *pos) {
    t.cells.clear();
    t.cells.reserve((size_t)(2 * n + 8));
    find_min_max(n, pos, t.btm, t.top);
    int err = 0;
    new_cell(t, -1, t.btm, t.top, 0, m, pos, &err);
    if (err) return err;
    for (int64_t b = 1; b < n; ++b) {
        int rc = tree_insert(t, b, m, pos);
        if (rc) return rc;
    }
    calc_com(t, 0, pos);
    return 0;
}

/* Corrected BH walk.  Visit order, own-leaf skip and MAC follow
 * cell_iterator::find_first_cell/find_next_cell/is_source_cell
 * (greg8tree.hpp:710-755): DFS from the root's children in octant order;
 * accept a leaf (not the own leaf) or an internal cell with
 * (top.x - btm.x)/distance(p, com) < theta (distance = sqrtl of the double sum,
 * vec_ops::distance gregvec.hpp:4701-4706); otherwise descend.  The force is
 * bh_acc's softened monopole (gregsys.hpp:371-375) but ACCUMULATED (+=) over
 * the accepted cells, not assigned as in the reference. */
static void walk_body(const Tree &t, int64_t own_leaf, const double *p, double G, double eps2, long double theta,
                      double *a, int64_t *count, std::vector<int64_t> *accepted) {
    a[0] = a[1] = a[2] = 0;
    int64_t stack[512];
    int sp = 0;
    const Cell &root = t.cells[0];
    for (int o = 7; o >= 0; --o)
        if (root.sub[o] >= 0) stack[sp++] = root.sub[o];
    while (sp) {
        int64_t c = stack[--sp];
        if (c == own_leaf) continue;
        const Cell &cell = t.cells[c];
        bool accept = cell.body >= 0;
        if (!accept) {
            double dx = cell.com[0] - p[0], dy = cell.com[1] - p[1], dz = cell.com[2] - p[2];
            long double dist = sqrtl((long double)(dx * dx + dy * dy + dz * dz));
            accept = (long double)(cell.top[0] - cell.btm[0]) / dist < theta;
        }
        if (accept) {
            double dx = cell.com[0] - p[0], dy = cell.com[1] - p[1], dz = cell.com[2] - p[2];
            double s = dx * dx + dy * dy + dz * dz + eps2;
            double sq = std::sqrt(s);
            double k = G * (cell.mass / s);
            a[0] += k * (dx / sq); a[1] += k * (dy / sq); a[2] += k * (dz / sq);
            ++*count;
            if (accepted) accepted->push_back(c);
        } else {
            for (int o = 7; o >= 0; --o)
                if (cell.sub[o] >= 0) stack[sp++] = cell.sub[o];
        }
    }
}

/* nn_iterator (greg8tree.hpp:615-700): the bodies of the leaf cells whose
 * closed box touches the own leaf's box (faces, edges and corners count),
 * found by a DFS from the root's children in octant order that descends into
 * touching internal cells and skips the own leaf; a body alone in the root
 * cell has none.  (The iterator's climb back up rescans the rest of the level
 * it just finished, which finds nothing new: it neither repeats nor skips a
 * cell, so the plain DFS yields the same sequence.) */
static inline bool touches(const Cell &a, const Cell &b) {
    return !(a.top[0] < b.btm[0] || a.btm[0] > b.top[0] || a.top[1] < b.btm[1] || a.btm[1] > b.top[1] ||
             a.top[2] < b.btm[2] || a.btm[2] > b.top[2]);
}
static void nn_bodies(const Tree &t, int64_t own, std::vector<int64_t> &out) {
    out.clear();
    const Cell &bc = t.cells[own];
    if (bc.parent < 0) return;
    std::vector<int64_t> stack;
    for (int o = 7; o >= 0; --o)
        if (t.cells[0].sub[o] >= 0) stack.push_back(t.cells[0].sub[o]);
    while (!stack.empty()) {
        int64_t c = stack.back();
        stack.pop_back();
        if (c == own) continue;
        const Cell &cell = t.cells[c];
        if (!touches(bc, cell)) continue;
        if (cell.body >= 0) {
            out.push_back(cell.body);
            continue;
        }
        for (int o = 7; o >= 0; --o)
            if (cell.sub[o] >= 0) stack.push_back(cell.sub[o]);
    }
}

/* Leaves in DFS pre-order, octants 0..7: the order of bh_tree::begin() and
 * body_iterator::operator++ (greg8tree.hpp:502-614, the Morton order of the
 * leaves). */
static void leaves_in_order(const Tree &t, std::vector<int64_t> &out) {
    out.clear();
    std::vector<int64_t> stack{0};
    while (!stack.empty()) {
        int64_t c = stack.back();
        stack.pop_back();
        const Cell &cell = t.cells[c];
        if (cell.body >= 0) {
            out.push_back(c);
            continue;
        }
        for (int o = 7; o >= 0; --o)
            if (cell.sub[o] >= 0) stack.push_back(cell.sub[o]);
    }
}

/* system::bh_s_coll (gregsys.hpp:631-686): for every body in leaf order, the
 * touching-cell neighbours j (nn_iterator order) with r_o + r_j > |p_j - p_o|
 * are keyed by that distance in an ordered map (an equal distance keeps the
 * first), with the axis momenta m v . n taken when found, and bounced in
 * distance order as in s_coll.  Every pair is met from both sides; no merges.
 * The tree is the one built this step (positions do not move after it). */
template <typename T>
void bh_s_coll(const Tree &t, std::vector<CBody<T>> &bods) {
    std::vector<int64_t> leaves, nb;
    leaves_in_order(t, leaves);
    std::map<T, Overlap<T>> overlapping;
    for (int64_t leaf : leaves) {
        overlapping.clear();
        CBody<T> *bo = &bods[(size_t)t.cells[leaf].body];
        nn_bodies(t, leaf, nb);
        for (int64_t j : nb) {
            CBody<T> &bi = bods[(size_t)j];
            T rad_dist = bo->r + bi.r;
            T d[3] = {bi.p[0] - bo->p[0], bi.p[1] - bo->p[1], bi.p[2] - bo->p[2]};
            T dist = t_sqrt<T>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            d[0] /= dist; d[1] /= dist; d[2] /= dist;
            if (rad_dist > dist) {
                T mom_o[3] = {bo->m * bo->v[0], bo->m * bo->v[1], bo->m * bo->v[2]};
                T mom_i[3] = {bi.m * bi.v[0], bi.m * bi.v[1], bi.m * bi.v[2]};
                Overlap<T> ov{j, dot3<T>(mom_o, d), dot3<T>(mom_i, d), {d[0], d[1], d[2]}};
                overlapping.emplace(dist, ov);
            }
        }
        if (!overlapping.empty()) bounce_all<T>(bo, overlapping, bods);
    }
}

template <typename T>
int bh_s_coll_arrays(int64_t n, const T *m, const T *r, const T *e, const T *pos, T *vel) {
    std::vector<double> pd((size_t)(3 * n)), md((size_t)n);
    for (int64_t i = 0; i < 3 * n; ++i) pd[(size_t)i] = (double)pos[i];
    for (int64_t i = 0; i < n; ++i) md[(size_t)i] = (double)m[i];
    Tree t;
    int rc = build_tree(t, n, md.data(), pd.data());
    if (rc) return rc;
    std::vector<CBody<T>> bods((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        CBody<T> &b = bods[(size_t)i];
        b.label = i; b.m = m[i]; b.r = r[i]; b.e = e[i];
        for (int k = 0; k < 3; ++k) { b.p[k] = pos[3 * i + k]; b.v[k] = vel[3 * i + k]; b.a[k] = 0; }
    }
    bh_s_coll<T>(t, bods);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) vel[3 * i + k] = bods[(size_t)i].v[k];
    return 0;
}

/* body_tracker (gregsys.hpp:3492-3611) over the bodies idx[0..k): total mass
 * (:3527-3531), centre of mass and COM velocity (:3533-3554: the mass sum,
 * then the ordered sum of m*r, divided by the mass unless it is zero),
 * mean distance to a point (:3555-3562) and mean pairwise separation
 * (:3563-3582: distance of every pair outer < inner, divided by the integer
 * k(k-1)/2).  The long-double form keeps the reference's single running sum
 * for the separation; the double form (ROWS) sums each outer body's row first
 * and then the rows in order, the association the GPU uses. */
template <typename T, bool ROWS>
void track(int64_t k, const int64_t *idx, const T *m, const T *pos, const T *vel, const T *point, T *out) {
    T tm = 0;
    for (int64_t q = 0; q < k; ++q) tm += m[idx[q]];
    T mr[3] = {0, 0, 0}, mv[3] = {0, 0, 0};
    for (int64_t q = 0; q < k; ++q) {
        const int64_t i = idx[q];
        for (int c = 0; c < 3; ++c) mr[c] += m[i] * pos[3 * i + c];
    }
    for (int64_t q = 0; q < k; ++q) {
        const int64_t i = idx[q];
        for (int c = 0; c < 3; ++c) mv[c] += m[i] * vel[3 * i + c];
    }
    out[0] = tm;
    for (int c = 0; c < 3; ++c) {
        out[1 + c] = (k == 0 || tm == 0) ? T(0) : mr[c] / tm;
        out[4 + c] = (k == 0 || tm == 0) ? T(0) : mv[c] / tm;
    }
    T sd = 0;
    for (int64_t q = 0; q < k; ++q) {
        const int64_t i = idx[q];
        T d[3] = {pos[3 * i] - point[0], pos[3 * i + 1] - point[1], pos[3 * i + 2] - point[2]};
        sd += t_sqrt<T>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    out[7] = k ? sd / (T)(uint64_t)k : T(0);
    T tot = 0;
    for (int64_t a = 0; a < k; ++a) {
        T row = 0;
        for (int64_t b = a + 1; b < k; ++b) {
            const int64_t o = idx[a], i = idx[b];
            T d[3] = {pos[3 * o] - pos[3 * i], pos[3 * o + 1] - pos[3 * i + 1], pos[3 * o + 2] - pos[3 * i + 2]};
            T dist = t_sqrt<T>(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (ROWS) row += dist;
            else tot += dist;
        }
        if (ROWS) tot += row;
    }
    out[8] = k > 1 ? tot / (T)(uint64_t)(((uint64_t)k * (uint64_t)(k - 1)) / 2) : T(0);
}

} // namespace

extern "C" {

/* ---------------- direct summation ---------------- */
/* long double scalars travel by pointer: a by-value long double argument from
 * ctypes would be rounded through a Python float first. */
int orc_direct_acc_ld(int64_t n, int form, const long double *G, const long double *eps, const long double *m,
                      const long double *pos, long double *acc) {
    direct_acc<long double>(n, form, *G, *eps, m, pos, acc);
    return 0;
}
int orc_direct_acc_rows_d(int64_t n, int form, double G, double eps, const double *m, const double *pos,
                          double *acc) {
    direct_acc_rows<double>(n, form, G, eps, m, pos, acc);
    return 0;
}
int orc_direct_acc_d(int64_t n, int form, double G, double eps, const double *m, const double *pos, double *acc) {
    direct_acc<double>(n, form, G, eps, m, pos, acc);
    return 0;
}
int orc_energy_ld(int64_t n, int form, const long double *G, const long double *eps, const long double *m,
                  const long double *pos, const long double *vel, long double *ke, long double *pe) {
    energy<long double>(n, form, *G, *eps, m, pos, vel, ke, pe);
    return 0;
}
int orc_energy_d(int64_t n, int form, double G, double eps, const double *m, const double *pos, const double *vel,
                 double *ke, double *pe) {
    energy<double>(n, form, G, eps, m, pos, vel, ke, pe);
    return 0;
}
/* coll: 0 = off, 1 = bounce only, 2 = bounce + mergers (GREGSYS_MERGERS). */
int orc_kdk_ld(int64_t *n, int form, const long double *G, const long double *dt, const long double *eps,
               int64_t steps, int coll, const long double *min_mom, long double *m, long double *r, long double *e,
               long double *pos, long double *vel, long double *acc, int64_t *labels, int64_t *log, int64_t log_cap,
               int64_t *log_len) {
    return integrate<long double>(METHOD_KDK, n, form, *G, *dt, *eps, steps, coll, *min_mom, m, r, e, pos, vel, acc,
                                  labels, log, log_cap, log_len);
}
int orc_kdk_d(int64_t *n, int form, double G, double dt, double eps, int64_t steps, int coll, double min_mom,
              double *m, double *r, double *e, double *pos, double *vel, double *acc, int64_t *labels, int64_t *log,
              int64_t log_cap, int64_t *log_len) {
    return integrate<double>(METHOD_KDK, n, form, G, dt, eps, steps, coll, min_mom, m, r, e, pos, vel, acc, labels,
                             log, log_cap, log_len);
}
/* The other direct-sum integrators of evolve (method = the reference's flag:
 * 2 Euler, 4 modified Euler, 8 midpoint, 16 leapfrog KDK, 32 leapfrog DKD),
 * with s_coll after every step when coll != 0. */
int orc_integrate_ld(int method, int64_t *n, int form, const long double *G, const long double *dt,
                     const long double *eps, int64_t steps, int coll, const long double *min_mom, long double *m,
                     long double *r, long double *e, long double *pos, long double *vel, long double *acc,
                     int64_t *labels, int64_t *log, int64_t log_cap, int64_t *log_len) {
    return integrate<long double>(method, n, form, *G, *dt, *eps, steps, coll, *min_mom, m, r, e, pos, vel, acc,
                                  labels, log, log_cap, log_len);
}
int orc_integrate_d(int method, int64_t *n, int form, double G, double dt, double eps, int64_t steps, int coll,
                    double min_mom, double *m, double *r, double *e, double *pos, double *vel, double *acc,
                    int64_t *labels, int64_t *log, int64_t log_cap, int64_t *log_len) {
    return integrate<double>(method, n, form, G, dt, eps, steps, coll, min_mom, m, r, e, pos, vel, acc, labels, log,
                             log_cap, log_len);
}
/* One collision pass only, on a given state (no force, no kick). */
int orc_s_coll_d(int64_t *n_io, int merge, double min_mom, int64_t step, double *m, double *r, double *e,
                 double *pos, double *vel, double *acc, int64_t *labels, int64_t *next_label, int64_t *log,
                 int64_t log_cap, int64_t *log_len) {
    int64_t n = *n_io;
    std::vector<CBody<double>> bods(n);
    for (int64_t i = 0; i < n; ++i) {
        CBody<double> &b = bods[i];
        b.label = labels[i]; b.m = m[i]; b.r = r[i]; b.e = e[i];
        for (int k = 0; k < 3; ++k) { b.p[k] = pos[3 * i + k]; b.v[k] = vel[3 * i + k]; b.a[k] = acc[3 * i + k]; }
    }
    std::vector<MergeRec> lg;
    s_coll<double>(bods, merge != 0, min_mom, step, *next_label, lg);
    n = (int64_t)bods.size();
    *n_io = n;
    for (int64_t i = 0; i < n; ++i) {
        const CBody<double> &b = bods[i];
        labels[i] = b.label; m[i] = b.m; r[i] = b.r; e[i] = b.e;
        for (int k = 0; k < 3; ++k) { pos[3 * i + k] = b.p[k]; vel[3 * i + k] = b.v[k]; acc[3 * i + k] = b.a[k]; }
    }
    *log_len = (int64_t)lg.size();
    for (size_t q = 0; q < lg.size() && (int64_t)q < log_cap; ++q) {
        log[4 * q + 0] = lg[q].step; log[4 * q + 1] = lg[q].outer_label;
        log[4 * q + 2] = lg[q].absorbed_label; log[4 * q + 3] = lg[q].new_label;
    }
    return 0;
}

/* check_overlap: merge = 0 -> *first = (outer, inner) of the first overlap or
 * (-1, -1); merge = 1 -> the merged bodies replace the arrays, *n_io the new
 * count, labels follow the surviving bodies. */
}  // extern "C"
template <typename T>
static int check_overlap_arrays(int64_t *n_io, int merge, T *m, T *r, T *e, T *pos, T *vel, int64_t *labels,
                                int64_t *first) {
    int64_t n = *n_io;
    std::vector<CBody<T>> bods(n);
    for (int64_t i = 0; i < n; ++i) {
        CBody<T> &b = bods[i];
        b.label = labels[i]; b.m = m[i]; b.r = r[i]; b.e = e[i];
        for (int k = 0; k < 3; ++k) { b.p[k] = pos[3 * i + k]; b.v[k] = vel[3 * i + k]; b.a[k] = T(0); }
    }
    check_overlap<T>(bods, merge != 0, first, first + 1);
    n = (int64_t)bods.size();
    *n_io = n;
    for (int64_t i = 0; i < n; ++i) {
        const CBody<T> &b = bods[i];
        labels[i] = b.label; m[i] = b.m; r[i] = b.r; e[i] = b.e;
        for (int k = 0; k < 3; ++k) { pos[3 * i + k] = b.p[k]; vel[3 * i + k] = b.v[k]; }
    }
    return 0;
}
extern "C" {
int orc_check_overlap_d(int64_t *n_io, int merge, double *m, double *r, double *e, double *pos, double *vel,
                        int64_t *labels, int64_t *first) {
    return check_overlap_arrays<double>(n_io, merge, m, r, e, pos, vel, labels, first);
}
int orc_check_overlap_ld(int64_t *n_io, int merge, long double *m, long double *r, long double *e, long double *pos,
                         long double *vel, int64_t *labels, int64_t *first) {
    return check_overlap_arrays<long double>(n_io, merge, m, r, e, pos, vel, labels, first);
}

/* ---------------- Barnes-Hut ---------------- */
/* Builds the reference octree and exports it in DFS pre-order (octants 0..7):
 * per node depth, octant index within its parent, body index (-1 internal),
 * parent (pre-order index), mass, com, side (top.x - btm.x).  Returns the node
 * count (>0) or a negative error (capacity too small, invalid placement,
 * indistinguishable bodies). */
int64_t orc_bh_tree_d(int64_t n, const double *m, const double *pos, int64_t cap, int32_t *depth, int32_t *octant,
                      int64_t *body, int64_t *parent, double *mass, double *com, double *side, double *box) {
    Tree t;
    int rc = build_tree(t, n, m, pos);
    if (rc) return rc;
    if ((int64_t)t.cells.size() > cap) return -1;
    std::vector<int64_t> order;
    order.reserve(t.cells.size());
    std::vector<int64_t> pre(t.cells.size(), -1);
    std::vector<std::pair<int64_t, int>> st;
    st.push_back({0, -1});
    while (!st.empty()) {
        auto [c, oct] = st.back();
        st.pop_back();
        int64_t idx = (int64_t)order.size();
        pre[c] = idx;
        order.push_back(c);
        const Cell &cell = t.cells[c];
        depth[idx] = cell.depth;
        octant[idx] = oct;
        body[idx] = cell.body;
        parent[idx] = cell.parent < 0 ? -1 : pre[cell.parent];
        mass[idx] = cell.mass;
        for (int k = 0; k < 3; ++k) com[3 * idx + k] = cell.com[k];
        side[idx] = cell.top[0] - cell.btm[0];
        for (int o = 7; o >= 0; --o)
            if (cell.sub[o] >= 0) st.push_back({cell.sub[o], o});
    }
    if (box) for (int k = 0; k < 3; ++k) { box[k] = t.btm[k]; box[3 + k] = t.top[k]; }
    return (int64_t)t.cells.size();
}

/* Corrected BH accelerations for all bodies; returns the total number of
 * accepted cells (the algorithmic interaction count) in *interactions. */
int orc_bh_acc_d(int64_t n, double G, double eps, double theta, const double *m, const double *pos, double *acc,
                 int64_t *interactions, int64_t *per_body) {
    Tree t;
    int rc = build_tree(t, n, m, pos);
    if (rc) return rc;
    std::vector<int64_t> leaf_of(n, -1);
    for (size_t c = 0; c < t.cells.size(); ++c)
        if (t.cells[c].body >= 0) leaf_of[t.cells[c].body] = (int64_t)c;
    int64_t total = 0;
    double eps2 = eps * eps;
    for (int64_t i = 0; i < n; ++i) {
        int64_t cnt = 0;
        walk_body(t, leaf_of[i], pos + 3 * i, G, eps2, (long double)theta, acc + 3 * i, &cnt, nullptr);
        if (per_body) per_body[i] = cnt;
        total += cnt;
    }
    if (interactions) *interactions = total;
    return 0;
}

/* ---------------- all-core CPU baselines (timing only) ----------------
 * FP64 restatements spread over host threads with OpenMP, for the all-core
 * CPU figure printed beside the GPU throughput.  They are not parity
 * references: the direct form sums all j for every i (no pair symmetry, the
 * GPU's work), softened pair force a_i += G m_j d / (s sqrt(s)),
 * s = |d|^2 + eps^2; the BH form builds the same insertion tree serially and
 * walks the bodies in parallel. */
}  // extern "C"

namespace {
__attribute__((target_clones("arch=haswell", "default")))
void direct_rows_mt(int64_t i0, int64_t i1, int64_t n, double G, double eps2, const double *m, const double *x,
                    const double *y, const double *z, double *acc) {
    for (int64_t i = i0; i < i1; ++i) {
        const double xi = x[i], yi = y[i], zi = z[i];
        double ax = 0, ay = 0, az = 0;
#pragma omp simd reduction(+ : ax, ay, az)
        for (int64_t j = 0; j < n; ++j) {
            const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
            const double s = dx * dx + dy * dy + dz * dz + eps2;
            const double w = m[j] / (s * std::sqrt(s));
            ax += w * dx;
            ay += w * dy;
            az += w * dz;
        }
        acc[3 * i] = G * ax;
        acc[3 * i + 1] = G * ay;
        acc[3 * i + 2] = G * az;
    }
}
}  // namespace

extern "C" {

int orc_direct_acc_mt(int64_t n, double G, double eps, const double *m, const double *pos, double *acc,
                      int threads) {
    std::vector<double> x(n), y(n), z(n);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = pos[3 * i];
        y[i] = pos[3 * i + 1];
        z[i] = pos[3 * i + 2];
    }
    const double eps2 = eps * eps;
    const int64_t rows = 64;
#pragma omp parallel for schedule(dynamic) num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t i0 = 0; i0 < n; i0 += rows)
        direct_rows_mt(i0, std::min(n, i0 + rows), n, G, eps2, m, x.data(), y.data(), z.data(), acc);
    return 0;
}

int orc_bh_acc_mt(int64_t n, double G, double eps, double theta, const double *m, const double *pos, double *acc,
                  int64_t *interactions, int threads) {
    Tree t;
    int rc = build_tree(t, n, m, pos);
    if (rc) return rc;
    std::vector<int64_t> leaf_of(n, -1);
    for (size_t c = 0; c < t.cells.size(); ++c)
        if (t.cells[c].body >= 0) leaf_of[t.cells[c].body] = (int64_t)c;
    int64_t total = 0;
    const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total) \
    num_threads(threads > 0 ? threads : omp_get_max_threads())
    for (int64_t i = 0; i < n; ++i) {
        int64_t cnt = 0;
        walk_body(t, leaf_of[i], pos + 3 * i, G, eps2, (long double)theta, acc + 3 * i, &cnt, nullptr);
        total += cnt;
    }
    if (interactions) *interactions = total;
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

/* Accepted-cell list of one body, as pre-order node indices (see orc_bh_tree_d). */
int64_t orc_bh_accepted_d(int64_t n, double theta, const double *m, const double *pos, int64_t body_idx,
                          int64_t cap, int64_t *out) {
    Tree t;
    int rc = build_tree(t, n, m, pos);
    if (rc) return rc;
    std::vector<int64_t> pre(t.cells.size());
    {
        std::vector<int64_t> st{0};
        int64_t k = 0;
        while (!st.empty()) {
            int64_t c = st.back();
            st.pop_back();
            pre[c] = k++;
            for (int o = 7; o >= 0; --o)
                if (t.cells[c].sub[o] >= 0) st.push_back(t.cells[c].sub[o]);
        }
    }
    int64_t own = -1;
    for (size_t c = 0; c < t.cells.size(); ++c)
        if (t.cells[c].body == body_idx) own = (int64_t)c;
    std::vector<int64_t> acc_list;
    double a[3];
    int64_t cnt = 0;
    walk_body(t, own, pos + 3 * body_idx, 1.0, 0.0, (long double)theta, a, &cnt, &acc_list);
    for (size_t q = 0; q < acc_list.size() && (int64_t)q < cap; ++q) out[q] = pre[acc_list[q]];
    return (int64_t)acc_list.size();
}

/* BH leapfrog as evolve(leapfrog_kdk | barnes_hut) (gregsys.hpp:2893-2936):
 * initial a, v += half_dt*a, then per step: drift x += dt*v, rebuild, a,
 * v += dt*a (velocities stay half a step ahead).  Forces accumulated. */
int orc_bh_kdk_d(int64_t n, double G, double dt, double eps, double theta, int64_t steps, const double *m,
                 double *pos, double *vel, double *acc) {
    double half_dt = dt / 2;
    int64_t cnt;
    int rc = orc_bh_acc_d(n, G, eps, theta, m, pos, acc, &cnt, nullptr);
    if (rc) return rc;
    for (int64_t i = 0; i < 3 * n; ++i) vel[i] += half_dt * acc[i];
    for (int64_t s = 0; s < steps; ++s) {
        for (int64_t i = 0; i < 3 * n; ++i) pos[i] += dt * vel[i];
        rc = orc_bh_acc_d(n, G, eps, theta, m, pos, acc, &cnt, nullptr);
        if (rc) return rc;
        for (int64_t i = 0; i < 3 * n; ++i) vel[i] += dt * acc[i];
    }
    return 0;
}

/* One bh_s_coll pass on a state (velocities updated in place), in long double
 * (the reference's system<long double> arithmetic) or in double (the GPU's
 * FP64 form).  The octree is built from the positions in double, as in the
 * other BH functions here. */
int orc_bh_s_coll_ld(int64_t n, const long double *m, const long double *r, const long double *e,
                     const long double *pos, long double *vel) {
    return bh_s_coll_arrays<long double>(n, m, r, e, pos, vel);
}
int orc_bh_s_coll_d(int64_t n, const double *m, const double *r, const double *e, const double *pos, double *vel) {
    return bh_s_coll_arrays<double>(n, m, r, e, pos, vel);
}

/* body_tracker reductions: out = mass, com[3], com_vel[3], mean distance to
 * `point`, mean separation. */
int orc_track_ld(int64_t k, const int64_t *idx, const long double *m, const long double *pos, const long double *vel,
                 const long double *point, long double *out) {
    track<long double, false>(k, idx, m, pos, vel, point, out);
    return 0;
}
int orc_track_d(int64_t k, const int64_t *idx, const double *m, const double *pos, const double *vel,
                const double *point, double *out) {
    track<double, true>(k, idx, m, pos, vel, point, out);
    return 0;
}

/* Neighbour bodies of body b as nn_iterator yields them (for tests). */
int64_t orc_bh_neighbours_d(int64_t n, const double *m, const double *pos, int64_t b, int64_t cap, int64_t *out) {
    Tree t;
    int rc = build_tree(t, n, m, pos);
    if (rc) return rc;
    int64_t leaf = -1;
    for (size_t c = 0; c < t.cells.size(); ++c)
        if (t.cells[c].body == b) leaf = (int64_t)c;
    std::vector<int64_t> nb;
    nn_bodies(t, leaf, nb);
    for (size_t k = 0; k < nb.size() && (int64_t)k < cap; ++k) out[k] = nb[k];
    return (int64_t)nb.size();
}

/* orc_bh_kdk_d followed, after each step's kick, by bh_s_coll on that step's
 * tree when coll != 0 (gregsys.hpp:2914-2939). */
int orc_bh_kdk_coll_d(int64_t n, double G, double dt, double eps, double theta, int64_t steps, int coll,
                      const double *m, const double *r, const double *e, double *pos, double *vel, double *acc) {
    double half_dt = dt / 2;
    int64_t cnt;
    int rc = orc_bh_acc_d(n, G, eps, theta, m, pos, acc, &cnt, nullptr);
    if (rc) return rc;
    for (int64_t i = 0; i < 3 * n; ++i) vel[i] += half_dt * acc[i];
    for (int64_t s = 0; s < steps; ++s) {
        for (int64_t i = 0; i < 3 * n; ++i) pos[i] += dt * vel[i];
        rc = orc_bh_acc_d(n, G, eps, theta, m, pos, acc, &cnt, nullptr);
        if (rc) return rc;
        for (int64_t i = 0; i < 3 * n; ++i) vel[i] += dt * acc[i];
        if (coll) {
            rc = bh_s_coll_arrays<double>(n, m, r, e, pos, vel);
            if (rc) return rc;
        }
    }
    return 0;
}

/* ---------------- adjoint (FP64 reference gradient) ----------------
 * Reverse-mode derivative of the softened direct KDK above (form 1), written
 * from the chain rule on the three sub-steps of one KDK step
 *   v' = v + h a(x) ; x1 = x + dt v' ; v1 = v' + h a(x1)      (h = dt/2)
 * with, for a_i = G sum_j m_j d_ij s_ij^{-3/2} (d_ij = x_j - x_i, s = |d|^2+eps^2):
 *   (J^T u)_k   = G sum_j [ s^{-3/2} w - 3 s^{-5/2} d_kj (d_kj . w) ],  w = m_k u_j - m_j u_k
 *   (dA/dm)^T u_k = -G sum_j s^{-3/2} (u_j . d_kj)
 * All forward states are stored (small N only).  Loss cotangents come in
 * (gx, gv at time T) and the gradients w.r.t. x0, v0, m come out. */
static void vjp(int64_t n, double G, double eps2, const double *m, const double *x, const double *u, double *gx,
                double *gm, double scale) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double ax = 0, ay = 0, az = 0, am = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == k) continue;
            double dx = x[3 * j] - x[3 * k], dy = x[3 * j + 1] - x[3 * k + 1], dz = x[3 * j + 2] - x[3 * k + 2];
            double s = dx * dx + dy * dy + dz * dz + eps2;
            double inv = 1.0 / std::sqrt(s);
            double inv3 = inv * inv * inv;
            double inv5 = inv3 * inv * inv;
            double wx = m[k] * u[3 * j] - m[j] * u[3 * k];
            double wy = m[k] * u[3 * j + 1] - m[j] * u[3 * k + 1];
            double wz = m[k] * u[3 * j + 2] - m[j] * u[3 * k + 2];
            double dw = dx * wx + dy * wy + dz * wz;
            ax += inv3 * wx - 3.0 * inv5 * dx * dw;
            ay += inv3 * wy - 3.0 * inv5 * dy * dw;
            az += inv3 * wz - 3.0 * inv5 * dz * dw;
            am += -inv3 * (u[3 * j] * dx + u[3 * j + 1] * dy + u[3 * j + 2] * dz);
        }
        gx[3 * k] += scale * G * ax;
        gx[3 * k + 1] += scale * G * ay;
        gx[3 * k + 2] += scale * G * az;
        gm[k] += scale * G * am;
    }
}

int orc_adjoint_d(int64_t n, double G, double dt, double eps, int64_t steps, const double *m, const double *x0,
                  const double *v0, const double *gxT, const double *gvT, double *gx0, double *gv0, double *gm) {
    double h = dt / 2, eps2 = eps * eps;
    std::vector<std::vector<double>> xs(steps + 1, std::vector<double>(3 * n));
    std::vector<double> v(v0, v0 + 3 * n), a(3 * n);
    xs[0].assign(x0, x0 + 3 * n);
    direct_acc_rows<double>(n, FORM_SOFT, G, eps, m, xs[0].data(), a.data());
    for (int64_t s = 0; s < steps; ++s) {
        std::vector<double> &x = xs[s + 1];
        for (int64_t i = 0; i < 3 * n; ++i) {
            v[i] += h * a[i];
            x[i] = xs[s][i] + dt * v[i];
        }
        direct_acc_rows<double>(n, FORM_SOFT, G, eps, m, x.data(), a.data());
        for (int64_t i = 0; i < 3 * n; ++i) v[i] += h * a[i];
    }
    std::vector<double> lx(gxT, gxT + 3 * n), lv(gvT, gvT + 3 * n);
    std::fill(gm, gm + n, 0.0);
    for (int64_t s = steps - 1; s >= 0; --s) {
        /* v1 = v' + h a(x1) */
        vjp(n, G, eps2, m, xs[s + 1].data(), lv.data(), lx.data(), gm, h);
        /* x1 = x + dt v' */
        for (int64_t i = 0; i < 3 * n; ++i) lv[i] += dt * lx[i];
        /* v' = v + h a(x) */
        vjp(n, G, eps2, m, xs[s].data(), lv.data(), lx.data(), gm, h);
    }
    std::copy(lx.begin(), lx.end(), gx0);
    std::copy(lv.begin(), lv.end(), gv0);
    return 0;
}

void orc_trace(int on) {
    g_trace_on = on != 0;
    g_trace_hdr.clear();
    g_trace_vals.clear();
}
int64_t orc_trace_len(void) { return (int64_t)(g_trace_hdr.size() / 2); }
void orc_trace_copy(int64_t *hdr, long double *vals) {
    std::copy(g_trace_hdr.begin(), g_trace_hdr.end(), hdr);
    std::copy(g_trace_vals.begin(), g_trace_vals.end(), vals);
}

/* Forward-only helper for finite differences: final x, v of the softened KDK. */
int orc_forward_d(int64_t n, double G, double dt, double eps, int64_t steps, const double *m, double *x, double *v) {
    int64_t nn = n;
    std::vector<double> acc(3 * n);
    return integrate<double>(METHOD_KDK, &nn, FORM_SOFT, G, dt, eps, steps, 0, 0.0, const_cast<double *>(m), nullptr,
                             nullptr, x, v, acc.data(), nullptr, nullptr, 0, nullptr);
}

} // extern "C"
