/*
 * gtd/nbody.hpp -- C++ facade with the shape of the reference's gtd::body and
 * gtd::system (GregW10/GregPackagesCPP gregbod.hpp / gregsys.hpp), running the
 * force-and-step path on an MI355X through the C ABI of gregnbody.h.
 *
 * Kept from the reference: the class and method names a caller uses
 * (body(mass, radius, pos, vel, restitution), mass()/rad()/pos()/vel()/
 * acceleration()/get_id(); system(bodies, timestep, iterations, units_format,
 * allow_overlapping_bodies), set_iterations/set_timestep/set_softening/
 * set_bh_opening_angle/set_G_units/set_min_tot_com_mom (bool returns),
 * evolve(method, continue_ev_if) returning std::chrono::nanoseconds,
 * num_bodies(), G_val(), operator[], begin()/end()), the method bit flags
 * (gregsys.hpp:156-165) and the exception types (gregbod.hpp:19-93,
 * greg8tree.hpp:11-46), re-declared here.
 * Also kept: body_tracker(system, predicate) with mass/com/com_vel/
 * mean_dist_to/mean_sep.
 * Different by design: state is FP64 (M = R = T = double); the memFreq/
 * fileFreq template arguments stay 0 (recording is nb_set_recording on the
 * C ABI); evolve's per-body callbacks are not supported (a continue_ev_if
 * predicate is: it is polled once per step).
 */
#ifndef GTD_NBODY_HPP
#define GTD_NBODY_HPP

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <ios>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../gregnbody.h"

namespace gtd {

/* ---- exceptions (same names and hierarchy as the reference) ---- */
class nbody_error : public std::logic_error {
public:
    nbody_error() : std::logic_error{"Invalid parameters for an N-body simulation.\n"} {}
    explicit nbody_error(const char *m) : std::logic_error{m} {}
};
class overlapping_bodies_error : public nbody_error { using nbody_error::nbody_error; };
class negative_mass_error : public nbody_error { using nbody_error::nbody_error; };
class negative_radius_error : public nbody_error { using nbody_error::nbody_error; };
class two_body_error : public nbody_error { using nbody_error::nbody_error; };
class empty_system_error : public nbody_error { using nbody_error::nbody_error; };
class no_evolution_error : public nbody_error { using nbody_error::nbody_error; };
class barnes_hut_error : public nbody_error { using nbody_error::nbody_error; };
class invalid_body_placement : public barnes_hut_error { using barnes_hut_error::barnes_hut_error; };
class indistinguishable_bodies : public barnes_hut_error { using barnes_hut_error::barnes_hut_error; };
class nsys_load_error : public nbody_error { using nbody_error::nbody_error; };
class nsys_format_error : public nsys_load_error { using nsys_load_error::nsys_load_error; };
class device_error : public std::runtime_error { using std::runtime_error::runtime_error; };

namespace detail {
inline void raise(int code, const char *msg) {
    switch (code) {
        case NB_OK: return;
        case NB_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case NB_ERR_NEGATIVE_MASS: throw negative_mass_error(msg);
        case NB_ERR_NEGATIVE_RADIUS: throw negative_radius_error(msg);
        case NB_ERR_OVERLAPPING_BODIES: throw overlapping_bodies_error(msg);
        case NB_ERR_TWO_BODY: throw two_body_error(msg);
        case NB_ERR_EMPTY_SYSTEM: throw empty_system_error(msg);
        case NB_ERR_INVALID_PLACEMENT: throw invalid_body_placement(msg);
        case NB_ERR_INDISTINGUISHABLE: throw indistinguishable_bodies(msg);
        case NB_ERR_NO_EVOLUTION: throw no_evolution_error(msg);
        case NB_ERR_NSYS_LOAD: throw nsys_load_error(msg);
        case NB_ERR_NSYS_FORMAT: throw nsys_format_error(msg);
        case NB_ERR_DEVICE: throw device_error(msg);
        default: throw nbody_error(msg);
    }
}
}  // namespace detail

template <typename T = double>
struct vector3D {
    T x{}, y{}, z{};
    const T &get_x() const noexcept { return x; }
    const T &get_y() const noexcept { return y; }
    const T &get_z() const noexcept { return z; }
};

/* A body's state as the caller sees it (gregbod.hpp:890-1087). */
class body {
    double mass_ = 1, radius_ = 1, rest_c_ = 1;
    vector3D<double> pos_{}, vel_{}, acc_{};
    uint64_t id_ = 0;
    template <typename, typename, typename, bool, bool, int, uint64_t, uint64_t, bool>
    friend class system;

public:
    body() = default;
    body(double body_mass, double body_radius, const vector3D<double> &pos = {}, const vector3D<double> &vel = {},
         double restitution = 1)
        : mass_{body_mass}, radius_{body_radius}, rest_c_{restitution}, pos_{pos}, vel_{vel} {
        if (mass_ < 0) throw negative_mass_error("Mass cannot be negative.\n");
        if (radius_ < 0) throw negative_radius_error("Radius cannot be negative.\n");
    }
    const double &mass() const noexcept { return mass_; }
    const double &rad() const noexcept { return radius_; }
    const vector3D<double> &pos() const noexcept { return pos_; }
    const vector3D<double> &vel() const noexcept { return vel_; }
    const vector3D<double> &acceleration() const noexcept { return acc_; }
    double restitution() const noexcept { return rest_c_; }
    uint64_t get_id() const noexcept { return id_; }
};

template <typename M, typename R, typename T, uint64_t rF>
class body_tracker;

template <typename M = double, typename R = double, typename T = double, bool prog = false,
          bool mergeOverlappingBodies = false, int collisions = 0, uint64_t memFreq = 0, uint64_t fileFreq = 0,
          bool binaryFile = false>
class system {
    template <typename, typename, typename, uint64_t>
    friend class body_tracker;
    static_assert(std::is_same_v<M, double> && std::is_same_v<R, double> && std::is_same_v<T, double>,
                  "the MI355X engine keeps masses, radii and positions in FP64");
    static_assert(memFreq == 0 && fileFreq == 0, "trajectory recording is not on the engine's path");
    static_assert(collisions == 0 || collisions == 3, "collisions = 0 or 3 (overlap check)");
    nb_system *h_ = nullptr;
    mutable std::vector<body> cache_;
    mutable bool stale_ = true;
    uint64_t iterations_ = 1000;

    void check(int rc) const { detail::raise(rc, nb_last_error(h_)); }
    void refresh() const {
        if (!stale_) return;
        int64_t n = 0;
        check(nb_num_bodies(h_, &n));
        std::vector<double> m(n), r(n), e(n), p(3 * n), v(3 * n), a(3 * n);
        std::vector<int64_t> ids(n);
        check(nb_get_state(h_, n, m.data(), r.data(), p.data(), v.data(), a.data(), e.data()));
        check(nb_body_ids(h_, n, ids.data()));
        cache_.resize(n);
        for (int64_t i = 0; i < n; ++i) {
            body &b = cache_[i];
            b.mass_ = m[i]; b.radius_ = r[i]; b.rest_c_ = e[i];
            b.pos_ = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
            b.vel_ = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
            b.acc_ = {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
            b.id_ = (uint64_t)ids[i];
        }
        stale_ = false;
    }
    static void type_check(int is_fp, int is_signed, int size, const char *type, const char *what) {
        (void)is_signed;
        char msg[256];
        if (!is_fp) {
            // the reference's wording for this case (NSYS_TYPE_CHECK, floating template parameter)
            snprintf(msg, sizeof msg,
                     "Invalid .nsys format: d_types field states that %s values are stored in floating-point format, "
                     "but template parameter %s is integral.\n",
                     what, type);
            throw nsys_load_error(msg);
        }
        if (size != (int)sizeof(double)) {
            snprintf(msg, sizeof msg,
                     "Size mismatch error: size of %s template parameter (%zu bytes) does not match its reported size "
                     "(%d bytes) in the .nsys file.\n",
                     type, sizeof(double), size);
            throw nsys_load_error(msg);
        }
    }
    void upload(const std::vector<body> &bods, bool allow_overlapping) {
        const size_t n = bods.size();
        std::vector<double> m(n), r(n), e(n), p(3 * n), v(3 * n);
        for (size_t i = 0; i < n; ++i) {
            m[i] = bods[i].mass_; r[i] = bods[i].radius_; e[i] = bods[i].rest_c_;
            p[3 * i] = bods[i].pos_.x; p[3 * i + 1] = bods[i].pos_.y; p[3 * i + 2] = bods[i].pos_.z;
            v[3 * i] = bods[i].vel_.x; v[3 * i + 1] = bods[i].vel_.y; v[3 * i + 2] = bods[i].vel_.z;
        }
        if (allow_overlapping || !mergeOverlappingBodies)
            check(nb_set_bodies(h_, (int64_t)n, m.data(), r.data(), p.data(), v.data(), e.data(),
                                allow_overlapping ? 0 : 1));
        else
            throw nbody_error("merging overlapping bodies at construction is not on the engine's path");
        stale_ = true;
    }

public:
    static constexpr int two_body = NB_TWO_BODY;
    static constexpr int euler = NB_EULER;
    static constexpr int modified_euler = NB_MODIFIED_EULER;
    static constexpr int midpoint = NB_MIDPOINT;
    static constexpr int leapfrog_kdk = NB_LEAPFROG_KDK;
    static constexpr int leapfrog_dkd = NB_LEAPFROG_DKD;
    static constexpr int rk4 = NB_RK4;
    static constexpr int rk3_8 = NB_RK3_8;
    static constexpr int barnes_hut = NB_BARNES_HUT;
    static constexpr int fast_multipole = NB_FAST_MULTIPOLE;
    static constexpr int no_coll_check = NB_NO_COLL_CHECK;
    static constexpr int overlap_coll_check = NB_OVERLAP_COLL_CHECK;
    static constexpr long double G_SI = 0.000000000066743L;

    explicit system(int device = 0) {
        detail::raise(nb_create(device, &h_), nb_last_error(nullptr));
        if constexpr (collisions == 3) check(nb_set_collisions(h_, NB_OVERLAP_COLL_CHECK, 0, 1e10));
    }
    system(const std::vector<body> &bodies, long double timestep = 1, uint64_t num_iterations = 1000,
           const char *units_format = "M1:1,D1:1,T1:1", bool allow_overlapping_bodies = false, int device = 0)
        : system(device) {
        if (timestep <= 0) throw std::invalid_argument{"Time-step cannot be zero or negative.\n"};
        check(nb_set_timestep(h_, (double)timestep));
        set_iterations(num_iterations);
        set_G_units(units_format);
        upload(bodies, allow_overlapping_bodies);
    }
    /* .nsys constructor (gregsys.hpp:1479-1490).  The stored types must be this
     * system's (floating, 8 bytes), checked in load_from_nsys's order and words
     * (NSYS_TYPE_CHECK :989-1046). */
    explicit system(const char *nsys_path, bool alloc_chunks = true, bool check_nan_inf = true,
                    bool allow_overlapping_bodies = true, int device = 0)
        : system(device) {
        (void)alloc_chunks;  // the engine reads the chunks in one pass either way
        if (nsys_path == nullptr) throw std::invalid_argument{"nullptr passed as .nsys file path.\n"};
        nb_nsys_info info{};
        const int rc = nb_nsys_read(nsys_path, check_nan_inf, &info, 0, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc != NB_OK && rc != NB_ERR_CAPACITY) detail::raise(rc, nb_last_error(nullptr));
        type_check(info.d_types & 0b10, info.d_types & 0b10000, info.m_size, "M", "mass");
        type_check(info.d_types & 0b100, info.d_types & 0b100000, info.r_size, "R", "radius");
        type_check(info.d_types & 0b1000, info.d_types & 0b1000000, info.t_size, "T", "position and velocity");
        check(nb_load_nsys(h_, nsys_path, check_nan_inf, allow_overlapping_bodies));
        check(nb_set_iterations(h_, iterations_));
        stale_ = true;
    }
    system(const system &) = delete;
    system &operator=(const system &) = delete;
    ~system() { nb_destroy(h_); }

    /* engine extensions */
    void set_precision(int precision) { check(nb_set_precision(h_, precision)); }
    void set_mergers(bool on, double min_tot_com_mom = 1e10) {
        check(nb_set_collisions(h_, collisions, on ? 1 : 0, min_tot_com_mom));
    }
    nb_system *handle() noexcept { return h_; }

    bool set_iterations(uint64_t number) noexcept {
        if (nb_set_iterations(h_, number) != NB_OK) return false;
        iterations_ = number;
        return true;
    }
    bool set_timestep(const long double &delta_t) noexcept { return nb_set_timestep(h_, (double)delta_t) == NB_OK; }
    bool set_softening(long double eps) noexcept { return nb_set_softening(h_, (double)eps) == NB_OK; }
    bool set_bh_opening_angle(long double theta) { return nb_set_bh_opening_angle(h_, (double)theta) == NB_OK; }
    bool set_min_tot_com_mom(double total_momentum) requires (collisions != 0) {
        return nb_set_collisions(h_, collisions, 1, total_momentum) == NB_OK;
    }
    void set_G_units(const char *units_format) { check(nb_set_units(h_, units_format)); }
    long double G_val() const {
        double g;
        check(nb_get_G(h_, &g));
        return g;
    }
    typename std::vector<body>::size_type num_bodies() const {
        int64_t n = 0;
        check(nb_num_bodies(h_, &n));
        return (typename std::vector<body>::size_type)n;
    }
    void add_bodies(const std::vector<body> &extra, bool allow_overlapping_bodies = false) {
        refresh();
        std::vector<body> all = cache_;
        all.insert(all.end(), extra.begin(), extra.end());
        upload(all, allow_overlapping_bodies);
    }

    /* evolve(int integration_method, continue_ev_if) (gregsys.hpp:2593-2607) */
    const std::chrono::nanoseconds evolve(int integration_method = leapfrog_kdk) {
        int64_t ns = 0;
        check(nb_evolve(h_, integration_method, &ns));
        stale_ = true;
        return std::chrono::nanoseconds{ns};
    }
    template <typename evFuncT>
    const std::chrono::nanoseconds evolve(int integration_method, const evFuncT &continue_ev_if) {
        static_assert(std::is_invocable_r_v<bool, evFuncT>, "continue_ev_if must be callable as bool()");
        auto t0 = std::chrono::high_resolution_clock::now();
        const uint64_t saved = iterations_;
        if (!continue_ev_if()) return std::chrono::nanoseconds::zero();
        check(nb_set_iterations(h_, 1));
        check(nb_evolve(h_, integration_method, nullptr));  // a(x0) and the first step
        stale_ = true;
        while (continue_ev_if()) {
            check(nb_advance(h_, 1, nullptr));
            stale_ = true;
        }
        check(nb_set_iterations(h_, saved));
        return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::high_resolution_clock::now() - t0);
    }

    /* to_nsys (gregsys.hpp:3214-3258): the file size, or 0 if it could not be
     * written. */
    std::streamoff to_nsys(const char *path) const {
        if (path == nullptr) return 0;
        int64_t bytes = 0;
        if (nb_save_nsys(h_, path, (int)sizeof(T), &bytes) != NB_OK) return 0;
        return (std::streamoff)bytes;
    }
    long double elapsed_time() const {
        double t = 0;
        check(nb_time_elapsed(h_, &t));
        return t;
    }

    /* energies of the current state (FP64) */
    std::pair<double, double> kinetic_potential() const {
        double ke, pe;
        check(nb_energy(h_, &ke, &pe));
        return {ke, pe};
    }

    const body &operator[](typename std::vector<body>::size_type index) const {
        refresh();
        if (index >= cache_.size())
            throw std::out_of_range("The requested body does not exist (index out of range).\n");
        return cache_[index];
    }
    auto begin() const {
        refresh();
        return cache_.cbegin();
    }
    auto end() const {
        refresh();
        return cache_.cend();
    }
};

/* body_tracker (gregsys.hpp:3492-3611): the bodies of a system that a
 * predicate picks at construction, with the tracker's reductions (mass, com,
 * com_vel, mean_dist_to, mean_sep) computed on the GPU by nb_track.  As in
 * the reference, it assumes that no merger happens and no body is added to the
 * system after construction. */
template <typename M = double, typename R = double, typename T = double, uint64_t rF = 0>
class body_tracker {
    nb_system *h_ = nullptr;
    std::vector<int64_t> idx_;
    void check(int rc) const { detail::raise(rc, nb_last_error(h_)); }
    const int64_t *ids() const noexcept { return idx_.empty() ? nullptr : idx_.data(); }

public:
    template <bool prog, bool merge, int coll, uint64_t fF, bool bin, typename F>
    body_tracker(const system<M, R, T, prog, merge, coll, rF, fF, bin> &sys, const F &pred,
                 typename std::vector<int64_t>::size_type reserve_bods = 0)
        : h_{sys.h_} {
        if (reserve_bods) idx_.reserve(reserve_bods);
        int64_t k = 0;
        for (const body &b : sys) {
            if (pred(b)) idx_.push_back(k);
            ++k;
        }
    }
    uint64_t num_bods() const noexcept { return idx_.size(); }
    M mass() const {
        double m = 0;
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, &m, nullptr, nullptr, nullptr, nullptr));
        return m;
    }
    vector3D<T> com() const {
        double c[3] = {0, 0, 0};
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, nullptr, c, nullptr, nullptr, nullptr));
        return {c[0], c[1], c[2]};
    }
    vector3D<T> com_vel() const {
        double c[3] = {0, 0, 0};
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, nullptr, nullptr, c, nullptr, nullptr));
        return {c[0], c[1], c[2]};
    }
    T mean_dist_to(const vector3D<T> &point) const {
        const double p[3] = {point.x, point.y, point.z};
        double d = 0;
        check(nb_track(h_, ids(), (int64_t)idx_.size(), p, nullptr, nullptr, nullptr, &d, nullptr));
        return d;
    }
    T mean_sep() const {
        if (idx_.empty()) throw nbody_error{"Error: no bodies being tracked.\n"};
        double d = 0;
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, nullptr, nullptr, nullptr, nullptr, &d));
        return d;
    }
};

}  // namespace gtd

#endif  // GTD_NBODY_HPP
