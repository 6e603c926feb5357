/*
 * gtd/nbody.hpp -- C++ facade with the shape of the reference's gtd::body and
 * gtd::system (GregW10/GregPackagesCPP gregbod.hpp / gregsys.hpp), running the
 * force-and-step path on an MI355X through the C ABI of gregnbody.h.
 *
 * A reference driver builds against it unchanged for the parts listed here:
 *   - gtd::system<M, R, T, prog, mergeOverlappingBodies, collisions, memFreq,
 *     fileFreq, binaryFile> with the reference's defaults (long double, no
 *     progress, no merging, no collisions, memFreq 0, fileFreq 1, binary);
 *     M, R and T may be any arithmetic type and are converted to and from the
 *     engine's FP64 at the boundary;
 *   - its constructors (bodies, timestep, iterations, units, allow_overlapping;
 *     initializer lists; timestep/iterations/units; .nsys files), setters with
 *     bool returns, add_body / emplace_body / add_bodies / add_system,
 *     get_body / remove_body / front / back, operator[], begin()/end(),
 *     num_bodies(), G_val(), elapsed_time(), to_nsys();
 *   - evolve(method, continue_ev_if, for_each_bod, for_all_bods) with the
 *     callbacks run where the reference runs them (for_all_bods after each
 *     step's update and before the collision pass, for_each_bod after it);
 *     the engine stops at those points and the state makes a host round trip;
 *   - mergeOverlappingBodies (construction-time merging, nb_merge_overlapping),
 *     collisions = 3 with -DGREGSYS_MERGERS honoured as in the reference,
 *     memFreq (a device ring of states and energies: body::prev_pos_at /
 *     prev_vel_at / prev_ke_at, pe_extrema / ke_extrema / tot_e_extrema) and
 *     prog (the reference's progress lines and closing summary);
 *   - gtd::body<M, R, T, recFreq> with the reference's constructors, accessors
 *     and setters, process-wide unique ids (gtd::body_counter: the smallest
 *     free id), vector3D<T> with the arithmetic drivers use, the method bit
 *     flags (gregsys.hpp:156-165) and the exception types (gregbod.hpp:19-93,
 *     greg8tree.hpp:11-46), re-declared here;
 *   - body_tracker(system, predicate) with mass/com/com_vel/mean_dist_to/
 *     mean_sep computed on the GPU.
 * Differences, each stated where it applies: state is FP64 on the GPU (long
 * double values are rounded to double on the way in); the reference writes no
 * files for fileFreq (its branch is empty), and neither does this facade; a
 * for_all_bods callback may change bodies but not their number; Barnes-Hut
 * steps run no callbacks, as in the reference.
 */
#ifndef GTD_NBODY_HPP
#define GTD_NBODY_HPP

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <initializer_list>
#include <ios>
#include <iostream>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
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
/* evolve's option check (check_option, gregsys.hpp:1406-1411): one method
 * bit in the low word (or none), at most one force-approximation bit. */
inline bool check_option(int option) noexcept {
    const uint16_t lo = (uint16_t)(option & 0x0000ffff);
    const uint16_t hi = (uint16_t)((unsigned)option >> 16);
    return !(lo & (lo - 1)) || !lo || lo > NB_RK4 || !(hi & (hi - 1)) || hi > 2;
}
}  // namespace detail

/* ---- unique body ids (body_counter, gregbod.hpp:99-207): a process-wide
 *      set; a new body takes the smallest free id, a moved-from body a fresh
 *      id from the top of the range ---- */
class body_counter {
    static inline uint64_t count = 0;
    static inline std::set<uint64_t> ids;
    static inline std::mutex id_mutex;

protected:
    uint64_t id;

public:
    body_counter() {
        std::lock_guard<std::mutex> lock{id_mutex};
        id = 0;
        while (ids.count(id)) ++id;
        ids.insert(id);
        ++count;
    }
    body_counter(const body_counter &) : body_counter() {}  // a copy is a new body with its own id
    body_counter(body_counter &&other) noexcept {
        std::lock_guard<std::mutex> lock{id_mutex};
        id = other.id;
        other.id = (uint64_t)-1;
        while (ids.count(other.id)) --other.id;
        ids.insert(other.id);
        ++count;
    }
    body_counter &operator=(const body_counter &) noexcept { return *this; }  // ids stay with their objects
    body_counter &operator=(body_counter &&other) noexcept {
        std::lock_guard<std::mutex> lock{id_mutex};
        ids.erase(id);
        id = other.id;
        other.id = (uint64_t)-1;
        while (ids.count(other.id)) --other.id;
        ids.insert(other.id);
        return *this;
    }
    virtual ~body_counter() {
        std::lock_guard<std::mutex> lock{id_mutex};
        --count;
        ids.erase(id);
    }
    uint64_t get_id() const noexcept { return id; }
    bool set_id(uint64_t new_id) {
        std::lock_guard<std::mutex> lock{id_mutex};
        if (!ids.insert(new_id).second) return false;
        ids.erase(id);
        id = new_id;
        return true;
    }
    static uint64_t body_count() noexcept { return count; }
    static std::vector<uint64_t> all_ids() {
        std::lock_guard<std::mutex> lock{id_mutex};
        return {ids.begin(), ids.end()};
    }
    explicit operator uint64_t() const noexcept { return id; }
};

/* ---- 3-vectors with the arithmetic the drivers use (gregvec.hpp) ---- */
template <typename T = long double>
struct vector3D {
    T x{}, y{}, z{};
    constexpr vector3D() noexcept = default;
    constexpr vector3D(const T &xc, const T &yc, const T &zc = T{}) noexcept : x{xc}, y{yc}, z{zc} {}
    template <typename U, typename = std::enable_if_t<!std::is_same_v<U, T>>>
    constexpr vector3D(const vector3D<U> &o) noexcept : x{(T)o.x}, y{(T)o.y}, z{(T)o.z} {}
    const T &get_x() const noexcept { return x; }
    const T &get_y() const noexcept { return y; }
    const T &get_z() const noexcept { return z; }
    vector3D &set_x(const T &v) noexcept { x = v; return *this; }
    vector3D &set_y(const T &v) noexcept { y = v; return *this; }
    vector3D &set_z(const T &v) noexcept { z = v; return *this; }
    long double magnitude() const noexcept {
        return std::sqrt((long double)x * x + (long double)y * y + (long double)z * z);
    }
    vector3D &make_zero() noexcept { x = y = z = T{}; return *this; }
    vector3D operator-() const noexcept { return {-x, -y, -z}; }
    vector3D &operator+=(const vector3D &o) noexcept { x += o.x; y += o.y; z += o.z; return *this; }
    vector3D &operator-=(const vector3D &o) noexcept { x -= o.x; y -= o.y; z -= o.z; return *this; }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    vector3D &operator*=(const S &s) noexcept { x *= s; y *= s; z *= s; return *this; }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    vector3D &operator/=(const S &s) noexcept { x /= s; y /= s; z /= s; return *this; }
    friend vector3D operator+(vector3D a, const vector3D &b) noexcept { return a += b; }
    friend vector3D operator-(vector3D a, const vector3D &b) noexcept { return a -= b; }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    friend vector3D operator*(vector3D a, const S &s) noexcept { return a *= s; }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    friend vector3D operator*(const S &s, vector3D a) noexcept { return a *= s; }
    template <typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
    friend vector3D operator/(vector3D a, const S &s) noexcept { return a /= s; }
    friend T operator*(const vector3D &a, const vector3D &b) noexcept { return a.x * b.x + a.y * b.y + a.z * b.z; }
    friend bool operator==(const vector3D &a, const vector3D &b) noexcept { return a.x == b.x && a.y == b.y && a.z == b.z; }
    friend bool operator!=(const vector3D &a, const vector3D &b) noexcept { return !(a == b); }
    /* the reference's str(): "xi + yj + zk" with the signs folded in */
    friend std::ostream &operator<<(std::ostream &os, const vector3D &v) {
        os << v.x << 'i' << (v.y < 0 ? " - " : " + ") << (v.y < 0 ? -v.y : v.y) << 'j'
           << (v.z < 0 ? " - " : " + ") << (v.z < 0 ? -v.z : v.z) << 'k';
        return os;
    }
};

template <typename M, typename R, typename T, bool, bool, int, uint64_t, uint64_t, bool>
class system;
template <typename M, typename R, typename T, uint64_t rF>
class body_tracker;

/* ---- a body (gtd::body<M, R, T, recFreq>, gregbod.hpp:890-1087) ---- */
template <typename M = long double, typename R = long double, typename T = long double, uint64_t recFreq = 0>
class body : public body_counter {
    M mass_{1};
    R radius{1};
    vector3D<T> curr_pos{}, curr_vel{}, acc{};
    T pe{};
    long double rest_c{1};
    // history kept every recFreq steps of an evolve (filled from the engine's records)
    std::vector<vector3D<T>> positions, velocities;
    std::vector<long double> energies;
    template <typename, typename, typename, bool, bool, int, uint64_t, uint64_t, bool>
    friend class system;
    template <typename, typename, typename, uint64_t>
    friend class body;
    void check_mass() const {
        if (mass_ < M{0}) throw negative_mass_error("Mass cannot be negative.\n");
    }
    void check_radius() const {
        if (radius < R{0}) throw negative_radius_error("Radius cannot be negative.\n");
    }

public:
    static inline bool allow_self_addition = false;
    body() = default;
    body(const M &body_mass, const R &body_radius) : mass_{body_mass}, radius{body_radius} {
        check_mass();
        check_radius();
    }
    body(const M &body_mass, const R &body_radius, const vector3D<T> &pos, const vector3D<T> &vel,
         const long double &restitution = 1)
        : mass_{body_mass}, radius{body_radius}, curr_pos{pos}, curr_vel{vel}, rest_c{restitution} {
        check_mass();
        check_radius();
    }
    body(const M &body_mass, const R &body_radius, const T &xpos, const T &ypos, const T &zpos, const T &xvel,
         const T &yvel, const T &zvel, const long double &restitution)
        : mass_{body_mass}, radius{body_radius}, curr_pos{xpos, ypos, zpos}, curr_vel{xvel, yvel, zvel},
          rest_c{restitution} {
        check_mass();
        check_radius();
    }
    body(const body &o)
        : body_counter{}, mass_{o.mass_}, radius{o.radius}, curr_pos{o.curr_pos}, curr_vel{o.curr_vel}, acc{o.acc},
          pe{o.pe}, rest_c{o.rest_c} {}
    template <typename m, typename r, typename t, uint64_t rF>
    body(const body<m, r, t, rF> &o)
        : body_counter{}, mass_{(M)o.mass_}, radius{(R)o.radius}, curr_pos{o.curr_pos}, curr_vel{o.curr_vel},
          acc{o.acc}, pe{(T)o.pe}, rest_c{o.rest_c} {}
    body(body &&) noexcept = default;
    body &operator=(const body &o) {
        if (&o == this) return *this;
        mass_ = o.mass_; radius = o.radius; curr_pos = o.curr_pos; curr_vel = o.curr_vel;
        acc = o.acc; pe = o.pe; rest_c = o.rest_c;
        return *this;
    }
    body &operator=(body &&o) noexcept = default;

    const M &mass() const noexcept { return mass_; }
    const R &rad() const noexcept { return radius; }
    const vector3D<T> &pos() const noexcept { return curr_pos; }
    const vector3D<T> &vel() const noexcept { return curr_vel; }
    vector3D<T> &acceleration() noexcept { return acc; }
    const vector3D<T> &acceleration() const noexcept { return acc; }
    long double ke() const noexcept {
        const long double v = curr_vel.magnitude();
        return 0.5l * (long double)mass_ * v * v;  // set_ke (gregbod.hpp:917-920)
    }
    const T &potential_energy() const noexcept { return pe; }
    long double restitution() const noexcept { return rest_c; }
    long double volume() const { return (4 / 3.0l) * 3.14159265358979323846264338327950288l * radius * radius * radius; }
    long double density() const { return (long double)mass_ / volume(); }
    vector3D<T> momentum() const { return curr_vel * mass_; }
    bool set_mass(const M &new_mass) {
        if (new_mass < 0) return false;
        mass_ = new_mass;
        return true;
    }
    bool set_radius(const R &new_radius) {
        if (new_radius < 0) return false;
        radius = new_radius;
        return true;
    }
    void set_acc(const vector3D<T> &new_acc) { acc = new_acc; }
    bool set_pe(const T &new_pe) {
        if (new_pe > 0) return false;
        pe = new_pe;
        return true;
    }
    bool set_restitution(const long double &c) noexcept {
        if (c < 0 || c > 1) return false;
        rest_c = c;
        return true;
    }
    void update(const vector3D<T> &new_position, const vector3D<T> &new_velocity) noexcept {
        curr_pos = new_position;
        curr_vel = new_velocity;
    }
    void shift(const vector3D<T> &position_shift, const vector3D<T> &velocity_shift) noexcept {
        curr_pos += position_shift;
        curr_vel += velocity_shift;
    }
    void reset_acc() { acc.make_zero(); }
    /* history (bodies of a system with memFreq != 0, gregbod.hpp:476-491) */
    const vector3D<T> &prev_pos_at(typename std::vector<vector3D<T>>::size_type index) const {
        if (index >= positions.size())
            throw std::out_of_range("Requested position does not exist (index out of range).\n");
        return positions[index];
    }
    const vector3D<T> &prev_vel_at(typename std::vector<vector3D<T>>::size_type index) const {
        if (index >= velocities.size())
            throw std::out_of_range("Requested velocity does not exist (index out of range).\n");
        return velocities[index];
    }
    const long double &prev_ke_at(typename std::vector<long double>::size_type index) const {
        if (index >= energies.size())
            throw std::out_of_range("Requested kinetic energy does not exist (index out of range).\n");
        return energies[index];
    }
    typename std::vector<vector3D<T>>::size_type num_recorded() const noexcept { return positions.size(); }
    friend std::ostream &operator<<(std::ostream &os, const body &b) {
        return os << "[gtd::body@" << &b << ":id=" << b.id << ",m=" << +b.mass_ << ",r=" << +b.radius
                  << ",pos=" << b.curr_pos << ",vel=" << b.curr_vel << ",acc=" << b.acc << ']';
    }
};

template <typename M = long double, typename R = long double, typename T = long double, bool prog = false,
          bool mergeOverlappingBodies = false, int collisions = 0, uint64_t memFreq = 0, uint64_t fileFreq = 1,
          bool binaryFile = true>
class system {
public:
    using bod_t = body<M, R, T, memFreq>;
    using vec_size_t = typename std::vector<bod_t>::size_type;
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
    static constexpr int pred_coll_check = 7;
    static constexpr long double G_SI = 0.000000000066743L;
    static_assert(collisions == 0 || collisions == 3, "collisions = 0 or 3 (overlap check) on this engine");
    static_assert(std::is_arithmetic_v<M> && std::is_arithmetic_v<R> && std::is_arithmetic_v<T>,
                  "M, R and T are arithmetic types");

private:
    template <typename, typename, typename, uint64_t>
    friend class body_tracker;
    template <typename, typename, typename, bool, bool, int, uint64_t, uint64_t, bool>
    friend class system;
    nb_system *h_ = nullptr;
    mutable std::vector<bod_t> bods_;       // the bodies as the caller sees them, engine slot order
    mutable std::vector<int64_t> labels_;   // engine label of each slot
    mutable bool stale_ = false;            // the engine holds newer state than bods_
    uint64_t iterations_ = 1000;
    long double dt_ = 1;
    long double prev_dt_ = 0;
    bool evolved_ = false;
    bool mergers_ =
#ifdef GREGSYS_MERGERS
        true;
#else
        false;
#endif
    double min_tot_com_mom_ = 1e10;  // BILLION*10 (gregsys.hpp:209)
    // memFreq records (tot_pe / tot_ke / tot_e, gregsys.hpp:420-476)
    std::vector<T> tot_pe_, tot_ke_, tot_e_;
    mutable T ext_[3][2][2]{};
    mutable std::pair<T *, long double> ext_pairs_[3][2]{};

    void check(int rc) const { detail::raise(rc, nb_last_error(h_)); }
    static double d(long double v) noexcept { return (double)v; }

    // ---- host <-> engine ----
    void push(bool check_overlap) {
        const size_t n = bods_.size();
        std::vector<double> m(n), r(n), e(n), p(3 * n), v(3 * n), a(3 * n);
        for (size_t i = 0; i < n; ++i) {
            const bod_t &b = bods_[i];
            m[i] = d(b.mass_); r[i] = d(b.radius); e[i] = d(b.rest_c);
            p[3 * i] = d(b.curr_pos.x); p[3 * i + 1] = d(b.curr_pos.y); p[3 * i + 2] = d(b.curr_pos.z);
            v[3 * i] = d(b.curr_vel.x); v[3 * i + 1] = d(b.curr_vel.y); v[3 * i + 2] = d(b.curr_vel.z);
            a[3 * i] = d(b.acc.x); a[3 * i + 1] = d(b.acc.y); a[3 * i + 2] = d(b.acc.z);
        }
        const int rc = nb_set_bodies(h_, (int64_t)n, m.data(), r.data(), p.data(), v.data(), e.data(),
                                     check_overlap ? 1 : 0);
        if (rc != NB_OK) {
            // as the reference's constructor throws, no bodies are kept
            const std::string msg = nb_last_error(h_);
            bods_.clear();
            labels_.clear();
            stale_ = false;
            detail::raise(rc, msg.c_str());
        }
        if (n) check(nb_update_state(h_, (int64_t)n, nullptr, nullptr, nullptr, nullptr, a.data(), nullptr));
        labels_.resize(n);
        for (size_t i = 0; i < n; ++i) labels_[i] = (int64_t)i;
        stale_ = false;
    }
    // Engine state into bods_, keeping each surviving body object (and its id)
    // by engine label; bodies created by merges get new objects (new ids).
    void pull() const {
        int64_t n = 0;
        check(nb_num_bodies(h_, &n));
        std::vector<double> m(n), r(n), e(n), p(3 * n), v(3 * n), a(3 * n);
        std::vector<int64_t> lab(n);
        if (n) {
            check(nb_get_state(h_, n, m.data(), r.data(), p.data(), v.data(), a.data(), e.data()));
            check(nb_body_ids(h_, n, lab.data()));
        }
        std::vector<bod_t> out;
        out.reserve((size_t)n);
        const std::unordered_map<int64_t, size_t> where = slot_map(labels_, lab);
        for (int64_t i = 0; i < n; ++i) {
            const size_t at = find_slot(labels_, where, lab[i], (size_t)i);
            if (at != (size_t)-1) out.emplace_back(std::move(bods_[at]));
            else out.emplace_back();
            bod_t &b = out.back();
            b.mass_ = (M)m[i]; b.radius = (R)r[i]; b.rest_c = (long double)e[i];
            b.curr_pos = {(T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2]};
            b.curr_vel = {(T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]};
            b.acc = {(T)a[3 * i], (T)a[3 * i + 1], (T)a[3 * i + 2]};
        }
        bods_ = std::move(out);
        labels_ = std::move(lab);
        stale_ = false;
    }
    void refresh() const {
        if (stale_) pull();
    }
    // Slot of engine label `lab` among `labels` (-1 if absent): the same slot
    // in the common case, else through a map built only when slots moved.
    static std::unordered_map<int64_t, size_t> slot_map(const std::vector<int64_t> &labels,
                                                        const std::vector<int64_t> &wanted) {
        std::unordered_map<int64_t, size_t> m;
        bool same = labels.size() >= wanted.size();
        for (size_t i = 0; same && i < wanted.size(); ++i) same = labels[i] == wanted[i];
        if (!same)
            for (size_t i = 0; i < labels.size(); ++i) m.emplace(labels[i], i);
        return m;
    }
    static size_t find_slot(const std::vector<int64_t> &labels, const std::unordered_map<int64_t, size_t> &m,
                            int64_t lab, size_t guess) {
        if (guess < labels.size() && labels[guess] == lab) return guess;
        auto it = m.find(lab);
        return it == m.end() ? (size_t)-1 : it->second;
    }
    void apply_params() {
        if constexpr (collisions == 3) check(nb_set_collisions(h_, NB_OVERLAP_COLL_CHECK, mergers_, min_tot_com_mom_));
    }
    void init_bodies(bool allow_overlapping_bodies) {
        if (dt_ <= 0) throw std::invalid_argument{"Time-step cannot be zero or negative.\n"};  // check_dt
        if constexpr (mergeOverlappingBodies) {
            push(false);
            if (!allow_overlapping_bodies) {
                int64_t removed = 0;
                check(nb_merge_overlapping(h_, &removed));
                if (removed) pull();
            }
        } else {
            push(!allow_overlapping_bodies);
        }
    }

    // ---- callbacks: the engine's step hooks ----
    struct HookCtx {
        const system *self;
        std::function<void(std::vector<bod_t> &)> all;
        std::function<void(bod_t &)> each;
        std::exception_ptr err;
    };
    struct Snap {
        std::vector<double> m, r, e, p, v, a;
    };
    static void snap(const std::vector<bod_t> &b, Snap &s) {
        const size_t n = b.size();
        s.m.resize(n); s.r.resize(n); s.e.resize(n); s.p.resize(3 * n); s.v.resize(3 * n); s.a.resize(3 * n);
        for (size_t i = 0; i < n; ++i) {
            s.m[i] = d(b[i].mass_); s.r[i] = d(b[i].radius); s.e[i] = d(b[i].rest_c);
            const vector3D<T> *vs[3] = {&b[i].curr_pos, &b[i].curr_vel, &b[i].acc};
            std::vector<double> *ds[3] = {&s.p, &s.v, &s.a};
            for (int q = 0; q < 3; ++q) {
                (*ds[q])[3 * i] = d(vs[q]->x);
                (*ds[q])[3 * i + 1] = d(vs[q]->y);
                (*ds[q])[3 * i + 2] = d(vs[q]->z);
            }
        }
    }
    static int hook(void *user, nb_system *, uint64_t, int where) {
        HookCtx *c = static_cast<HookCtx *>(user);
        try {
            const system &S = *c->self;
            S.pull();
            const size_t n = S.bods_.size();
            Snap before, after;
            snap(S.bods_, before);
            if (where == NB_HOOK_STEP) c->all(S.bods_);
            else for (bod_t &b : S.bods_) c->each(b);
            if (S.bods_.size() != n)
                throw nbody_error("for_all_bods changed the number of bodies, which the engine's step loop cannot "
                                  "follow\n");
            snap(S.bods_, after);
            if (before.m != after.m || before.r != after.r || before.e != after.e || before.p != after.p ||
                before.v != after.v || before.a != after.a)
                S.check(nb_update_state(S.h_, (int64_t)n, after.m.data(), after.r.data(), after.p.data(),
                                        after.v.data(), after.a.data(), after.e.data()));
        } catch (...) {
            c->err = std::current_exception();
            return 1;
        }
        return 0;
    }
    // Restores the iteration count and removes the hook whatever happens.
    struct EvolveGuard {
        system &s;
        bool restore_iters = false, hooked = false;
        ~EvolveGuard() {
            if (restore_iters) (void)nb_set_iterations(s.h_, s.iterations_);
            if (hooked) (void)nb_set_step_hook(s.h_, nullptr, nullptr, 0);
        }
    };

    // ---- prog: the reference's progress lines (print_progress :1417-1431,
    //      print_conclusion :1432-1440) ----
    void print_progress(uint64_t steps, bool with_total) const {
        if (with_total)
            std::printf("\033[36mIteration \033[34m%" PRIu64 "\033[31m/\033[35m%" PRIu64 "\r", steps, iterations_);
        else
            std::printf("\033[36mIteration \033[34m%" PRIu64 "\r", steps);
        std::fflush(stdout);
    }
    static const char *method_str(int method) {
        if ((method & barnes_hut) == barnes_hut) return "Leapfrog KDK and Barnes-Hut";
        if ((method & euler) == euler) return "Euler";
        if ((method & modified_euler) == modified_euler) return "Modified Euler";
        if ((method & midpoint) == midpoint) return "Midpoint";
        if ((method & leapfrog_kdk) == leapfrog_kdk) return "Leapfrog KDK";
        if ((method & leapfrog_dkd) == leapfrog_dkd) return "Leapfrog DKD";
        return "";
    }
    static void print_conclusion(int method, std::chrono::nanoseconds total) {
        std::cout << "\033[0m\033[30m\n--------------------Done--------------------\n\033[0m\033[4m\033[32m"
                  << method_str(method) << "\033[0m\033[37m - time elapsed: \033[0m\033[1m\033[33m"
                  << total.count() / 1000000000 << "\033[0m\033[37m second"
                  << (total == std::chrono::seconds{1} ? "" : "s") << "\033[0m" << std::endl;
    }

    // ---- memFreq: records into the energy vectors and body histories ----
    void arm_recording(uint64_t expected_steps) {
        if constexpr (memFreq != 0) {
            tot_pe_.clear(); tot_ke_.clear(); tot_e_.clear();
            for (bod_t &b : bods_) {
                b.positions.clear();
                b.velocities.clear();
                b.energies.clear();
            }
            check(nb_set_recording(h_, memFreq, expected_steps / memFreq + 2, NB_RECORD_STATE | NB_RECORD_ENERGY));
        } else {
            (void)expected_steps;
        }
    }
    void collect_records() {
        if constexpr (memFreq != 0) {
            uint64_t taken = 0, stored = 0;
            check(nb_record_count(h_, &taken, &stored));
            for (uint64_t k = 0; k < stored; ++k) {
                int64_t step = 0, n = 0;
                double ke = 0, pe = 0;
                check(nb_get_record(h_, k, &step, &n, nullptr, nullptr, nullptr, &ke, &pe));
                std::vector<double> p(3 * n), v(3 * n);
                std::vector<int64_t> ids(n);
                check(nb_get_record(h_, k, nullptr, &n, p.data(), v.data(), ids.data(), nullptr, nullptr));
                tot_ke_.push_back((T)ke);
                tot_pe_.push_back((T)pe);
                tot_e_.push_back((T)(ke + pe));
                const std::unordered_map<int64_t, size_t> where = slot_map(labels_, ids);
                for (int64_t i = 0; i < n; ++i) {
                    const size_t at = find_slot(labels_, where, ids[i], (size_t)i);
                    if (at == (size_t)-1) continue;  // merged away since
                    bod_t &b = bods_[at];
                    b.positions.push_back({(T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2]});
                    b.velocities.push_back({(T)v[3 * i], (T)v[3 * i + 1], (T)v[3 * i + 2]});
                    const long double vm = b.velocities.back().magnitude();
                    b.energies.push_back(0.5l * (long double)b.mass_ * vm * vm);
                }
            }
        }
    }
    std::pair<T *, long double> *extrema(int which, const std::vector<T> &vals, const char *name) const {
        refresh();
        if (bods_.empty())
            throw empty_system_error((std::string(name) + "() cannot be called on empty system objects (no bodies "
                                      "present).\n").c_str());
        if (!evolved_ || vals.empty())
            throw no_evolution_error((std::string(name) + "() can only be called after a system object has been "
                                      "evolved.\n").c_str());
        const T first = vals[0];
        T mn = first, mx = first;
        uint64_t imn = 0, imx = 0, k = 0;
        for (const T &x : vals) {
            if (x < mn) { mn = x; imn = k; }
            if (x > mx) { mx = x; imx = k; }
            ++k;
        }
        ext_[which][0][0] = mn; ext_[which][0][1] = first - mn;
        ext_[which][1][0] = mx; ext_[which][1][1] = mx - first;
        ext_pairs_[which][0] = {ext_[which][0], imn * prev_dt_ * memFreq};
        ext_pairs_[which][1] = {ext_[which][1], imx * prev_dt_ * memFreq};
        return ext_pairs_[which];
    }

public:
    explicit system(int device = 0) {
        detail::raise(nb_create(device, &h_), nb_last_error(nullptr));
        check(nb_set_iterations(h_, iterations_));
        check(nb_set_timestep(h_, d(dt_)));
        apply_params();
    }
    system(std::initializer_list<bod_t> list) : system(0) {
        bods_.assign(list.begin(), list.end());
        init_bodies(false);
    }
    explicit system(long double timestep, uint64_t num_iterations, const char *units_format) : system(0) {
        dt_ = timestep;
        if (dt_ <= 0) throw std::invalid_argument{"Time-step cannot be zero or negative.\n"};
        check(nb_set_timestep(h_, d(dt_)));
        set_iterations(num_iterations);
        set_G_units(units_format);
    }
    template <uint64_t mF>
    system(const std::vector<body<M, R, T, mF>> &bodies, long double timestep = 1, uint64_t num_iterations = 1000,
           const char *units_format = "M1:1,D1:1,T1:1", bool allow_overlapping_bodies = false, int device = 0)
        : system(device) {
        dt_ = timestep;
        if (dt_ <= 0) throw std::invalid_argument{"Time-step cannot be zero or negative.\n"};
        check(nb_set_timestep(h_, d(dt_)));
        set_iterations(num_iterations);
        bods_.reserve(bodies.size());
        for (const auto &b : bodies) bods_.emplace_back(b);
        set_G_units(units_format);
        init_bodies(allow_overlapping_bodies);
    }
    /* .nsys constructor (gregsys.hpp:1479-1490).  The stored types must be
     * this system's (floating-ness and size of M, R, T), checked in
     * load_from_nsys's order and words (NSYS_TYPE_CHECK :989-1046). */
    explicit system(const char *nsys_path, bool alloc_chunks = true, bool check_nan_inf = true,
                    bool allow_overlapping_bodies = true, int device = 0)
        : system(device) {
        (void)alloc_chunks;  // the engine reads the chunks in one pass either way
        if (nsys_path == nullptr) throw std::invalid_argument{"nullptr passed as .nsys file path.\n"};
        nb_nsys_info info{};
        const int rc = nb_nsys_read(nsys_path, check_nan_inf, &info, 0, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc != NB_OK && rc != NB_ERR_CAPACITY) detail::raise(rc, nb_last_error(nullptr));
        type_check<M>(info.d_types & 0b10, info.m_size, "M", "mass");
        type_check<R>(info.d_types & 0b100, info.r_size, "R", "radius");
        type_check<T>(info.d_types & 0b1000, info.t_size, "T", "position and velocity");
        // without allow_overlapping_bodies the load runs check_overlap: the
        // engine's overlap test, or its merging loop for mergeOverlappingBodies
        check(nb_load_nsys(h_, nsys_path, check_nan_inf, allow_overlapping_bodies || mergeOverlappingBodies));
        if (mergeOverlappingBodies && !allow_overlapping_bodies) check(nb_merge_overlapping(h_, nullptr));
        check(nb_set_iterations(h_, iterations_));
        dt_ = info.dt;
        labels_.clear();
        bods_.clear();
        pull();
    }
    system(const system &) = delete;
    system &operator=(const system &) = delete;
    ~system() {
        bods_.clear();
        nb_destroy(h_);
    }

    /* engine extensions */
    void set_precision(int precision) { check(nb_set_precision(h_, precision)); }
    void set_mergers(bool on, double min_tot_com_mom = 1e10) {
        mergers_ = on;
        min_tot_com_mom_ = min_tot_com_mom;
        check(nb_set_collisions(h_, collisions, on ? 1 : 0, min_tot_com_mom));
    }
    nb_system *handle() noexcept { return h_; }
    std::pair<double, double> kinetic_potential() const {
        double ke, pe;
        check(nb_energy(h_, &ke, &pe));
        return {ke, pe};
    }

    /* parameters (gregsys.hpp:1627-1680) */
    vec_size_t num_bodies() const {
        int64_t n = 0;
        check(nb_num_bodies(h_, &n));
        return (vec_size_t)n;
    }
    long double elapsed_time() const {
        double t = 0;
        check(nb_time_elapsed(h_, &t));
        return t;
    }
    long double G_val() const {
        double g;
        check(nb_get_G(h_, &g));
        return g;
    }
    bool set_iterations(uint64_t number) noexcept {
        if (nb_set_iterations(h_, number) != NB_OK) return false;
        iterations_ = number;
        return true;
    }
    bool set_timestep(const long double &delta_t) noexcept {
        if (nb_set_timestep(h_, d(delta_t)) != NB_OK) return false;
        dt_ = delta_t;
        return true;
    }
    void set_G_units(const char *units_format) { check(nb_set_units(h_, units_format)); }
    bool set_softening(long double eps) noexcept { return nb_set_softening(h_, d(eps)) == NB_OK; }
    bool set_bh_opening_angle(long double theta) { return nb_set_bh_opening_angle(h_, d(theta)) == NB_OK; }
    template <int c = collisions, typename = std::enable_if_t<c != 0>>
    bool set_min_tot_com_mom(const long double &total_momentum) {
        if (total_momentum < 0) return false;
        min_tot_com_mom_ = d(total_momentum);
        return nb_set_collisions(h_, collisions, mergers_, min_tot_com_mom_) == NB_OK;
    }

    /* bodies (gregsys.hpp:1681-1760) */
    system &add_body(const bod_t &bod, bool chk_overlap = true) {
        refresh();
        bods_.emplace_back(bod);
        init_bodies(!chk_overlap);
        return *this;
    }
    template <typename... Args>
    system &emplace_body(bool chk_overlap, Args &&...args) {
        refresh();
        bods_.emplace_back(std::forward<Args>(args)...);
        init_bodies(!chk_overlap);
        return *this;
    }
    system &add_bodies(const std::vector<bod_t> &extra, bool chk_overlap = true) {
        if (extra.empty()) return *this;
        refresh();
        for (const auto &b : extra) bods_.emplace_back(b);
        init_bodies(!chk_overlap);
        return *this;
    }
    template <bool p2, bool m2, int c2, uint64_t f2, bool b2>
    system &add_system(const system<M, R, T, p2, m2, c2, memFreq, f2, b2> &other, bool chk_overlap = true) {
        other.refresh();
        return add_bodies(other.bods_, chk_overlap);
    }
    const bod_t &back() const {
        refresh();
        if (!bods_.empty()) return bods_.back();
        throw std::logic_error{"gtd::system<>::back cannot be called on an empty gtd::system object.\n"};
    }
    const bod_t &front() const {
        refresh();
        if (!bods_.empty()) return bods_.front();
        throw std::logic_error{"gtd::system<>::front cannot be called on an empty gtd::system object.\n"};
    }
    const bod_t &get_body(uint64_t id) const {
        refresh();
        for (const bod_t &b : bods_)
            if (b.get_id() == id) return b;
        throw std::invalid_argument("The id passed does not correspond to any body present in this system object.\n");
    }
    bool remove_body(uint64_t id) {
        refresh();
        for (size_t i = 0; i < bods_.size(); ++i)
            if (bods_[i].get_id() == id) {
                bods_.erase(bods_.begin() + (std::ptrdiff_t)i);
                push(false);
                return true;
            }
        return false;
    }

    /* evolve(int integration_method, continue_ev_if, for_each_bod, for_all_bods)
     * (gregsys.hpp:2593-3070) */
    template <typename evFuncT = std::nullptr_t, typename bodFuncT = std::nullptr_t,
              typename bodsFuncT = std::nullptr_t>
    const std::chrono::nanoseconds evolve(int integration_method = leapfrog_kdk,
                                          const evFuncT &continue_ev_if = nullptr,
                                          const bodFuncT &for_each_bod = nullptr,
                                          const bodsFuncT &for_all_bods = nullptr) {
        constexpr bool has_pred = !std::is_same_v<evFuncT, std::nullptr_t>;
        constexpr bool has_each = !std::is_same_v<bodFuncT, std::nullptr_t>;
        constexpr bool has_all = !std::is_same_v<bodsFuncT, std::nullptr_t>;
        if constexpr (has_pred) static_assert(std::is_invocable_r_v<bool, evFuncT>, "continue_ev_if: bool()");
        if constexpr (has_each) static_assert(std::is_invocable_v<bodFuncT, bod_t &>, "for_each_bod: void(body&)");
        if constexpr (has_all)
            static_assert(std::is_invocable_v<bodsFuncT, std::vector<bod_t> &>, "for_all_bods: void(vector&)");
        refresh();
        if (bods_.empty() || !detail::check_option(integration_method)) return std::chrono::nanoseconds::zero();
        const bool bh = (integration_method & barnes_hut) == barnes_hut;
        const auto t0 = std::chrono::high_resolution_clock::now();
        EvolveGuard guard{*this};
        HookCtx ctx{this, {}, {}, nullptr};
        if constexpr (has_each || has_all) {
            if (!bh) {  // the reference's BH branch runs neither callback
                int mask = 0;
                if constexpr (has_all) {
                    ctx.all = [&](std::vector<bod_t> &v) { for_all_bods(v); };
                    mask |= NB_HOOK_STEP;
                }
                if constexpr (has_each) {
                    ctx.each = [&](bod_t &b) { for_each_bod(b); };
                    mask |= NB_HOOK_BODIES;
                }
                check(nb_set_step_hook(h_, &system::hook, &ctx, mask));
                guard.hooked = true;
            }
        }
        auto rethrow = [&](int rc) {
            if (rc == NB_OK) return;
            stale_ = true;
            if (ctx.err) std::rethrow_exception(ctx.err);
            check(rc);
        };
        if constexpr (prog) std::cout << "\033[1m";
        if (!has_pred || bh) {  // the reference's BH branch steps `iterations` times whatever the predicate
            arm_recording(iterations_);
            if constexpr (prog) {
                // one step, then the rest in slices, each followed by the progress line
                guard.restore_iters = true;
                rethrow(nb_set_iterations(h_, 1));
                rethrow(nb_evolve(h_, integration_method, nullptr));
                uint64_t done = 1;
                print_progress(done, true);
                const uint64_t slice = std::max<uint64_t>(1, iterations_ / 100);
                while (done < iterations_) {
                    const uint64_t k = std::min<uint64_t>(slice, iterations_ - done);
                    rethrow(nb_advance(h_, k, nullptr));
                    done += k;
                    print_progress(done, true);
                }
            } else {
                rethrow(nb_evolve(h_, integration_method, nullptr));
            }
        } else if constexpr (has_pred) {
            arm_recording(4096 * (memFreq ? memFreq : 1));
            uint64_t steps = 0;
            if (continue_ev_if()) {
                guard.restore_iters = true;
                rethrow(nb_set_iterations(h_, 1));
                rethrow(nb_evolve(h_, integration_method, nullptr));  // a(x0) and the first step
                ++steps;
                if constexpr (prog) print_progress(steps, false);
                while (continue_ev_if()) {
                    rethrow(nb_advance(h_, 1, nullptr));
                    ++steps;
                    if constexpr (prog) print_progress(steps, false);
                }
            }
        }
        const auto total =
            std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::high_resolution_clock::now() - t0);
        stale_ = true;
        evolved_ = true;
        prev_dt_ = dt_;
        if constexpr (memFreq != 0) {
            pull();
            collect_records();
        }
        if constexpr (prog) print_conclusion(integration_method, total);
        return total;
    }

    /* memFreq accessors (gregsys.hpp:3072-3180): pairs of (value, difference
     * from the first record) for the minimum and the maximum, with the time of
     * each; the arrays live in the system until the next call. */
    template <uint64_t f = memFreq, typename = std::enable_if_t<f != 0>>
    std::pair<T *, long double> *pe_extrema() const { return extrema(0, tot_pe_, "pe_extrema"); }
    template <uint64_t f = memFreq, typename = std::enable_if_t<f != 0>>
    std::pair<T *, long double> *ke_extrema() const { return extrema(1, tot_ke_, "ke_extrema"); }
    template <uint64_t f = memFreq, typename = std::enable_if_t<f != 0>>
    std::pair<T *, long double> *tot_e_extrema() const { return extrema(2, tot_e_, "tot_e_extrema"); }
    template <uint64_t f = memFreq, typename = std::enable_if_t<f != 0>>
    const std::vector<T> &total_pe() const noexcept { return tot_pe_; }
    template <uint64_t f = memFreq, typename = std::enable_if_t<f != 0>>
    const std::vector<T> &total_ke() const noexcept { return tot_ke_; }
    template <uint64_t f = memFreq, typename = std::enable_if_t<f != 0>>
    const std::vector<T> &total_e() const noexcept { return tot_e_; }

    /* to_nsys (gregsys.hpp:3214-3258): the file size, or 0 if it could not be
     * written (the engine writes 8- and 16-byte floating values). */
    std::streamoff to_nsys(const char *path) const {
        if (path == nullptr) return 0;
        if (!std::is_floating_point_v<T> || (sizeof(T) != 8 && sizeof(T) != 16)) return 0;
        int64_t bytes = 0;
        if (nb_save_nsys(h_, path, (int)sizeof(T), &bytes) != NB_OK) return 0;
        return (std::streamoff)bytes;
    }

    const bod_t &operator[](vec_size_t index) const {
        refresh();
        if (index >= bods_.size())
            throw std::out_of_range("The requested body does not exist (index out of range).\n");
        return bods_[index];
    }
    auto begin() const {
        refresh();
        return bods_.cbegin();
    }
    auto end() const {
        refresh();
        return bods_.cend();
    }

private:
    template <typename X>
    static void type_check(int is_fp, int size, const char *type, const char *what) {
        char msg[256];
        if (std::is_floating_point_v<X> && !is_fp) {
            std::snprintf(msg, sizeof msg,
                          "Invalid .nsys format: d_types field states that %s values are stored in integral format, "
                          "but template parameter %s is floating-point.\n",
                          what, type);
            throw nsys_load_error(msg);
        }
        if (!std::is_floating_point_v<X> && is_fp) {
            std::snprintf(msg, sizeof msg,
                          "Invalid .nsys format: d_types field states that %s values are stored in floating-point "
                          "format, but template parameter %s is integral.\n",
                          what, type);
            throw nsys_load_error(msg);
        }
        if (size != (int)sizeof(X)) {
            std::snprintf(msg, sizeof msg,
                          "Size mismatch error: size of %s template parameter (%zu bytes) does not match its reported "
                          "size (%d bytes) in the .nsys file.\n",
                          type, sizeof(X), size);
            throw nsys_load_error(msg);
        }
    }
};

/* body_tracker (gregsys.hpp:3492-3611): the bodies of a system that a
 * predicate picks at construction, with the tracker's reductions (mass, com,
 * com_vel, mean_dist_to, mean_sep) computed on the GPU by nb_track.  As in
 * the reference, it assumes that no merger happens and no body is added to the
 * system after construction. */
template <typename M = long double, typename R = long double, typename T = long double, uint64_t rF = 0>
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
        for (const auto &b : sys) {
            if (pred(b)) idx_.push_back(k);
            ++k;
        }
    }
    uint64_t num_bods() const noexcept { return idx_.size(); }
    M mass() const {
        double m = 0;
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, &m, nullptr, nullptr, nullptr, nullptr));
        return (M)m;
    }
    vector3D<T> com() const {
        double c[3] = {0, 0, 0};
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, nullptr, c, nullptr, nullptr, nullptr));
        return {(T)c[0], (T)c[1], (T)c[2]};
    }
    vector3D<T> com_vel() const {
        double c[3] = {0, 0, 0};
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, nullptr, nullptr, c, nullptr, nullptr));
        return {(T)c[0], (T)c[1], (T)c[2]};
    }
    T mean_dist_to(const vector3D<T> &point) const {
        const double p[3] = {(double)point.x, (double)point.y, (double)point.z};
        double dd = 0;
        check(nb_track(h_, ids(), (int64_t)idx_.size(), p, nullptr, nullptr, nullptr, &dd, nullptr));
        return (T)dd;
    }
    T mean_sep() const {
        if (idx_.empty()) throw nbody_error{"Error: no bodies being tracked.\n"};
        double dd = 0;
        check(nb_track(h_, ids(), (int64_t)idx_.size(), nullptr, nullptr, nullptr, nullptr, nullptr, &dd));
        return (T)dd;
    }
};

}  // namespace gtd

#endif  // GTD_NBODY_HPP
