/*
 * ref_harness.cpp -- drives the compiled REFERENCE (GregW10/GregPackagesCPP
 * headers, included from /root/reference by oracle/Makefile) to produce golden
 * vectors for the oracle.  TEST INFRASTRUCTURE ONLY: built into oracle/_ref/
 * (git-ignored) and run only by tests/golden/make_fixtures.py and the CPU
 * tests.  No reference source is copied here; the headers are included from
 * where they lie.
 *
 * Usage:  ref_harness <mode> <in.bin> <out.bin>
 *   mode kdk  : gtd::system<ld,ld,ld,false,false,0,0,0,false>::evolve(leapfrog_kdk)
 *               (gregsys.hpp:2946-2990); output = final pos, vel (long double).
 *   mode euler / modeuler / midpoint / dkd : the same with evolve(euler),
 *               evolve(modified_euler), evolve(midpoint), evolve(leapfrog_dkd)
 *               (:2666-2891, :2993-3039).
 *   mode coll, coll_euler, coll_modeuler, coll_midpoint, coll_dkd :
 *               system<ld,ld,ld,false,false,3,0,0,false> (built with and without
 *               -DGREGSYS_MERGERS); evolve(leapfrog_kdk, nullptr, per_body,
 *               per_step) dumps every body after each step's s_coll.
 *   mode track : body_tracker<ld,ld,ld,0> (gregsys.hpp:3492-3611) over the bodies
 *               with x > 0: mass, com, com_vel, mean_dist_to(point), mean_sep.
 *   mode tree : bh_tree<double,double,double,0>::add_bodies (greg8tree.hpp:1066)
 *               and its operator<< dump (:975-1001) parsed to DFS pre-order
 *               records (depth, octant, body index, mass).
 *   mode nsys_ld / nsys_d : system<T,T,T> (T = long double / double) built
 *               from in.bin with units "M3:2,D5:7,T11:13", evolved `steps`
 *               KDK steps when steps > 0, then to_nsys(out) (gregsys.hpp:3214).
 *   mode nsys_copy / nsys_copy_d : the .nsys constructor (gregsys.hpp:1479)
 *               on the file given in place of in.bin, then to_nsys(out): the
 *               reference reads a snapshot and writes it back.
 * Input (little endian): int64 n, int64 steps, double dt, then n records of
 * 9 doubles (m, r, e, x, y, z, vx, vy, vz).
 */
#include "gregsys.hpp"
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

using ld = long double;

struct In {
    int64_t n = 0, steps = 0;
    double dt = 1;
    std::vector<double> rec;
};

static bool read_in(const char *path, In &in) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    bool ok = fread(&in.n, 8, 1, f) == 1 && fread(&in.steps, 8, 1, f) == 1 && fread(&in.dt, 8, 1, f) == 1;
    if (ok) {
        in.rec.resize((size_t)in.n * 9);
        ok = fread(in.rec.data(), 8, in.rec.size(), f) == in.rec.size();
    }
    fclose(f);
    return ok;
}

static int run_kdk(const In &in, FILE *out, int method) {
    using bod_t = gtd::body<ld, ld, ld, 0>;
    std::vector<bod_t> bods;
    bods.reserve((size_t)in.n);
    for (int64_t i = 0; i < in.n; ++i) {
        const double *r = &in.rec[(size_t)i * 9];
        bods.emplace_back((ld)r[0], (ld)r[1], gtd::vector3D<ld>{(ld)r[3], (ld)r[4], (ld)r[5]},
                          gtd::vector3D<ld>{(ld)r[6], (ld)r[7], (ld)r[8]}, (ld)r[2]);
    }
    gtd::system<ld, ld, ld, false, false, 0, 0, 0, false> sys{bods, (ld)in.dt, (uint64_t)in.steps,
                                                              "M1:1,D1:1,T1:1", true};
    /* evolve's return value is the reference's own wall-clock timer of the step loop (:2604, :3061-3065) */
    auto ns = sys.evolve(method);
    printf("evolve_ns %lld\n", (long long)ns.count());
    for (const auto &b : sys) {
        ld v[6] = {b.pos().get_x(), b.pos().get_y(), b.pos().get_z(), b.vel().get_x(), b.vel().get_y(), b.vel().get_z()};
        fwrite(v, sizeof(ld), 6, out);
    }
    return 0;
}

static int run_coll(const In &in, FILE *out, int method) {
    using bod_t = gtd::body<ld, ld, ld, 0>;
    using sys_t = gtd::system<ld, ld, ld, false, false, 3, 0, 0, false>;
    std::vector<bod_t> bods;
    bods.reserve((size_t)in.n);
    for (int64_t i = 0; i < in.n; ++i) {
        const double *r = &in.rec[(size_t)i * 9];
        bods.emplace_back((ld)r[0], (ld)r[1], gtd::vector3D<ld>{(ld)r[3], (ld)r[4], (ld)r[5]},
                          gtd::vector3D<ld>{(ld)r[6], (ld)r[7], (ld)r[8]}, (ld)r[2]);
    }
    sys_t sys{bods, (ld)in.dt, (uint64_t)in.steps, "M1:1,D1:1,T1:1", true};
    /* initial slot ids, so that the reader can map ids to input indices */
    int64_t n0 = (int64_t)sys.num_bodies();
    fwrite(&n0, 8, 1, out);
    for (const auto &b : sys) {
        int64_t id = (int64_t)b.get_id();
        fwrite(&id, 8, 1, out);
    }
    int64_t step = 0;
    auto per_step = [&](std::vector<bod_t> &) { ++step; };
    auto per_body = [&](bod_t &b) {
        int64_t hdr[2] = {step, (int64_t)b.get_id()};
        fwrite(hdr, 8, 2, out);
        ld v[8] = {b.mass(), b.rad(), b.pos().get_x(), b.pos().get_y(), b.pos().get_z(), b.vel().get_x(), b.vel().get_y(), b.vel().get_z()};
        fwrite(v, sizeof(ld), 8, out);
    };
    sys.evolve(method, nullptr, per_body, per_step);
    return 0;
}

/* body_tracker (gregsys.hpp:3492-3611) over the bodies with x > 0: count, then
 * mass(), com(), com_vel(), mean_dist_to(1.5e8, -2.5e8, 0.5e8), mean_sep() in
 * long double. */
static int run_track(const In &in, FILE *out) {
    using bod_t = gtd::body<ld, ld, ld, 0>;
    std::vector<bod_t> bods;
    bods.reserve((size_t)in.n);
    for (int64_t i = 0; i < in.n; ++i) {
        const double *r = &in.rec[(size_t)i * 9];
        bods.emplace_back((ld)r[0], (ld)r[1], gtd::vector3D<ld>{(ld)r[3], (ld)r[4], (ld)r[5]},
                          gtd::vector3D<ld>{(ld)r[6], (ld)r[7], (ld)r[8]}, (ld)r[2]);
    }
    gtd::system<ld, ld, ld, false, false, 0, 0, 0, false> sys{bods, (ld)in.dt, (uint64_t)1, "M1:1,D1:1,T1:1", true};
    gtd::body_tracker<ld, ld, ld, 0> tr{sys, [](const bod_t &b) { return b.pos().get_x() > 0; }};
    int64_t cnt = (int64_t)tr.num_bods();
    fwrite(&cnt, 8, 1, out);
    auto com = tr.com();
    auto cv = tr.com_vel();
    gtd::vector3D<ld> point{(ld)1.5e8, (ld)-2.5e8, (ld)0.5e8};
    ld v[9] = {tr.mass(), com.get_x(), com.get_y(), com.get_z(), cv.get_x(), cv.get_y(), cv.get_z(),
               tr.mean_dist_to(point), tr.mean_sep()};
    fwrite(v, sizeof(ld), 9, out);
    return 0;
}

/* Recursive-descent reader of the bh_cube operator<< text (greg8tree.hpp:975-1001). */
struct Dump {
    const std::string &s;
    size_t p = 0;
    std::vector<int64_t> rec; /* depth, octant, body id */
    std::vector<double> mass;
    bool expect(const char *lit) {
        size_t l = strlen(lit);
        if (s.compare(p, l, lit) != 0) return false;
        p += l;
        return true;
    }
    bool cube(int depth, int octant) {
        if (!expect("[gtd::bh_cube@")) return false;
        p = s.find(":mass=", p);
        if (p == std::string::npos) return false;
        p += 6;
        size_t e = s.find(",com=", p);
        double m = strtod(s.c_str() + p, nullptr);
        p = s.find(",body=", e);
        p += 6;
        int64_t id = -1;
        if (!expect("NULL")) {
            size_t q = s.find(":id=", p);
            id = strtoll(s.c_str() + q + 4, nullptr, 10);
            q = s.find("current_ke=", q);
            p = s.find(']', q) + 1;
        }
        rec.push_back(depth);
        rec.push_back(octant);
        rec.push_back(id);
        mass.push_back(m);
        if (!expect(",cubes:\n")) return false;
        for (int i = 0; i < 8; ++i) {
            std::string lab = "cube_" + std::to_string(i) + "=";
            if (!expect(lab.c_str())) return false;
            if (!expect("NULL")) {
                if (!expect("\n")) return false;
                if (!cube(depth + 1, i)) return false;
            }
            if (!expect("\n")) return false;
        }
        return expect("]");
    }
};

static int run_tree(const In &in, FILE *out) {
    using bod_t = gtd::body<double, double, double, 0>;
    std::vector<bod_t> bods;
    bods.reserve((size_t)in.n);
    for (int64_t i = 0; i < in.n; ++i) {
        const double *r = &in.rec[(size_t)i * 9];
        bods.emplace_back(r[0], r[1], gtd::vector3D<double>{r[3], r[4], r[5]},
                          gtd::vector3D<double>{r[6], r[7], r[8]}, (ld)r[2]);
    }
    std::vector<int64_t> id_of(in.n);
    for (int64_t i = 0; i < in.n; ++i) id_of[i] = (int64_t)bods[i].get_id();
    gtd::bh_tree<double, double, double, 0> tree;
    tree.add_bodies(bods);
    std::ostringstream os;
    os << std::setprecision(17) << tree;
    std::string txt = os.str();
    Dump d{txt};
    if (!d.cube(0, -1)) {
        fprintf(stderr, "tree dump parse failed at %zu\n", d.p);
        return 3;
    }
    /* map body ids back to input indices */
    std::map<int64_t, int64_t> idx;
    for (int64_t i = 0; i < in.n; ++i) idx[id_of[i]] = i;
    int64_t cnt = (int64_t)d.mass.size();
    fwrite(&cnt, 8, 1, out);
    for (int64_t k = 0; k < cnt; ++k) {
        int64_t r3[3] = {d.rec[3 * k], d.rec[3 * k + 1], d.rec[3 * k + 2] < 0 ? -1 : idx[d.rec[3 * k + 2]]};
        fwrite(r3, 8, 3, out);
        fwrite(&d.mass[k], 8, 1, out);
    }
    return 0;
}

template <typename T>
static int run_nsys_write(const In &in, const char *out_path) {
    using bod_t = gtd::body<T, T, T, 0>;
    std::vector<bod_t> bods;
    bods.reserve((size_t)in.n);
    for (int64_t i = 0; i < in.n; ++i) {
        const double *r = &in.rec[(size_t)i * 9];
        bods.emplace_back((T)r[0], (T)r[1], gtd::vector3D<T>{(T)r[3], (T)r[4], (T)r[5]},
                          gtd::vector3D<T>{(T)r[6], (T)r[7], (T)r[8]}, (ld)r[2]);
    }
    gtd::system<T, T, T, false, false, 0, 0, 0, false> sys{bods, (ld)in.dt, (uint64_t)std::max<int64_t>(in.steps, 1),
                                                           "M3:2,D5:7,T11:13", true};
    if (in.steps > 0) sys.evolve(gtd::system<T, T, T, false, false, 0, 0, 0, false>::leapfrog_kdk);
    return sys.to_nsys(out_path) > 0 ? 0 : 3;
}

template <typename T>
static int run_nsys_copy(const char *in_path, const char *out_path) {
    gtd::system<T, T, T, false, false, 0, 0, 0, false> sys{in_path};
    return sys.to_nsys(out_path) > 0 ? 0 : 3;
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s <mode> in out (modes: see the file header)\n", argv[0]);
        return 2;
    }
    const std::string m0 = argv[1];
    if (m0 == "nsys_copy" || m0 == "nsys_copy_d") {
        try {
            return m0 == "nsys_copy" ? run_nsys_copy<ld>(argv[2], argv[3]) : run_nsys_copy<double>(argv[2], argv[3]);
        } catch (const std::exception &e) {
            fprintf(stderr, "reference threw: %s\n", e.what());
            return 4;
        }
    }
    In in;
    if (!read_in(argv[2], in)) {
        fprintf(stderr, "cannot read %s\n", argv[2]);
        return 2;
    }
    if (m0 == "nsys_ld" || m0 == "nsys_d") {
        try {
            return m0 == "nsys_ld" ? run_nsys_write<ld>(in, argv[3]) : run_nsys_write<double>(in, argv[3]);
        } catch (const std::exception &e) {
            fprintf(stderr, "reference threw: %s\n", e.what());
            return 4;
        }
    }
    FILE *out = fopen(argv[3], "wb");
    if (!out) return 2;
    int rc = 2;
    std::string mode = argv[1];
    try {
        using S = gtd::system<ld, ld, ld, false, false, 0, 0, 0, false>;
        const std::pair<const char *, int> methods[] = {{"kdk", S::leapfrog_kdk}, {"euler", S::euler},
                                                        {"modeuler", S::modified_euler}, {"midpoint", S::midpoint},
                                                        {"dkd", S::leapfrog_dkd}};
        for (const auto &mm : methods) {
            if (mode == mm.first) rc = run_kdk(in, out, mm.second);
            if (mode == std::string(mm.first == std::string("kdk") ? "coll" : "coll_") +
                           (mm.first == std::string("kdk") ? "" : mm.first))
                rc = run_coll(in, out, mm.second);
        }
        if (mode == "tree") rc = run_tree(in, out);
        if (mode == "track") rc = run_track(in, out);
    } catch (const std::exception &e) {
        fprintf(stderr, "reference threw: %s\n", e.what());
        rc = 4;
    }
    fclose(out);
    return rc;
}
