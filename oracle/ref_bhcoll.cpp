/*
 * ref_bhcoll.cpp -- runs the compiled REFERENCE's Barnes-Hut collision pass,
 * system::bh_s_coll (gtd/gregsys.hpp:631-686, with bh_cube::nn_iterator,
 * greg8tree.hpp:615-700), on one given state, for the oracle's fixtures.
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ (git-ignored) by
 * oracle/Makefile from the headers where they lie, and run only by
 * tests/golden/make_fixtures.py.
 *
 * bh_s_coll and the tree member are private and evolve reaches them only
 * after a gravity step, so this harness opens the class (access specifiers
 * become public; the standard headers are included first so that only the
 * reference's own classes are affected), builds the tree with
 * btree.add_bodies(bods) as evolve does (gregsys.hpp:2919), and calls the pass.
 *
 * Usage: ref_bhcoll <in.bin> <out.bin>
 * Input as ref_harness (int64 n, int64 steps (unused), double dt (unused),
 * n records of 9 doubles m, r, e, x, y, z, vx, vy, vz).
 * Output: n records of 3 long doubles, the velocities after the pass, in
 * input order.  Type: system<long double, long double, long double, false,
 * false, 3, 0, 0, false>.
 */
#include <algorithm>
#include <array>
#include <chrono>
#include <cinttypes>
#include <climits>
#include <cmath>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <pwd.h>
#include <random>
#include <regex>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <tuple>
#include <type_traits>
#include <unistd.h>
#include <utility>
#include <vector>

#define private public
#define protected public
#include "gregsys.hpp"
#undef private
#undef protected

using ld = long double;

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: ref_bhcoll <in.bin> <out.bin>\n");
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 3;
    int64_t n = 0, steps = 0;
    double dt = 1;
    bool ok = fread(&n, 8, 1, f) == 1 && fread(&steps, 8, 1, f) == 1 && fread(&dt, 8, 1, f) == 1;
    std::vector<double> rec((size_t)n * 9);
    ok = ok && fread(rec.data(), 8, rec.size(), f) == rec.size();
    fclose(f);
    if (!ok) return 3;
    using bod_t = gtd::body<ld, ld, ld, 0>;
    using sys_t = gtd::system<ld, ld, ld, false, false, 3, 0, 0, false>;
    std::vector<bod_t> bods;
    bods.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const double *r = &rec[(size_t)i * 9];
        bods.emplace_back((ld)r[0], (ld)r[1], gtd::vector3D<ld>{(ld)r[3], (ld)r[4], (ld)r[5]},
                          gtd::vector3D<ld>{(ld)r[6], (ld)r[7], (ld)r[8]}, (ld)r[2]);
    }
    try {
        sys_t sys{bods, (ld)dt, (uint64_t)1, "M1:1,D1:1,T1:1", true};
        sys.btree.delete_tree();
        sys.btree.add_bodies(sys.bods);
        sys.bh_s_coll();
        FILE *o = fopen(argv[2], "wb");
        if (!o) return 3;
        for (const auto &b : sys.bods) {
            ld v[3] = {b.vel().get_x(), b.vel().get_y(), b.vel().get_z()};
            fwrite(v, sizeof(ld), 3, o);
        }
        fclose(o);
    } catch (const std::exception &ex) {
        fprintf(stderr, "reference threw: %s\n", ex.what());
        return 4;
    }
    return 0;
}
