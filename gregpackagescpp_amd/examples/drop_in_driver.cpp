// drop_in_driver.cpp -- one reference-style driver that builds unchanged
// against the reference's gtd::system (-DUSE_REFERENCE_HEADERS, gregsys.hpp)
// and against this engine's facade (gtd/nbody.hpp).  It uses long double
// masses, radii and positions, the types of every reference driver.
//
//   drop_in_driver <mode> <in.bin> <out.bin>
//   in.bin: int64 n, int64 steps, double dt, then n records of 9 doubles
//           (m, r, restitution, x, y, z, vx, vy, vz)
//   modes:
//     cor       collisions = 3; evolve(leapfrog_kdk, nullptr, per_body,
//               per_step), where per_step rewrites every body's coefficient of
//               restitution from its distance to a running centre the way
//               system::random_comet's for_all_bods does (gregsys.hpp:1958-
//               1984) and per_body records every body after each step's
//               collision pass;
//     cor_pred  the same with a continue_ev_if predicate that stops after
//               `steps` steps (the random_comet calling pattern);
//     merge     mergeOverlappingBodies: the bodies after construction;
//     check     the construction overlap check: 1 and the overlapping pair's
//               slot order if it throws overlapping_bodies_error, else 0.
//   out.bin: doubles -- cor/cor_pred: per record (step, slot, m, r, x, y, z,
//   vx, vy, vz, restitution); merge: n, then per body (m, r, x, y, z, vx, vy,
//   vz, restitution); check: flag.
#ifdef USE_REFERENCE_HEADERS
#include "gregsys.hpp"
#else
#include "gtd/nbody.hpp"
#endif
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using ld = long double;
using bod_t = gtd::body<ld, ld, ld, 0>;

static bool read_in(const char *path, int64_t &n, int64_t &steps, double &dt, std::vector<double> &rec) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    bool ok = std::fread(&n, 8, 1, f) == 1 && std::fread(&steps, 8, 1, f) == 1 && std::fread(&dt, 8, 1, f) == 1;
    if (ok) {
        rec.resize((size_t)n * 9);
        ok = std::fread(rec.data(), 8, rec.size(), f) == rec.size();
    }
    std::fclose(f);
    return ok;
}

static std::vector<bod_t> make_bodies(int64_t n, const std::vector<double> &rec) {
    std::vector<bod_t> bods;
    bods.reserve((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const double *r = &rec[(size_t)i * 9];
        bods.emplace_back((ld)r[0], (ld)r[1], gtd::vector3D<ld>{(ld)r[3], (ld)r[4], (ld)r[5]},
                          gtd::vector3D<ld>{(ld)r[6], (ld)r[7], (ld)r[8]}, (ld)r[2]);
    }
    return bods;
}

static void put(FILE *out, const bod_t &b, double step, double slot) {
    const double v[11] = {step, slot, (double)b.mass(), (double)b.rad(), (double)b.pos().get_x(),
                          (double)b.pos().get_y(), (double)b.pos().get_z(), (double)b.vel().get_x(),
                          (double)b.vel().get_y(), (double)b.vel().get_z(), (double)b.restitution()};
    std::fwrite(v, 8, 11, out);
}

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <cor|cor_pred|merge|check> in.bin out.bin\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    int64_t n = 0, steps = 0;
    double dt = 1;
    std::vector<double> rec;
    if (!read_in(argv[2], n, steps, dt, rec)) return 2;
    FILE *out = std::fopen(argv[3], "wb");
    if (!out) return 2;
    try {
        if (mode == "cor" || mode == "cor_pred") {
            using sys_t = gtd::system<ld, ld, ld, false, false, 3, 0, 0, false>;
            sys_t sys{make_bodies(n, rec), (ld)dt, (uint64_t)steps, "M1:1,D1:1,T1:1", true};
            // random_comet's restitution profile: max_cor at the centre falling
            // towards min_cor with distance (gregsys.hpp:1903-1926, 1958-1984)
            const ld min_cor = 0.1l, max_cor = 0.9l, n_exp = 2.0l, n_ = (ld)n;
            gtd::vector3D<ld> com{};
            for (const auto &b : sys) com += b.pos();
            com /= n_;
            ld d_scale = 0;
            for (const auto &b : sys) d_scale = std::max(d_scale, (ld)(b.pos() - com).magnitude());
            const ld scale_factor = powl(d_scale / 2, n_exp);
            int64_t step = 0;
            auto per_step = [&](std::vector<bod_t> &bodies) {
                ++step;
                for (const bod_t &b : bodies) com += b.pos();
                com /= n_;
                for (bod_t &b : bodies) {
                    const ld dist_exp = powl((b.pos() - com).magnitude(), n_exp);
                    b.set_restitution(max_cor - (max_cor - min_cor) * (dist_exp / (dist_exp + scale_factor)));
                }
            };
            int64_t slot = 0, last_step = -1;
            auto per_body = [&](bod_t &b) {
                if (step != last_step) {
                    slot = 0;
                    last_step = step;
                }
                put(out, b, (double)step, (double)slot++);
            };
            if (mode == "cor") {
                sys.evolve(sys_t::leapfrog_kdk, nullptr, per_body, per_step);
            } else {
                int64_t polls = 0;
                sys.evolve(sys_t::leapfrog_kdk, [&]() { return polls++ < steps; }, per_body, per_step);
            }
        } else if (mode == "merge") {
            gtd::system<ld, ld, ld, false, true, 0, 0, 0, false> sys{make_bodies(n, rec), (ld)dt, (uint64_t)steps,
                                                                     "M1:1,D1:1,T1:1", false};
            const double cnt = (double)sys.num_bodies();
            std::fwrite(&cnt, 8, 1, out);
            for (const auto &b : sys) {
                const double v[9] = {(double)b.mass(), (double)b.rad(), (double)b.pos().get_x(),
                                     (double)b.pos().get_y(), (double)b.pos().get_z(), (double)b.vel().get_x(),
                                     (double)b.vel().get_y(), (double)b.vel().get_z(), (double)b.restitution()};
                std::fwrite(v, 8, 9, out);
            }
        } else if (mode == "check") {
            double flag = 0;
            try {
                gtd::system<ld, ld, ld, false, false, 0, 0, 0, false> sys{make_bodies(n, rec), (ld)dt,
                                                                          (uint64_t)steps, "M1:1,D1:1,T1:1", false};
            } catch (const gtd::overlapping_bodies_error &e) {
                flag = 1;
                std::fprintf(stderr, "%s", e.what());
            }
            std::fwrite(&flag, 8, 1, out);
        } else {
            std::fclose(out);
            return 2;
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "driver threw: %s\n", e.what());
        std::fclose(out);
        return 4;
    }
    std::fclose(out);
    return 0;
}
