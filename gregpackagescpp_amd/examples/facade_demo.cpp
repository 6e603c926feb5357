// facade_demo.cpp -- a reference-style driver (cf. sl9.cpp / ran.cpp) written
// against the gtd::system facade: a moon on a circular orbit around a
// Jupiter-mass body, evolved with leapfrog KDK, then a small Plummer sphere
// through the Barnes-Hut path.  Exits non-zero when a check fails.
#include <cmath>
#include <cstdio>
#include <vector>

#include "gtd/nbody.hpp"

int main() {
    using sys_t = gtd::system<double, double, double, false, false, 0, 0, 0, false>;
    const double M = 1.89813e27, R = 4.217e8, G = 6.6743e-11;  // Jupiter and Io's orbital radius
    const double v = std::sqrt(G * M / R);
    std::vector<gtd::body<double, double, double, 0>> bods;
    bods.emplace_back(M, 6.9911e7);
    bods.emplace_back(8.93e22, 1.8216e6, gtd::vector3D<double>{R, 0, 0}, gtd::vector3D<double>{0, v, 0});
    const double period = 2 * M_PI * std::sqrt(R * R * R / (G * M));
    const uint64_t steps = 20000;
    sys_t sys{bods, period / steps, steps, "M1:1,D1:1,T1:1", true};
    auto ns = sys.evolve(sys_t::leapfrog_kdk);
    const auto &moon = sys[1];
    const auto &jup = sys[0];
    double dx = moon.pos().get_x() - jup.pos().get_x(), dy = moon.pos().get_y() - jup.pos().get_y();
    double r = std::sqrt(dx * dx + dy * dy);
    std::printf("one orbit in %llu steps: %.3f ms, radius drift %.3e, end angle %.3e rad\n",
                (unsigned long long)steps, ns.count() / 1e6, r / R - 1, std::atan2(dy, dx));
    if (std::fabs(r / R - 1) > 1e-6 || std::fabs(std::atan2(dy, dx)) > 1e-3) return 1;

    // memFreq = 10: energies and body histories every 10 steps (the
    // reference's pe_extrema / prev_pos_at accessors), in long double
    {
        using ld = long double;
        std::vector<gtd::body<ld, ld, ld, 10>> lb;
        lb.emplace_back((ld)M, (ld)6.9911e7);
        lb.emplace_back((ld)8.93e22, (ld)1.8216e6, gtd::vector3D<ld>{(ld)R, 0, 0}, gtd::vector3D<ld>{0, (ld)v, 0});
        gtd::system<ld, ld, ld, false, false, 0, 10, 0, false> rec{lb, (ld)(period / steps), 200, "M1:1,D1:1,T1:1",
                                                                   true};
        rec.evolve(decltype(rec)::leapfrog_kdk);
        const auto *pe = rec.pe_extrema();
        const auto *te = rec.tot_e_extrema();
        std::printf("memFreq 10: %zu records, PE in [%.6Le, %.6Le], total energy spread %.3Le\n",
                    rec.total_pe().size(), pe[0].first[0], pe[1].first[0], te[1].first[0] - te[0].first[0]);
        if (rec.total_pe().size() != 21 || rec[1].num_recorded() != 21 || rec[1].prev_pos_at(0).get_x() != (ld)R ||
            !(pe[0].first[0] <= pe[1].first[0]) || std::fabs((double)((te[1].first[0] - te[0].first[0]) /
                                                                      te[0].first[0])) > 1e-6)
            return 8;
    }
    // prog: the reference's progress lines and closing summary
    {
        gtd::system<double, double, double, true, false, 0, 0, 0, false> p{bods, period / steps, 300,
                                                                           "M1:1,D1:1,T1:1", true};
        p.evolve(decltype(p)::leapfrog_kdk);
    }

    // continue_ev_if: stop after 100 polled steps
    int polls = 0;
    sys.evolve(sys_t::leapfrog_kdk, [&]() { return polls++ < 100; });
    std::printf("predicate-driven evolve ran %d steps\n", polls - 1);

    // Barnes-Hut on a Plummer sphere (seeded IC generator of the C ABI)
    const int64_t n = 20000;
    std::vector<double> m(n), p(3 * n), w(3 * n);
    if (nb_ic_plummer(7, n, 1.0, nullptr, nullptr, m.data(), p.data(), w.data()) != NB_OK) return 2;
    std::vector<gtd::body<double, double, double, 0>> pl;
    pl.reserve(n);
    for (int64_t i = 0; i < n; ++i)
        pl.emplace_back(m[i], 0.0, gtd::vector3D<double>{p[3 * i], p[3 * i + 1], p[3 * i + 2]},
                        gtd::vector3D<double>{w[3 * i], w[3 * i + 1], w[3 * i + 2]});
    sys_t bh{pl, 1e-3, 10, "M1:1,D1:1,T1:1", true};
    bh.set_G_units("M1:1,D1:1,T1:1");
    bh.set_softening(0.01);
    bh.set_bh_opening_angle(0.5);
    bh.set_precision(NB_PREC_MIXED);
    auto bns = bh.evolve(sys_t::leapfrog_kdk | sys_t::barnes_hut);
    std::printf("Barnes-Hut: %lld bodies, 10 steps in %.3f ms\n", (long long)bh.num_bodies(), bns.count() / 1e6);
    // body_tracker over the inner bodies (the reference's COM-of-a-subset helper)
    gtd::body_tracker<double, double, double, 0> core{bh, [](const auto &b) {
        return b.pos().get_x() * b.pos().get_x() + b.pos().get_y() * b.pos().get_y() < 0.25;
    }};
    const auto c = core.com();
    std::printf("tracker: %llu bodies, mass %.4f, com (%.2e, %.2e, %.2e), mean separation %.4f\n",
                (unsigned long long)core.num_bods(), core.mass(), c.x, c.y, c.z, core.mean_sep());
    if (!(core.num_bods() > 0 && core.mass() > 0 && core.mean_sep() > 0)) return 5;
    try {
        sys.set_timestep(-1);
        if (sys.set_timestep(-1)) return 3;  // setters answer false, as the reference's do
        std::vector<gtd::body<double, double, double, 0>> bad{{1.0, 1.0}, {1.0, 1.0}};
        sys_t overlapping{bad, 1.0, 1, "M1:1,D1:1,T1:1", false};
        return 4;
    } catch (const gtd::overlapping_bodies_error &e) {
        std::printf("overlapping bodies rejected: %s\n", e.what());
    }
    // .nsys snapshot: to_nsys, then the .nsys constructor (the sl9.cpp frame
    // pattern: write a frame, reload it later)
    const char *snap = "facade_demo_snapshot.nsys";
    const std::streamoff bytes = sys.to_nsys(snap);
    sys_t again{snap};
    std::remove(snap);
    if (bytes != 96 + 2 * 80 || again.num_bodies() != 2 || again.elapsed_time() != sys.elapsed_time()) return 5;
    for (int k = 0; k < 2; ++k)
        if (again[k].pos().get_x() != sys[k].pos().get_x() || again[k].vel().get_y() != sys[k].vel().get_y() ||
            again[k].mass() != sys[k].mass())
            return 6;
    std::printf(".nsys snapshot: %lld bytes, time elapsed %.6e s, reloaded identically\n", (long long)bytes,
                (double)again.elapsed_time());
    try {
        sys_t missing{"no_such_file.nsys"};
        return 7;
    } catch (const gtd::nsys_load_error &e) {
        std::printf("missing snapshot rejected: %s", e.what());
    }

    std::printf("facade demo OK\n");
    return 0;
}
