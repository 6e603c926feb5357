// nb_ics.cpp -- seeded, bit-reproducible initial conditions (host code).
//
// Random stream: SplitMix64 (Steele, Lea & Flood 2014; constants as published
// by Vigna), U = (u64 >> 11) * 2^-53.  The reference's own generators are
// unseeded (gregalg.hpp:392-404), so every seeded test input comes from here.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/gregnbody.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform_open() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }  // (0, 1)
    double gauss() {  // Box-Muller, first variate
        double u1 = uniform_open(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Isotropic unit vector: cos(theta) = 1 - 2U, phi = 2 pi U (gregsys.hpp:1818-1824)
void iso(SplitMix64 &g, double r, double *out) {
    double ct = 1.0 - 2.0 * g.uniform();
    double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
    double ph = 2.0 * M_PI * g.uniform();
    out[0] = r * st * std::cos(ph);
    out[1] = r * st * std::sin(ph);
    out[2] = r * ct;
}

void com_frame(int64_t n, const double *m, double *pos, double *vel) {
    double M = 0, cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        M += m[i];
        for (int k = 0; k < 3; ++k) {
            cx[k] += m[i] * pos[3 * i + k];
            cv[k] += m[i] * vel[3 * i + k];
        }
    }
    if (!(M > 0)) return;
    for (int k = 0; k < 3; ++k) {
        cx[k] /= M;
        cv[k] /= M;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] -= cx[k];
            vel[3 * i + k] -= cv[k];
        }
}

// Plummer sphere by Aarseth, Henon & Wielen (1974): r from the inverted mass
// profile (X <= 0.999), speed by von Neumann rejection on g(q) = q^2 (1-q^2)^3.5.
void plummer_body(SplitMix64 &g, double a, double *p, double *v) {
    double X;
    do {
        X = g.uniform_open();
    } while (X > 0.999);
    double r = a / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
    iso(g, r, p);
    double q;
    while (true) {
        q = g.uniform();
        double y = 0.1 * g.uniform();
        if (y < q * q * std::pow(1.0 - q * q, 3.5)) break;
    }
    double ve = std::sqrt(2.0) * std::pow(1.0 + r * r / (a * a), -0.25) / std::sqrt(a);
    iso(g, q * ve, v);
}

// Solve Barker's equation D + D^3/3 = W for the parabolic anomaly D = tan(nu/2).
double barker(double W) {
    double y = std::cbrt(1.5 * W + std::sqrt(2.25 * W * W + 1.0));
    return y - 1.0 / y;
}

}  // namespace

extern "C" {

int nb_ic_uniform_ball(uint64_t seed, int64_t n, double mlo, double mhi, double radius, double sigma, double *m,
                       double *pos, double *vel) {
    if (n < 0 || !m || !pos || !vel || mhi < mlo || radius < 0 || sigma < 0) return NB_ERR_INVALID_ARGUMENT;
    SplitMix64 g(seed);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = mlo + (mhi - mlo) * g.uniform();
        iso(g, radius * std::cbrt(g.uniform()), pos + 3 * i);
        for (int k = 0; k < 3; ++k) vel[3 * i + k] = sigma * g.gauss();
    }
    com_frame(n, m, pos, vel);
    return NB_OK;
}

int nb_ic_plummer(uint64_t seed, int64_t n, double a, const double *offset, const double *vel_offset, double *m,
                  double *pos, double *vel) {
    if (n < 0 || !m || !pos || !vel || !(a > 0)) return NB_ERR_INVALID_ARGUMENT;
    SplitMix64 g(seed);
    for (int64_t i = 0; i < n; ++i) {
        m[i] = 1.0 / (double)n;
        plummer_body(g, a, pos + 3 * i, vel + 3 * i);
    }
    com_frame(n, m, pos, vel);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (offset) pos[3 * i + k] += offset[k];
            if (vel_offset) vel[3 * i + k] += vel_offset[k];
        }
    return NB_OK;
}

// Comet-string ensemble (the SL9 scenario of sl9.cpp / simsup.hpp at scale):
// system s = one central body at rest at the origin plus n_s ~ U{nmin..nmax}
// fragments of mass log-uniform in [1e9, 1e13] kg, radius from density
// 500 kg/m^3, strung along one parabolic orbit of pericentre q ~ U[5e7, 2e8] m
// that reaches pericentre 14-17 h after t = 0; consecutive fragments trail by
// 0.2-0.8 s of orbital time (a few km) plus 0.05 m/s velocity jitter.
int nb_ic_comet_ensemble(uint64_t seed, int64_t nsys, int64_t nmin, int64_t nmax, double central_mass,
                         double central_radius, double G, int64_t *counts, int64_t *offsets, double *m, double *r,
                         double *pos, double *vel, double *rest_c) {
    if (nsys < 0 || nmin < 1 || nmax < nmin || !counts || !offsets) return NB_ERR_INVALID_ARGUMENT;
    SplitMix64 sizes(seed);
    offsets[0] = 0;
    for (int64_t s = 0; s < nsys; ++s) {
        int64_t nf = nmin + (int64_t)(sizes.uniform() * (double)(nmax - nmin + 1));
        if (nf > nmax) nf = nmax;
        counts[s] = nf + 1;
        offsets[s + 1] = offsets[s] + nf + 1;
    }
    if (!m) return NB_OK;
    if (!r || !pos || !vel || !rest_c) return NB_ERR_INVALID_ARGUMENT;
    const double mu = G * central_mass;
    for (int64_t s = 0; s < nsys; ++s) {
        SplitMix64 g(seed ^ (0xA24BAED4963EE407ULL * (uint64_t)(s + 1)));
        int64_t o = offsets[s];
        int64_t nb = counts[s];
        m[o] = central_mass;
        r[o] = central_radius;
        rest_c[o] = 0.5;
        for (int k = 0; k < 3; ++k) pos[3 * o + k] = vel[3 * o + k] = 0.0;
        double q = 5e7 + 1.5e8 * g.uniform();
        double inc = std::acos(1.0 - 2.0 * g.uniform());
        double node = 2.0 * M_PI * g.uniform();
        double t_peri = 5.0e4 + 1.0e4 * g.uniform();  // seconds until pericentre
        double tscale = std::sqrt(2.0 * q * q * q / mu);
        double ci = std::cos(inc), si = std::sin(inc), cn = std::cos(node), sn = std::sin(node);
        double lag = 0.0;
        for (int64_t k = 1; k < nb; ++k) {
            int64_t b = o + k;
            double lm = 9.0 + 4.0 * g.uniform();
            m[b] = std::pow(10.0, lm);
            r[b] = std::cbrt(3.0 * m[b] / (4.0 * M_PI * 500.0));
            rest_c[b] = 0.1 + 0.8 * g.uniform();
            lag += 0.2 + 0.6 * g.uniform();
            double t = -(t_peri + lag);
            double D = barker(t / tscale);
            double nu = 2.0 * std::atan(D);
            double rr = q * (1.0 + D * D);
            double px = rr * std::cos(nu), py = rr * std::sin(nu);
            double vfac = std::sqrt(mu / (2.0 * q));  // parabolic speed components
            double vx = -vfac * std::sin(nu);
            double vy = vfac * (1.0 + std::cos(nu));
            // rotate from the orbital plane: inclination about x, then node about z
            double p3[3] = {px, py * ci, py * si}, v3[3] = {vx, vy * ci, vy * si};
            double P[3] = {cn * p3[0] - sn * p3[1], sn * p3[0] + cn * p3[1], p3[2]};
            double V[3] = {cn * v3[0] - sn * v3[1], sn * v3[0] + cn * v3[1], v3[2]};
            for (int c = 0; c < 3; ++c) {
                pos[3 * b + c] = P[c];
                vel[3 * b + c] = V[c] + 0.05 * g.gauss();
            }
        }
    }
    return NB_OK;
}

}  // extern "C"
