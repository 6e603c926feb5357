// nb_record.hip -- state and energy recording every k steps (gfx950).
//
// Reference: the memFreq template argument of gtd::system (gregsys.hpp:135-170)
// keeps, every memFreq steps, each body's position, velocity and kinetic
// energy (body::add_pos_vel_ke, gregbod.hpp:311-315) and the system totals
// tot_ke / tot_pe / tot_e (calc_acc_and_e :439-476, leapfrog_kdk_acc_e_and_step
// :514-526).  In the reference only the initial energies are actually
// recorded on the paths that compile (the step loop is instantiated with
// mem = false, FUNC_TEMPL_SELECT :16-18), so this is the documented intent
// built as an engine extension: every `every` steps, and before the first
// step of nb_evolve, the state (positions, velocities, labels) and the
// totals KE = sum 0.5 m |v| |v| and PE = sum of the halved per-body
// potentials (set_ke gregbod.hpp:917-920, cumulative_pe gregsys.hpp:358-363,
// calc_energy :608-630), summed in body order, go into a ring of device
// records.  Copies stay on the device (one D2D copy per vector); the totals
// come from the energy kernel and a one-lane body-order sum, so no host
// synchronisation is added to the step loop.
#include <algorithm>

#include "nb_system.hpp"

#pragma clang fp contract(off)

namespace nbe {

namespace {

// KE and PE totals in body order with calc_energy's operation order (the
// same arithmetic as nb_energy's host loop).
__global__ void k_energy_totals(int64_t n, const double *__restrict__ m, const double *__restrict__ v,
                                const double *__restrict__ pe, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double tke = 0, tpe = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double vx = v[i], vy = v[n + i], vz = v[2 * n + i];
        double vm2 = vx * vx;
        vm2 = vm2 + vy * vy;
        vm2 = vm2 + vz * vz;
        const double vm = __builtin_sqrt(vm2);
        double k = 0.5 * m[i];
        k = k * vm;
        k = k * vm;
        tpe = tpe + pe[i];
        tke = tke + k;
    }
    out[0] = tke;
    out[1] = tpe;
}

}  // namespace

void record_state(nb_system *s, uint64_t step) {
    const int64_t n = s->n;
    const uint64_t cap = s->rec_cap;
    if (!s->rec_every || !cap) return;
    const bool state = s->rec_what & NB_RECORD_STATE, energy = s->rec_what & NB_RECORD_ENERGY;
    const size_t per = (size_t)3 * (size_t)std::max<int64_t>(s->rec_n0, n);
    if (state && s->rec_x.n < per * cap) {
        // (re)sizing drops what was recorded
        s->rec_x.ensure(per * cap);
        s->rec_v.ensure(per * cap);
        s->rec_total = 0;
        s->rec_n0 = std::max<int64_t>(s->rec_n0, n);
    }
    s->rec_e.ensure(2 * cap);
    if (s->rec_step.size() != cap) {
        s->rec_step.assign(cap, 0);
        s->rec_n.assign(cap, 0);
        s->rec_ids.assign(cap, {});
    }
    const uint64_t slot = s->rec_total % cap;
    const size_t stride = (size_t)3 * (size_t)s->rec_n0;
    hipStream_t st = s->stream;
    if (n > 0 && state) {
        NB_HIP(hipMemcpyAsync(s->rec_x.p + slot * stride, s->x.p, (size_t)3 * n * 8, hipMemcpyDeviceToDevice, st));
        NB_HIP(hipMemcpyAsync(s->rec_v.p + slot * stride, s->v.p, (size_t)3 * n * 8, hipMemcpyDeviceToDevice, st));
    }
    if (energy) {
        if (n > 0) launch_pe_fp64(st, n, s->x.p, s->m.p, s->G, s->eps * s->eps, form_of(s), s->pe.p);
        hipLaunchKernelGGL(k_energy_totals, dim3(1), dim3(64), 0, st, n, s->m.p, s->v.p, s->pe.p,
                           s->rec_e.p + 2 * slot);
    } else {
        hipLaunchKernelGGL(k_energy_totals, dim3(1), dim3(64), 0, st, (int64_t)0, s->m.p, s->v.p, s->pe.p,
                           s->rec_e.p + 2 * slot);  // zero totals
    }
    NB_HIP(hipGetLastError());
    s->rec_step[slot] = (int64_t)step;
    s->rec_n[slot] = n;
    s->rec_ids[slot].assign(s->labels.begin(), s->labels.begin() + std::min<size_t>(s->labels.size(), (size_t)n));
    ++s->rec_total;
}

}  // namespace nbe
