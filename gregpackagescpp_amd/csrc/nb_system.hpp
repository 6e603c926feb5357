// nb_system.hpp -- the context behind the C ABI handle (one gtd::system on one device).
#pragma once
#include "nb_internal.hpp"

namespace nbe {
struct BHWork;    // nb_bh.hip
struct CollWork;  // nb_coll.hip
void bh_work_free(BHWork *w);
void coll_work_free(CollWork *w);
}  // namespace nbe

struct nb_system {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // parameters (gregsys.hpp:169-212)
    double G = 6.6743e-11;
    uint64_t G_vals[6] = {1, 1, 1, 1, 1, 1};
    double dt = 1.0, half_dt = 0.5;
    uint64_t iterations = 1000;
    double eps = 0.0;
    double theta = 1.0471975511965977;  // PI/3 (gregsys.hpp:212)
    int coll_mode = NB_NO_COLL_CHECK;
    bool mergers = false;
    double min_tot_com_mom = 1e10;       // BILLION*10 (gregsys.hpp:209)
    int precision = NB_PREC_FP64;

    // state: SoA in HBM, [x block | y block | z block] for vectors
    int64_t n = 0;
    nbe::DBuf<double> m, r, e, x, v, a, pe;
    std::vector<int64_t> labels;
    int64_t next_label = 0;
    bool evolved = false;
    bool acc_valid = false;
    int acc_method = 0;  // method whose accelerations are cached
    uint64_t step_count = 0;
    long double time_elapsed = 0;  // system::time_elapsed (gregsys.hpp:206)
    std::vector<int64_t> merge_log;  // 4 per record

    // state recording (nb_set_recording): a ring of rec_cap records taken
    // every rec_every steps; a slot holds 3 * rec_n0 doubles per vector
    uint64_t rec_every = 0, rec_cap = 0, rec_total = 0;
    int rec_what = 0;  // NB_RECORD_STATE | NB_RECORD_ENERGY
    int64_t rec_n0 = 0;
    nbe::DBuf<double> rec_x, rec_v, rec_e;   // [slot][x block | y | z], same for v; [slot][ke, pe]
    std::vector<int64_t> rec_step, rec_n;    // per slot: step index, body count
    std::vector<std::vector<int64_t>> rec_ids;  // per slot: labels

    // tuning knobs (0 = engine default); set by nb_set_tuning / nb_autotune
    int tune_bh_bucket = 0;
    int tune_direct_ipl = 0;
    int tune_direct_kernel = 0;  // 1: scalar-load j stream, 2: LDS j tiles (default)
    int tune_direct_waves = 0;   // target waves per SIMD for the direct split
    int tune_step_graphs = 0;    // 0: graphs for n <= kGraphMaxBodies, 1: always, 2: never
    int tune_bh_walk = 0;        // 0/1: frontier walk (FP32 path), 2: per-wave DFS walk
    int tune_bh_stack = 0;       // frontier walk stack entries per wave (0: engine default)
    int tune_bh_sort = 0;        // octree sort: radix passes over the high key (4..7; 0: adaptive)

    // captured direct step loop (kGraphSteps steps) and what it was captured for
    nbe::StepGraph step_graph;
    nbe::GraphKey step_graph_key;

    // step hooks (nb_set_step_hook): called where evolve's callbacks run
    nb_step_hook hook_fn = nullptr;
    void *hook_user = nullptr;
    int hook_mask = 0;
    bool in_hook = false;

    nbe::DirectWork dw;
    nbe::BHWork *bh = nullptr;
    nbe::CollWork *coll = nullptr;
    nbe::Profiler prof;

    ~nb_system();
};

namespace nbe {
// direct summation driver pieces used by several modules (nb_api.hip)
void direct_forces(nb_system *s);                    // a(x) into s->a (and pos32 for the mixed path)
void direct_kdk_steps(nb_system *s, uint64_t steps); // steps of KDK from valid a
int form_of(const nb_system *s);                     // 0 = unsoftened reference form, 1 = softened
GraphKey graph_key(const nb_system *s);
void set_static_error(const std::string &msg);  // message for nb_last_error(NULL)
// step hooks: call the hook for `where` when set (synchronises the stream)
void run_hook(nb_system *s, uint64_t step, int where);
inline bool hooks_on(const nb_system *s) { return s->hook_fn != nullptr && s->hook_mask != 0; }
// construction-time overlap check (nb_overlap.hip)
bool first_overlap(nb_system *s, int64_t from, int64_t *outer, int64_t *inner);
int64_t merge_overlapping(nb_system *s);
// state recording (nb_record.hip): a record of the current state when `step`
// is a multiple of the recording interval
void record_state(nb_system *s, uint64_t step);
inline bool record_due(const nb_system *s, uint64_t step) { return s->rec_every && step % s->rec_every == 0; }
// body_tracker reductions over bodies idx[0..k) (nb_track.hip): out9 = mass,
// com[3], com_vel[3], mean distance to point (0 when point is NULL), mean separation
void track_bodies(nb_system *s, const int64_t *idx, int64_t k, const double *point, double *out9);
// collisions after a step's closing kick (nb_coll.hip)
void collision_pass(nb_system *s, uint64_t step);
// Barnes-Hut (nb_bh.hip)
void bh_forces(nb_system *s, int64_t *interactions);  // a(x) into s->a by the tree walk
void bh_steps(nb_system *s, uint64_t steps);
// Barnes-Hut steps leave the bodies in the tree's order; this puts them back
// in input order (every entry point does so before it returns)
void bh_restore_order(nb_system *s);
void bh_forget_order(nb_system *s);  // the bodies were replaced
int bh_next_sort_passes(const nb_system *s);  // radix passes the next octree sort takes when adaptive
inline void restore_order_quietly(nb_system *s) {
    try {
        bh_restore_order(s);
    } catch (...) {
    }
}
int bh_topology(nb_system *s, int64_t cap, int32_t *depth, int32_t *octant, int64_t *body, int64_t *parent,
                double *mass, double *com, double *side, int64_t *count);
// adjoint (nb_adjoint.hip)
void adjoint_run(nb_system *s, uint64_t steps, uint64_t ckpt, const double *gxT, const double *gvT, double *gx0,
                 double *gv0, double *gm, int64_t *fwd_ns, int64_t *bwd_ns);
}  // namespace nbe
