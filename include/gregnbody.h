/*
 * gregnbody.h -- C ABI of the MI355X (gfx950) N-body engine.
 *
 * This is the drop-in boundary for the force-and-step path of the reference
 * library's gtd::system (GregW10/GregPackagesCPP, gregsys.hpp).  Every entry
 * point names the reference interface it replaces.  Conventions:
 *   - every function returns an int status: NB_OK (0) or a negative code that
 *     mirrors one of the reference's exception classes; nb_last_error() gives
 *     the message;
 *   - host arrays are caller-owned and copied; device buffers are owned by the
 *     library; a context owns one HIP stream on one device;
 *   - a context is not thread-safe (as gtd::system is not, gregbod.hpp:277);
 *     distinct contexts may be driven from distinct host threads.
 * Positions/velocities/accelerations are passed as interleaved (x, y, z)
 * triples of length 3*n; masses, radii and restitution coefficients as n
 * values.  All physical state is IEEE double.
 */
#ifndef GREGNBODY_H
#define GREGNBODY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (exception classes of gregbod.hpp:19-93, greg8tree.hpp:11-46,
 *      gregsys.hpp:125-134) ---- */
#define NB_OK 0
#define NB_ERR_INVALID_ARGUMENT (-1)      /* std::invalid_argument                 */
#define NB_ERR_NEGATIVE_MASS (-2)         /* negative_mass_error                   */
#define NB_ERR_NEGATIVE_RADIUS (-3)       /* negative_radius_error                 */
#define NB_ERR_OVERLAPPING_BODIES (-4)    /* overlapping_bodies_error              */
#define NB_ERR_TWO_BODY (-5)              /* two_body_error                        */
#define NB_ERR_EMPTY_SYSTEM (-6)          /* empty_system_error                    */
#define NB_ERR_INVALID_PLACEMENT (-7)     /* invalid_body_placement (barnes_hut)   */
#define NB_ERR_INDISTINGUISHABLE (-8)     /* indistinguishable_bodies (barnes_hut) */
#define NB_ERR_NO_EVOLUTION (-9)          /* no_evolution_error                    */
#define NB_ERR_NOT_SUPPORTED (-10)        /* method outside this engine's path     */
#define NB_ERR_CAPACITY (-11)             /* caller buffer or engine limit too small */
#define NB_ERR_DEVICE (-12)               /* HIP runtime failure                   */
#define NB_ERR_NSYS_LOAD (-13)            /* nsys_load_error                       */
#define NB_ERR_NSYS_FORMAT (-14)          /* nsys_format_error                     */
#define NB_ERR_CALLBACK (-15)             /* a step hook returned nonzero          */

/* ---- integration method bit flags: gtd::system constants gregsys.hpp:156-165 ---- */
#define NB_TWO_BODY 1
#define NB_EULER 2
#define NB_MODIFIED_EULER 4
#define NB_MIDPOINT 8
#define NB_LEAPFROG_KDK 16
#define NB_LEAPFROG_DKD 32
#define NB_RK4 64
#define NB_RK3_8 128
#define NB_BARNES_HUT 65536
#define NB_FAST_MULTIPOLE 131072

/* ---- collision modes: gtd::system template argument `collisions` (gregsys.hpp:166-168) ---- */
#define NB_NO_COLL_CHECK 0
#define NB_OVERLAP_COLL_CHECK 3

/* ---- arithmetic of the pairwise force (an engine choice, not in the reference) ---- */
#define NB_PREC_FP64 0       /* FP64 forces in the reference's operation order (parity) */
#define NB_PREC_MIXED 1      /* FP32 pair forces, Kahan-compensated FP32 sums, FP64 state */
#define NB_PREC_FP64_FAST 2  /* FP64 forces, j range split over the GPU, fixed combine order */

typedef struct nb_system nb_system;

/* Lifecycle.  Replaces gtd::system's constructors (gregsys.hpp:1478-1578) and
 * destructor; G starts at the SI value (system() : G{G_SI}, :1478). */
int nb_create(int device, nb_system **out);
int nb_destroy(nb_system *sys);
/* Message of the last failing call on this context (static storage if sys is NULL). */
const char *nb_last_error(const nb_system *sys);

/* State in.  Replaces system(const std::vector<body>&, ..., allow_overlapping_bodies)
 * (gregsys.hpp:1526-1545) and body validation check_mass/check_radius
 * (gregbod.hpp:329-336).  r and rest_c may be NULL (radius 1, restitution 1, the
 * body<> defaults gregbod.hpp:907-916).  When check_overlap != 0 the O(N^2)
 * construction-time overlap test of system::check_overlap (gregsys.hpp:307-342)
 * runs on the GPU and NB_ERR_OVERLAPPING_BODIES is returned on the first
 * overlapping pair in (outer, inner > outer) order, the context then holding
 * no bodies (as the reference's constructor throws). */
int nb_set_bodies(nb_system *sys, int64_t n, const double *m, const double *r, const double *pos,
                  const double *vel, const double *rest_c, int check_overlap);
/* system::check_overlap with mergeOverlappingBodies (gregsys.hpp:307-342)
 * on the current bodies: the reference's construction-time loop, which
 * merges an overlapping pair (body::operator+=, gregbod.hpp:1065-1082), erases
 * the inner body and restarts, with its exact visiting order after a merge.
 * The first overlapping outer body is found on the GPU; the loop is followed
 * on the host from there.  *removed (may be NULL) receives the number of
 * bodies removed; labels follow the surviving bodies (a merged body keeps the
 * outer body's label). */
int nb_merge_overlapping(nb_system *sys, int64_t *removed);
/* Overwrite the state of the current bodies in place (engine extension; the
 * per-body setters of gtd::body -- set_mass, set_radius, set_restitution,
 * set_acc, update, shift -- gregbod.hpp:498-600, applied from the host).  n
 * must equal the body count; any pointer may be NULL (that field is kept).
 * Labels, time elapsed, the merge log and the evolved state are kept, and
 * acc is the acceleration the next opening half kick uses, as the body's
 * stored acceleration is in the reference.  Negative masses or radii return
 * NB_ERR_NEGATIVE_MASS / NB_ERR_NEGATIVE_RADIUS with nothing changed. */
int nb_update_state(nb_system *sys, int64_t n, const double *m, const double *r, const double *pos,
                    const double *vel, const double *acc, const double *rest_c);
/* Step hooks: where evolve's callbacks run in the reference
 * (gregsys.hpp:2593-2607, 2946-2990).  For the direct-sum integrators, after
 * each step's update and before the collision pass (for_all_bods, passed into
 * leapfrog_kdk_acc_e_and_step :527-539 and FUNC_TEMPL_SELECT :16-57), and
 * after the collision pass (for_each_bod, :2961-2963), the engine
 * synchronises its stream and calls fn(user, sys, step, where) with where =
 * NB_HOOK_STEP or NB_HOOK_BODIES (selected by where_mask).  Inside the hook
 * the caller may read the state (nb_get_state, nb_body_ids, nb_num_bodies)
 * and change it (nb_update_state); nb_evolve, nb_advance and nb_set_bodies
 * are refused there.  A nonzero return stops the evolve with
 * NB_ERR_CALLBACK.  Steps that run hooks end with their closing half kick and
 * replay no captured graph.  Barnes-Hut steps call no hook, as the
 * reference's BH branch calls neither callback (:2893-2945).  fn = NULL
 * removes the hook. */
#define NB_HOOK_STEP 1
#define NB_HOOK_BODIES 2
typedef int (*nb_step_hook)(void *user, nb_system *sys, uint64_t step, int where);
int nb_set_step_hook(nb_system *sys, nb_step_hook fn, void *user, int where_mask);
/* State out (any pointer may be NULL).  Replaces iteration / operator[] over
 * bodies (gregsys.hpp:3178-3213, 3383-3387) with body::pos()/vel()/mass()/rad()
 * (gregbod.hpp:444-459).  cap is the capacity in bodies. */
int nb_get_state(const nb_system *sys, int64_t cap, double *m, double *r, double *pos, double *vel, double *acc,
                 double *rest_c);
/* num_bodies() (gregsys.hpp:1627). */
int nb_num_bodies(const nb_system *sys, int64_t *n);
/* Body labels: input index for original bodies; merged bodies get n0, n0+1, ...
 * in creation order (stands in for body_counter ids, gregbod.hpp:99-207). */
int nb_body_ids(const nb_system *sys, int64_t cap, int64_t *ids);

/* Parameters.  Setters return NB_ERR_INVALID_ARGUMENT where the reference's
 * returns false. */
int nb_set_units(nb_system *sys, const char *units_format); /* set_G_units / parse_units_format :245-282, 1655 */
int nb_set_G(nb_system *sys, double G);                     /* engine extension: G directly (G = 1 configs)   */
int nb_get_G(const nb_system *sys, double *G);              /* G_val (:1633)                                    */
int nb_set_timestep(nb_system *sys, double dt);             /* set_timestep (:1642-1648)                        */
int nb_set_iterations(nb_system *sys, uint64_t iterations); /* set_iterations (:1636-1641)                      */
/* set_softening (:1660-1666): eps >= sqrt(DBL_MIN); eps == 0 restores the
 * unsoftened direct kernel (engine extension).  eps > 0 selects the reference's
 * softened (bh_acc, :371-375) pair form for direct summation too. */
int nb_set_softening(nb_system *sys, double eps);
int nb_set_bh_opening_angle(nb_system *sys, double theta); /* set_bh_opening_angle (:1667-1672) */
/* Collision handling = the `collisions` template argument plus -DGREGSYS_MERGERS
 * and set_min_tot_com_mom (:1674-1679, default 1e10 at :209). */
int nb_set_collisions(nb_system *sys, int mode, int mergers, double min_tot_com_mom);
int nb_set_precision(nb_system *sys, int precision);

/* evolve(int integration_method) (gregsys.hpp:2593-2607): runs `iterations`
 * steps of NB_LEAPFROG_KDK (direct summation, or | NB_BARNES_HUT), or of the
 * reference's other direct-sum integrators NB_EULER, NB_MODIFIED_EULER,
 * NB_MIDPOINT (:2666-2891) and NB_LEAPFROG_DKD (:2993-3039), with s_coll
 * after every step when collisions are on (with Barnes-Hut, bh_s_coll
 * :631-686 instead: bounces between bodies of touching tree cells, never a
 * merge, :2937-2939); the first matching flag wins, as
 * in the reference's dispatch.  Like the reference, a KDK call first
 * recomputes a(x0).  elapsed_ns (may be NULL) receives the host wall time of
 * the call (evolve's return value).  NB_ERR_NOT_SUPPORTED: rk4 / rk3_8 (empty
 * in the reference) and Barnes-Hut with any integrator but KDK. */
int nb_evolve(nb_system *sys, int method, int64_t *elapsed_ns);
/* Engine extension: continue the previous evolve for `steps` more steps of
 * the same method, without recomputing a(x0) for KDK (the cached
 * accelerations are those of the current positions).  Returns
 * NB_ERR_NO_EVOLUTION if nothing was evolved. */
int nb_advance(nb_system *sys, uint64_t steps, int64_t *elapsed_ns);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Continues a direct-sum NB_LEAPFROG_KDK evolve exactly as
 * nb_advance does (same kernels, same split counts, bit-identical states), in
 * NB_PREC_MIXED or NB_PREC_FP64_FAST, with collisions and recording off; the
 * force pass's split partials live in a static device array of
 * NB_NOALLOC_SCRATCH_BYTES per GPU instead of context buffers, and the call
 * makes no hipMalloc / hipFree and captures no graph.  NB_ERR_CAPACITY when
 * the partials (6 floats or 3 doubles per body per split) exceed the array,
 * NB_ERR_NO_EVOLUTION without a previous direct KDK nb_evolve, and
 * NB_ERR_NOT_SUPPORTED for other precisions or with collisions or recording.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (contexts on several host threads) race and must be serialised
 * by the caller. */
#define NB_NOALLOC_SCRATCH_BYTES (32u << 20)
int nb_advance_noalloc(nb_system *sys, uint64_t steps, int64_t *elapsed_ns);
/* Total kinetic and potential energy of the current state in FP64, reference
 * formulae (set_ke gregbod.hpp:917-920, cumulative_pe gregsys.hpp:358-363,
 * calc_energy :608-630; bh_pe's softened form when eps > 0). */
int nb_energy(nb_system *sys, double *ke, double *pe);
#define NB_RECORD_STATE 1   /* positions, velocities, labels */
#define NB_RECORD_ENERGY 2  /* total KE and PE (PE: the direct O(N^2) FP64 sum) */
/* Engine extension in place of the memFreq template argument (gregsys.hpp:135-170,
 * body::add_pos_vel_ke gregbod.hpp:311-315, tot_ke/tot_pe/tot_e :420-476, whose
 * step loop records nothing on the paths that compile): every `every` steps of
 * nb_evolve / nb_advance, and before the first step of nb_evolve, keep the
 * positions, velocities and labels of the bodies and the total KE and PE (the
 * nb_energy formulae, summed in body order) in a ring of `capacity` records in
 * device memory; the oldest is overwritten first.  Recording adds no host
 * synchronisation but turns the captured step graphs off; with Barnes-Hut the
 * velocities are those of the loop, half a step ahead of the positions, and
 * the PE is the direct O(N^2) sum.  `what` is NB_RECORD_STATE and/or
 * NB_RECORD_ENERGY (an energy-only ring keeps no state and returns zero
 * positions and velocities; a state-only ring reports KE = PE = 0).  every = 0 stops recording and frees the ring;
 * nb_set_bodies drops the records. */
int nb_set_recording(nb_system *sys, uint64_t every, uint64_t capacity, int what);
/* taken: records made since recording was set; stored: how many are kept. */
int nb_record_count(const nb_system *sys, uint64_t *taken, uint64_t *stored);
/* Record k of the stored ones (0 = the oldest kept): step index since the last
 * nb_evolve, body count n, positions and velocities (3n interleaved), labels
 * (n), total KE and PE.  Any output may be NULL; call with pos = vel = ids =
 * NULL to learn n first. */
int nb_get_record(nb_system *sys, uint64_t k, int64_t *step, int64_t *n, double *pos, double *vel, int64_t *ids,
                  double *ke, double *pe);
/* body_tracker (gregsys.hpp:3492-3611) over the bodies at slots idx[0..k), in
 * that order (the tracker's predicate is the caller's: it picks the slots).
 * Outputs (each may be NULL): total mass (mass()), centre of mass com[3]
 * (com()), COM velocity com_vel[3] (com_vel()), mean distance to point[3]
 * (mean_dist_to(point); needs point), mean pairwise separation (mean_sep()).
 * All FP64: the O(k) sums in the subset's order, the separation summed per
 * outer body and then over the rows in order.  k = 0 with mean_sep requested
 * returns NB_ERR_INVALID_ARGUMENT, as mean_sep() throws on an empty tracker. */
int nb_track(nb_system *sys, const int64_t *idx, int64_t k, const double *point, double *mass, double *com,
             double *com_vel, double *mean_dist, double *mean_sep);
/* Merge records since nb_set_bodies: 4 int64 per record
 * (step, outer body label, absorbed body label, merged body label). */
int nb_merge_log(const nb_system *sys, int64_t cap, int64_t *records, int64_t *count);

/* Barnes-Hut octree of the current positions (bh_tree::add_bodies
 * greg8tree.hpp:1066-1078, calc_com :211-282) exported in DFS pre-order
 * (octants 0..7): depth, octant within parent (-1 root), body index (-1
 * internal), parent pre-order index, mass, com (3 per node), side (top.x-btm.x).
 * *count receives the node count; NB_ERR_CAPACITY if cap is too small. */
int nb_tree_topology(nb_system *sys, int64_t cap, int32_t *depth, int32_t *octant, int64_t *body, int64_t *parent,
                     double *mass, double *com, double *side, int64_t *count);
/* Accelerations of the current positions by the Barnes-Hut walk without
 * stepping (acc: 3n), and the number of accepted cells. */
int nb_bh_accelerations(nb_system *sys, double *acc, int64_t *interactions);

/* Adjoint of `steps` softened direct KDK steps from the current state (no
 * counterpart in the reference): given dL/dx_T and dL/dv_T (3n each), writes
 * dL/dx_0, dL/dv_0 (3n) and dL/dm (n).  Checkpoints every ckpt_every steps,
 * recomputing each segment forward.  The state is left at time T. */
int nb_adjoint(nb_system *sys, uint64_t steps, uint64_t ckpt_every, const double *dl_dx_T, const double *dl_dv_T,
               double *dl_dx0, double *dl_dv0, double *dl_dm, int64_t *fwd_ns, int64_t *bwd_ns);

/* ---- .nsys snapshots (system::to_nsys gregsys.hpp:3214-3258, the .nsys
 *      constructor :1479-1490 and load_from_nsys :943-1400) ----
 * Layout: a 96-byte header -- "NS", d_types (bit 0 float time, bits 1-3 float
 * M/R/T, bits 4-6 signed integral M/R/T), time_size, time elapsed and dt as
 * time_size-byte values in 16-byte fields, the sizes of M, R, T and of the
 * restitution coefficient, the six G-ratio integers (m_num, m_denom, d_num,
 * d_denom, t_num, t_denom) and the body count -- then one chunk per body:
 * restitution (16-byte field), mass, radius, x, y, z, vx, vy, vz in their
 * declared types and natural alignment; the file is zero-padded to a multiple
 * of 4 bytes.  Values travel as FP64 through this API (x87 long double fields
 * are converted on the host). */
typedef struct nb_nsys_info {
    int32_t d_types, time_size, m_size, r_size, t_size, rest_size;
    uint64_t g_vals[6];
    uint64_t num_bodies;
    double time_elapsed;
    double dt;
} nb_nsys_info;

/* Parse an .nsys file without a context (host only).  Fills *info, then, if
 * cap >= num_bodies, the bodies (pos/vel interleaved xyz).  Errors mirror
 * load_from_nsys: NB_ERR_NSYS_LOAD / NB_ERR_NSYS_FORMAT with the reference's
 * messages (nb_last_error(NULL)), NB_ERR_INVALID_ARGUMENT for dt <= 0
 * (check_dt :1401-1405), NB_ERR_CAPACITY when cap is too small (info is
 * still filled). */
int nb_nsys_read(const char *path, int check_nan_inf, nb_nsys_info *info, int64_t cap, double *m, double *r,
                 double *pos, double *vel, double *rest_c);
/* Write an .nsys file (host only) with mass, radius, position and velocity
 * all of value_size bytes: 8 (double, 80-byte chunks) or 16 (x87 long double,
 * 144-byte chunks); time and restitution as long double.  info supplies the
 * time elapsed, dt and G ratios (other fields ignored).  *bytes = file size. */
int nb_nsys_write(const char *path, const nb_nsys_info *info, int64_t n, const double *m, const double *r,
                  const double *pos, const double *vel, const double *rest_c, int value_size, int64_t *bytes);
/* The context forms: load replaces bodies, dt, time elapsed and the G ratios
 * (G recomputed as parse_units_format does); allow_overlapping = 0 runs the
 * construction overlap check.  save writes the current state (to_nsys). */
int nb_load_nsys(nb_system *sys, const char *path, int check_nan_inf, int allow_overlapping);
int nb_save_nsys(const nb_system *sys, const char *path, int value_size, int64_t *bytes);
/* Time elapsed (system::elapsed_time :1630-1632): set to iterations*dt by
 * nb_evolve, advanced by steps*dt by nb_advance, 0 after nb_set_bodies, read
 * from the file by nb_load_nsys. */
int nb_time_elapsed(const nb_system *sys, double *t);

/* Tuning knobs (engine extension; 0 restores the default):
 *   "direct_kernel" FP32 direct kernel: 1 j stream through scalar loads,
 *                  2 j tiles through LDS (identical sums)
 *   "direct_ipl"   i-bodies per lane of the FP32 direct kernel (1, 2, 4)
 *   "direct_waves" target waves per SIMD for the direct kernel's j split
 *   "bh_bucket"    rejected tree nodes with at most this many bodies are summed
 *                  body by body in the FP32 Barnes-Hut walk
 *   "bh_walk"      FP32 Barnes-Hut walk: 1 the frontier walk (64 cell tests per
 *                  step, the default), 2 the per-wave depth-first walk; both
 *                  accept the same cells and differ only in summation order
 *   "bh_stack"     node entries of each wave's frontier stack (64..4096); when
 *                  a group would overflow it, that step's walk is redone
 *                  depth-first (same cells accepted)
 *   "bh_sort_passes" 9-bit radix passes of the octree sort over the high key
 *                  (4..7; the default adapts them to how many bodies shared
 *                  the coarser cells in the last build, and nb_get_tuning then
 *                  reports the passes the next build takes; the tree is the same)
 *   "step_graphs"  1 capture step loops of small systems in HIP graphs, 2 off */
int nb_set_tuning(nb_system *sys, const char *key, int64_t value);
int nb_get_tuning(const nb_system *sys, const char *key, int64_t *value);
/* Autotuner (engine extension): times the candidate settings of the knobs
 * above on the current bodies with HIP events and keeps the fastest.  With a
 * cache_path, results are read from / appended to that text file, keyed by
 * (kernel, body count, precision). */
int nb_autotune(nb_system *sys, int method, const char *cache_path);

/* Kernel timing (engine extension): when enabled, every launch of the named
 * kernel class is bracketed by HIP events on the context's stream.
 * names: "force", "update", "energy", "bh_build", "bh_walk", "bh_coll", "adjoint". */
int nb_profile_enable(nb_system *sys, int on);
int nb_profile_read(nb_system *sys, const char *name, double *total_ms, int64_t *launches);

/* ---- batched ensembles: many independent small systems in one persistent
 *      launch (the run_sim thread-per-system pattern, simsup.hpp:209-391) ---- */
typedef struct nb_batch nb_batch;
/* offsets[nsys+1] partitions the body arrays; G, dt, eps shared. */
int nb_batch_create(int device, int64_t nsys, const int64_t *offsets, const double *m, const double *r,
                    const double *pos, const double *vel, const double *rest_c, nb_batch **out);
int nb_batch_destroy(nb_batch *b);
const char *nb_batch_last_error(const nb_batch *b);
int nb_batch_set_params(nb_batch *b, double G, double dt, double eps, int mode, int mergers, double min_tot_com_mom,
                        int precision);
int nb_batch_evolve(nb_batch *b, uint64_t steps, int64_t *elapsed_ns);
/* Engine extension: device time of the stepping launches of the last
 * nb_batch_evolve (HIP events on the batch's stream around them, the a(x0)
 * launch excluded) and the number of those launches. */
int nb_batch_kernel_time(const nb_batch *b, double *total_ms, int64_t *launches);
/* Per-system body counts (nsys) and the state packed per system in body
 * order; cap is in bodies. */
int nb_batch_get_state(const nb_batch *b, int64_t cap, int64_t *counts, double *m, double *r, double *pos,
                       double *vel, int64_t *ids);
/* Merge records of all systems: 5 int64 per record
 * (system, step, outer label, absorbed label, merged label). */
int nb_batch_merge_log(const nb_batch *b, int64_t cap, int64_t *records, int64_t *count);

/* ---- seeded initial conditions (SplitMix64 stream; engine utility) ---- */
/* Uniform ball (the reference's own recipe r = R cbrt(U), cos(th) = 1-2U,
 * phi = 2 pi U, gregsys.hpp:1818-1824), masses U[mlo, mhi], Gaussian velocities
 * (Box-Muller, sigma), shifted to the centre-of-mass frame. */
int nb_ic_uniform_ball(uint64_t seed, int64_t n, double mlo, double mhi, double radius, double sigma, double *m,
                       double *pos, double *vel);
/* Plummer sphere, G = M = 1, scale a, equal masses 1/n (Aarseth, Henon &
 * Wielen 1974), centre-of-mass frame.  offset/vel_offset shift the sphere. */
int nb_ic_plummer(uint64_t seed, int64_t n, double a, const double *offset, const double *vel_offset, double *m,
                  double *pos, double *vel);
/* Ensemble of comet strings around a central body: system s has 1 + n_s bodies
 * (central body first); counts[nsys] receives n_s + 1 and offsets[nsys+1] the
 * packing.  When m == NULL only counts/offsets are produced (sizing pass). */
int nb_ic_comet_ensemble(uint64_t seed, int64_t nsys, int64_t nmin, int64_t nmax, double central_mass,
                         double central_radius, double G, int64_t *counts, int64_t *offsets, double *m, double *r,
                         double *pos, double *vel, double *rest_c);

#ifdef __cplusplus
}
#endif
#endif /* GREGNBODY_H */
