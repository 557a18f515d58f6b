/*
 * rockable_amd.h -- C ABI of the MI355X DEM time-step engine.
 *
 * The engine runs Rockable's velocity-Verlet step (OBB neighbour list,
 * sphero-polyhedral sub-element contacts, ForceLaw with breakable bonds,
 * per-body force/moment reduction, quaternion update) on one AMD Instinct
 * MI355X (gfx950), in FP64, with results identical to Rockable's CPU path.
 *
 * Conventions: plain pointers and sizes; every call returns 0 on success and
 * a negative code on error (rk_last_error() then describes it); nothing calls
 * exit().  One engine per host thread; calls on one engine are serialised.
 * Each engine owns one device stream and all its device memory.
 *
 * Reference interfaces replaced (paths relative to the Rockable tree):
 *   rk_engine_create/destroy   class Rockable ctor/dtor (src/Core/Rockable.cpp:78-127)
 *   rk_load_conf               Rockable::loadConf (Rockable.cpp:1085-1145) incl.
 *                              the trailing accelerations() "fake step"
 *   rk_load_driving_system     DrivingSystem::read (src/Core/DrivingSystem.cpp:44-268)
 *   rk_set_force_law           conf keyword forceLaw -> Factory<ForceLaw>::Create,
 *                              plug, init (Rockable.cpp:763-781; ForceLaw.hpp:44-55)
 *   rk_set_body_force          conf keyword BodyForce (Rockable.cpp:1026-1040;
 *                              src/BodyForces/BodyForce.hpp:45-57)
 *   rk_update_nl               Rockable::UpdateNL -> UpdateNL_bruteForce /
 *                              UpdateNL_linkCells (Rockable.cpp:1614-1809)
 *   rk_accelerations           Rockable::accelerations (Rockable.cpp:3562-3719)
 *   rk_step                    body of the integrate() loop (Rockable.cpp:2584-2745)
 *   rk_integrate               Rockable::integrate (Rockable.cpp:2538-2746), with
 *                              its output files perf.txt, kineticEnergy.txt and
 *                              staticBalance.txt in the output directory
 *   rk_load_data_extractors    Rockable::readDataExtractors (Rockable.cpp:1152-1180)
 *   rk_save_conf               Rockable::saveConf (Rockable.cpp:440-610)
 *   rk_get_bodies              Particles[i].pos/Q/vel/vrot/acc/arot/force/moment
 *   rk_get_interactions        the Interactions sets (Rockable.hpp:107) and
 *                              activeInteractions, lexicographic order
 *                              (Interaction.hpp:104-122)
 */
#ifndef ROCKABLE_AMD_H
#define ROCKABLE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rk_engine rk_engine;

/* Interaction record: the fields of a conf "Interactions" line
 * (Rockable.cpp:571-578) plus the interface index of a bond (-1 if none). */
typedef struct {
  int32_t i, j, type, isub, jsub, stick;
  double n[3], dn, pos[3], vel[3], fn, ft[3], mom[3], damp;
} rk_interaction;

typedef struct {
  int64_t nbodies, ndriven, nlisted, nactive, ninterfaces, nl_updates, steps;
  double t, dt, total_broken_area;
  int64_t last_candidates; /* body pairs passed by the last broad phase */
  int64_t guard_reads;     /* host reads of the device displacement guard */
  int64_t guard_wasted;    /* iterations enqueued after a rebuild fell due (done nothing) */
  int64_t graph_launches;  /* runs of graph_steps steps launched as one captured graph */
  int64_t graph_captures;  /* graphs captured (one per device view, cached) */
  int64_t guard_polls;     /* host waits on the guard's progress word (with a rebuild network) */
} rk_counts;

/* Statistics of the packing, reduced on the device in a fixed order
 * (DataExtractor_dnStat.cpp and DataExtractor_TrackDamage.cpp report the
 * overlap and damage figures; kinetic energies as in the reference's
 * kineticEnergy output, Rockable.cpp:3958-3970).  Free bodies only for the
 * body sums; "contacts" are the active interactions with dn < 0. */
typedef struct {
  double kinetic_translational;  /* sum m |v|^2 / 2 */
  double kinetic_rotational;     /* sum w_b . (I w_b) / 2, w_b in the body frame */
  double vmax;                   /* largest |v| */
  int64_t nfree;                 /* free (not driven) bodies */
  int64_t nactive;               /* active interactions (dn < 0 or bonded) */
  int64_t ncontacts;             /* active with dn < 0 */
  int64_t active_pairs;          /* body pairs with an active interaction */
  double coordination;           /* 2 active_pairs / nfree */
  double dn_mean, dn_min;        /* over the contacts (0 if none) */
  int64_t ninterfaces;           /* live breakable interfaces */
  int64_t broken_interfaces;     /* interfaces broken since the load */
  double total_broken_area;      /* TrackDamage's broken area */
} rk_stats;

/* What rk_inspect_conf reports of a deck. */
typedef struct {
  int64_t nbodies, ndriven, nshapes, ninteractions, ninterfaces, ngroups;
  int32_t law;       /* 0 Default, 1 Avalanche, 2 StickedLinks */
  int32_t linkcells; /* UpdateNL linkCells (1) or bruteForce (0) */
} rk_conf_info;

/* The deck reader of rk_load_conf alone, on the host, with no device and no
 * engine (initParser and the interface linking of loadConf,
 * Rockable.cpp:643-1069, 956-1002): parses path (its redirection and shape
 * library included) and fills *out.  On error returns a negative code and
 * writes the message into err (at most errlen bytes, NUL-terminated). */
int rk_inspect_conf(const char* path, rk_conf_info* out, char* err, size_t errlen);

int rk_engine_create(rk_engine** out, int device);
void rk_engine_destroy(rk_engine* eng);
const char* rk_last_error(const rk_engine* eng);

int rk_load_conf(rk_engine* eng, const char* path);
/* Batched form (no counterpart in the reference): n decks, independent
 * packings that must lie apart in space, loaded as one model -- every
 * deck's driven bodies first, then every deck's free bodies; shapes shared by
 * name (a name must mean one shape), groups renumbered per deck so that each
 * keeps its own parameters; each deck's interactions, interfaces and
 * controls (driving_paths[p], may be NULL) renumbered.  The decks must agree
 * on forceLaw, dt, interVerlet, interConf, DVerlet, dVerlet,
 * dynamicUpdateNL, ParamsInInterfaces, UpdateNL, cellMinSizes, gravity,
 * BodyForce, t and tmax.  Each packing then computes what it computes alone
 * (with dynamicUpdateNL the check is the batch's: one packing's motion
 * rebuilds every list).  Then as rk_load_conf (the fake step included). */
int rk_load_batch(rk_engine* eng, const char* const* conf_paths, const char* const* driving_paths, int n);
/* Body ranges of packing p: out[0] first driven body, out[1] driven count,
 * out[2] first free body, out[3] free count (one packing after rk_load_conf). */
int rk_batch_packing(rk_engine* eng, int p, int64_t* out);
int rk_load_driving_system(rk_engine* eng, const char* path);
/* readDataExtractors (Rockable.cpp:1152-1180): a dataExtractors.txt file of
 * "dnStat <file> <nrec>" and "TrackDamage <file> <nrec>" lines (other names
 * are skipped, as the reference skips an unknown name); the files are opened
 * now, relative to the directory of path, and rk_integrate records them every
 * nrec steps of its step counter (Rockable.cpp:2719-2726).  A missing file
 * adds no extractor. */
int rk_load_data_extractors(rk_engine* eng, const char* path);
/* Returns 1 (not an error) when the name is unknown and Default is used. */
int rk_set_force_law(rk_engine* eng, const char* name);
/* name: AttractingPoint (x y z acceleration), PreferredDirection (axisBody.xyz
 * axis.xyz momentMax), ViscousFluid (fluidDensity), or "" to remove. */
int rk_set_body_force(rk_engine* eng, const char* name, const double* args, int nargs);
/* Options: "record_geometry" (1: keep n/pos/vel of every listed interaction
 * current each step; 0: only at conf-save steps and for bonds),
 * "profile" (device events on the engine stream: 1 brackets the force phase
 * "force" and the neighbour-list rebuild "nl"; 2 brackets every kernel),
 * "heavy_min" (slot count above which a body's resultant is summed by a
 * whole workgroup; default 256),
 * "exact_sums" (1: the resultants of fixed and velocity-driven walls are
 * summed in the reference's serial order too; 0, the default: by a fixed
 * tree, which changes only their reported force and moment by rounding),
 * "gravity_x", "gravity_y", "gravity_z" (the conf keyword gravity, changed
 * between steps),
 * "guard_batch" (with dynamicUpdateNL or live bonds and no network:
 * iterations enqueued between two host reads of the device check; default 16;
 * while bonds live the batches adapt between 2 and it: twice the iterations
 * a batch ended by a freeze executed, doubled after one that was not),
 * "graph_steps" (n > 1: runs of n steps with no host action between them go
 * out as one HIP graph, captured once per device view and replayed; 0 or 1,
 * the default: every kernel launched on its own),
 * "model_scale" (the rebuild-interval network's proposals are multiplied by
 * it; default 0.75),
 * "skip_far" (1, the default: in steps that write no geometry of inactive
 * entries, record_geometry 0, an inactive entry with zero history whose gap
 * is proven positive by the bodies' accumulated motion bounds is not
 * evaluated; its gap is recomputed when the list is read out or rebuilt;
 * 0: every listed entry every step), "skip_segment" (entries a workgroup
 * tests and compacts before evaluating them under skip_far: 512, 1024, 2048,
 * 4096; 256: no compaction; default 1024), "count_skips" (1: the geometry pass
 * counts entries skipped, inactive with zero history, and looked at, read
 * back with rk_get_option as "skipped_entries", "zero_history_entries" and
 * "entries_looked_at"; setting it again clears them; 0 stops counting),
 * launch shapes: "fuse_law" (1, the default: contact kinematics and force law
 * in one kernel; 0: two), "type_sort" (1, the default: the fused kernel takes
 * a block's entries by contact type; 0: in list order), "law_group" (geometry blocks per force-law
 * wavefront when not fused: 1, 2, 4, 8; default 1), "geom_block"
 * (contact-kinematics workgroup when not fused: 64, 128, 256, 512; default
 * 256), "gather_lanes" (lanes per body in the resultant sums: 8, 16; default
 * 16), "gather_groups" (groups of 64 / gather_lanes bodies per resultant
 * wavefront: 1, 2, 4, 8, 16, at most gather_lanes; 0, the default: 1 below
 * 262144 bodies, gather_lanes from there), "fuse_glob" (1: the predictor
 * writes the free bodies' global vertices, one thread per body; 2: one
 * thread per vertex; 3: workgroups of 64 bodies, their vertices shared by the
 * workgroup; 0: a kernel of its own; -1, the default: 0 from 262144
 * bodies, else 3), "gather_waves" (wavefronts per resultant
 * workgroup: 1, 2, 4; default 4), "body_block" (per-body kernels: 64, 128,
 * 256; default 256), "persistent" (0, the default: one tile per workgroup;
 * n in [8, 1048576]: the step kernels' grids capped at n workgroups, each
 * looping over its tiles), "narrow_pairs" (candidate body pairs per
 * narrow-phase wavefront: 1, 2; 0, the default: 2 for shapes of up to 8
 * vertices, else 1).  No option changes results. */
int rk_set_option(rk_engine* eng, const char* name, double value);
/* Also reads the deck's scalars "dt", "interVerlet", "interConf", "DVerlet",
 * "dVerlet" and "tmax". */
int rk_get_option(rk_engine* eng, const char* name, double* value);
/* Autotuner (no counterpart in the reference) of fuse_law, law_group,
 * geom_block, gather_lanes, gather_groups, fuse_glob, gather_waves,
 * body_block, persistent, graph_steps, type_sort and skip_segment, one after the other: the
 * simulation runs on, trial after trial of steps_per_trial steps, each
 * candidate value twice, and the fastest is kept: by the step kernels' own
 * durations per step (events around each kernel) for the kernel shapes, by
 * wall time per step (rebuilds excluded) for persistent and graph_steps;
 * the shapes it started from are timed (wall time) before and after the
 * search and kept if the search ends slower.  Every candidate computes
 * the same state, so the steps taken are those rk_step would take;
 * *steps_done receives their number. */
int rk_autotune(rk_engine* eng, int steps_per_trial, int64_t* steps_done);

int rk_update_nl(rk_engine* eng);
int rk_accelerations(rk_engine* eng);
int rk_step(rk_engine* eng, int64_t nsteps);
int rk_integrate(rk_engine* eng, const char* outdir, int64_t max_steps, int64_t* steps_done);
int rk_save_conf(rk_engine* eng, const char* path);
int rk_synchronize(rk_engine* eng);

int rk_get_counts(rk_engine* eng, rk_counts* out);
int rk_get_stats(rk_engine* eng, rk_stats* out);
/* rk_get_stats without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * reduction's partial sums live in a fixed scratch array of 2^20 doubles
 * (8 MiB) per device, held until the process ends; a packing whose
 * statistics need more (about 4 per 256 bodies plus 5 per 256 listed
 * interactions, over 40 million listed interactions) is rejected with an
 * error.  The array is shared by every engine on the device: calls that
 * overlap on one device race.  Results equal rk_get_stats' bit for bit. */
int rk_get_stats_noalloc(rk_engine* eng, rk_stats* out);
/* Any pointer may be NULL; arrays hold 3 doubles per body (Q: w x y z). */
int rk_get_bodies(rk_engine* eng, double* pos, double* Q, double* vel, double* vrot, double* acc, double* arot,
                  double* force, double* moment, size_t nbodies);
int rk_set_bodies(rk_engine* eng, const double* pos, const double* Q, const double* vel, const double* vrot,
                  size_t nbodies);
/* active_only: 1 for the active interactions (dn < 0 or bonded), 0 for every
 * listed one.  *n receives the total; at most cap records are written. */
int rk_get_interactions(rk_engine* eng, int active_only, rk_interaction* out, size_t cap, size_t* n);
/* Accumulated device time of a bracket while "profile" is on: "force" and
 * "nl" at level 1; "predict", "interact", "gather", "breakage", "nl" at
 * level 2. */
int rk_get_kernel_time(rk_engine* eng, const char* kernel, int64_t* launches, double* total_ms);
int rk_reset_kernel_time(rk_engine* eng);

/* ---------------------------------------------------------------------------
 * Rebuild-interval proposal network.  No counterpart in the reference.  With
 * dynamicUpdateNL the displacement check (dynamicCheckUpdateNL,
 * Rockable.cpp:1571-1596) runs on the device and the host reads it once per
 * batch of iterations; the network proposes the size of the next batch from
 * 16 statistics of the packing (rk_guard_features).  The guard alone decides
 * every rebuild, so results do not depend on the network.
 * The network: 16 -> 64 (tanh) -> 64 (tanh) -> 1, FP32, trained on the
 * device (Adam, mean squared error) by the library's own kernels.
 * Parameters: W1[64][16] b1[64] W2[64][64] b2[64] W3[64] b3[1].
 * ------------------------------------------------------------------------- */
typedef struct rk_mlp rk_mlp;
int rk_mlp_create(rk_mlp** out, int device, uint64_t seed); /* Glorot-uniform weights */
void rk_mlp_destroy(rk_mlp* mlp);
const char* rk_mlp_last_error(const rk_mlp* mlp);
int rk_mlp_nparams(void);
int rk_mlp_get_params(rk_mlp* mlp, float* out, size_t n);
int rk_mlp_set_params(rk_mlp* mlp, const float* in, size_t n);
/* y[s] = network(X[s][0..15]) for s < n */
int rk_mlp_forward(rk_mlp* mlp, const float* X, size_t n, float* y);
/* loss = mean (y - t)^2 over the n samples, grad = its gradient (nparams) */
int rk_mlp_loss_grad(rk_mlp* mlp, const float* X, const float* t, size_t n, float* loss, float* grad);
/* Adam on shuffled minibatches; *final_loss = loss over all n samples */
int rk_mlp_train(rk_mlp* mlp, const float* X, const float* t, size_t n, int epochs, int batch, float lr,
                 uint64_t seed, float* final_loss);
/* The engine sizes its batches by the network (NULL: fixed guard_batch).
 * The network predicts log(1 + iterations until the next rebuild falls due). */
int rk_set_rebuild_model(rk_engine* eng, rk_mlp* mlp);
/* The 16 statistics of the current state that the network reads. */
int rk_guard_features(rk_engine* eng, float* out16);

#ifdef __cplusplus
}
#endif
#endif
