/*
 * rk_oracle.h -- CPU oracle for the Rockable velocity-Verlet DEM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (C99 + OpenMP) restatement of
 * Rockable's DefaultParallelMode hot path, used as the parity checker for the
 * HIP engine and as the CPU baseline timed by bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product library (librockable_amd.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the Rockable source tree).  The arithmetic of the toofus primitives that the
 * reference takes from an external, unvendored header library (vec3r, quat,
 * OBB, linkCells, DataTable) is re-derived here; the forms chosen are the ones
 * that reproduce the reference's own golden conf files (see rk_oracle.c).
 */
#ifndef RK_ORACLE_H
#define RK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rko_sim rko_sim;

/* Interaction record as written in the conf "Interactions" block
 * (Rockable.cpp:571-578): i j type isub jsub n dn pos vel fn ft mom damp. */
typedef struct {
  int32_t i, j, type, isub, jsub, stick;
  double n[3], dn, pos[3], vel[3], fn, ft[3], mom[3], damp;
} rko_interaction;

rko_sim* rko_create(void);
void rko_destroy(rko_sim* s);
const char* rko_last_error(rko_sim* s);
void rko_set_threads(rko_sim* s, int n);
void rko_set_verbose(rko_sim* s, int v);

/* loadConf (Rockable.cpp:1085-1145), including the trailing accelerations()
 * "fake step" when run_fake_step != 0. */
int rko_load_conf(rko_sim* s, const char* path, int run_fake_step);
/* DrivingSystem::read (DrivingSystem.cpp:44-497): Control lines and the
 * servos (tritri*, shakers, ramp); the file is read again at each conf save
 * of rko_integrate / rko_run_steps. */
int rko_read_driving(rko_sim* s, const char* path);

/* readDataExtractors (Rockable.cpp:1152-1180): dnStat and TrackDamage, their
 * files relative to the directory of path; recorded by rko_integrate. */
int rko_load_data_extractors(rko_sim* s, const char* path);

void rko_update_nl(rko_sim* s);            /* UpdateNL() */
void rko_accelerations(rko_sim* s);        /* accelerations() */
/* update_interactions() alone (Rockable.cpp:2934-2977): kinematics of every listed interaction. */
void rko_update_kinematics(rko_sim* s);
/* resultants + accelerations from the stored interaction forces (no force law). */
void rko_resultants_only(rko_sim* s);
void rko_vv_step(rko_sim* s);              /* velocityVerletStep() */
/* integrate() (Rockable.cpp:2538-2746).  Writes conf<i> files into outdir
 * when outdir != NULL.  max_steps < 0: run until t > tmax. Returns steps. */
int64_t rko_integrate(rko_sim* s, const char* outdir, int64_t max_steps);
/* Run nsteps of the integrate() loop body without any conf output. */
int64_t rko_run_steps(rko_sim* s, int64_t nsteps);
int rko_save_conf(rko_sim* s, const char* path);

size_t rko_nbodies(rko_sim* s);
size_t rko_ndriven(rko_sim* s);
size_t rko_nlisted(rko_sim* s);
size_t rko_nactive(rko_sim* s);
size_t rko_ninterfaces(rko_sim* s);
double rko_time(rko_sim* s);
int64_t rko_nl_updates(rko_sim* s);
double rko_total_broken_area(rko_sim* s);
/* Copy per-body arrays (3 doubles per body, Q: 4 doubles w x y z). Any
 * pointer may be NULL. */
void rko_get_bodies(rko_sim* s, double* pos, double* vel, double* acc, double* Q, double* vrot, double* arot,
                    double* force, double* moment);
void rko_set_bodies(rko_sim* s, const double* pos, const double* vel, const double* Q, const double* vrot);
/* masses (nb) and principal inertia moments (3 nb) */
void rko_get_mass(rko_sim* s, double* mass, double* inertia);
/* All listed interactions (active == 0) or the active ones (active == 1),
 * lexicographic order (Interaction.hpp:104-122). Returns the count; writes at
 * most cap records. */
size_t rko_get_interactions(rko_sim* s, int active, rko_interaction* out, size_t cap);

/* Select the toofus floating-point forms (used only by the pinning tests). */
void rko_set_variant(int which, int value);
int rko_get_variant(int which);

#ifdef __cplusplus
}
#endif
#endif
