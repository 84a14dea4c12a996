/*
 * gdem_oracle.h -- CPU restatement of the Grains3D sphere DEM time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
 * leg, always as the checker or as the timed CPU baseline, never as part of the
 * product path.  The product (pacific_amd/, libgdem.so) never links or calls it.
 *
 * Each function follows a reference function operation for operation (same
 * operand order, no FMA contraction: built with -ffp-contract=off); the
 * citations are given in gdem_oracle.c next to each function.
 *
 * Deliberate restatement choices (identical in the HIP path):
 *   - periodic boundaries use the minimum image on a grid without ghost
 *     layers instead of the reference's clone particles; the periodic partner
 *     is always the second body of the pair and is shifted by +/- L;
 *   - positions leaving a periodic box are wrapped by -/+ L right after the
 *     drift; a cell index rounded to n is clamped to n-1;
 *   - walls are tested against every particle (the reference's per-cell wall
 *     lists are a conservative broad-phase filter and select the same set);
 *   - a ContactError is reported, not retried with a ten-fold crust.
 */
#ifndef GDEM_ORACLE_H
#define GDEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_ERR_CONTACT 1
#define ORC_ERR_MOTION 2
#define ORC_ERR_SIMULATION 3
#define ORC_ERR_ARG 5

#define ORC_MAX_MATERIALS 32
#define ORC_MAX_WALLS 64

#define ORC_MODEL_NONE 0
#define ORC_MODEL_HOOKE 1
#define ORC_MODEL_HERTZ 2

typedef struct orc_sys orc_sys;

typedef struct {
  double origin[3];      /* domain origin */
  double size[3];        /* domain size */
  int periodic[3];       /* 1 = periodic in that direction */
  double cell_size;      /* requested edge: 2 * CellSizeFactor * max radius */
  double gravity[3];
} orc_config;

int orc_create(const orc_config* cfg, orc_sys** out);
void orc_destroy(orc_sys* s);

/* Particles (spheres).  q is (x, y, z, w).  w is the world angular velocity. */
int orc_set_particles(orc_sys* s, int n, const double* x, const double* v,
                      const double* w, const double* q, const double* radius,
                      const double* density, const double* crust,
                      const int32_t* material);

/* Box wall obstacle: basis (row-major 3x3 rotation), centre, half extents. */
int orc_add_wall(orc_sys* s, const double* basis, const double* centre,
                 const double* half_extent, double crust, int material);

/* kind HOOKE: p = {kn, en, etat, muc, kr}; kind HERTZ: p = {Es, en, Gs, muc, kr} */
int orc_set_contact_model(orc_sys* s, int mat_a, int mat_b, int kind,
                          const double* p);

/* Phases of one leap-frog step, as in Grains::Simulation. */
int orc_link(orc_sys* s);                    /* initial LinkedCell::Link */
int orc_link_update(orc_sys* s);             /* LinkedCell::LinkUpdate */
int orc_compute_forces(orc_sys* s);          /* InitializeForces + ComputeForces */
int orc_move(orc_sys* s, double dt_vel, double dt_disp);   /* AllComponents::Move */
int orc_advance_velocity(orc_sys* s, double dt_vel);
/* Initial binning + a0 (Grains::do_before_time_stepping), then nsteps KDK. */
int orc_prime(orc_sys* s);
int orc_step(orc_sys* s, int nsteps, double dt);

int orc_get_state(const orc_sys* s, double* x, double* v, double* w,
                  double* q, double* F, double* T);
int orc_set_state(orc_sys* s, const double* x, const double* v,
                  const double* w, const double* q, const double* F,
                  const double* T);
int orc_get_grid(const orc_sys* s, int* nc, double* edge, double* origin);

/* Contact log of the last orc_compute_forces call (record must be enabled). */
void orc_enable_contact_log(orc_sys* s, int capacity);
int orc_get_contacts(const orc_sys* s, int capacity, int* n_out, int32_t* id0,
                     int32_t* id1, double* point, double* overlap_vec,
                     double* overlap, double* force);
int orc_last_error(const orc_sys* s, int* code, int32_t* id0, int32_t* id1,
                   double* value, double* limit);
/* Candidate pairs tested in the last ComputeForces (statistics). */
int64_t orc_candidates(const orc_sys* s);

/* ---- Stand-alone kernels of the restatement (used by the unit tests) ---- */
/* returns 1 contact, 0 no contact, -1 ContactError */
int orc_sphere_sphere(const double* a, double ra, double cra, const double* b,
                      double rb, double crb, double* point, double* ovec,
                      double* dist);
int orc_sphere_box(const double* centre, double r, double cr,
                   const double* basis, const double* origin,
                   const double* half_extent, double crb, double* point,
                   double* ovec, double* dist);
/* Force of one contact on body 0 (body 1 receives the opposite).  kind as above;
   body b: x, v, w (world), m, req. */
void orc_contact_force(int kind, const double* p, const double* point,
                       const double* ovec, double dist, const double* x0,
                       const double* v0, const double* w0, double m0,
                       double req0, const double* x1, const double* v1,
                       const double* w1, double m1, double req1, double* f);

#ifdef __cplusplus
}
#endif
#endif
