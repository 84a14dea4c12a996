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
#define ORC_MODEL_HOOKE_MEMORY 3
#define ORC_MODEL_HERTZ_MEMORY 4

/* Contacts a particle can remember at once (memory models). */
#define ORC_MEM_CAP 32

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
enum { ORC_WALL_BOX = 0, ORC_WALL_RECTANGLE = 1 };
/* Rectangle obstacle: sides lx, ly in its local xy plane, normal local z */
int orc_add_rectangle_wall(orc_sys* s, const double* basis, const double* centre, double lx,
                           double ly, double crust, int material);
int orc_add_wall(orc_sys* s, const double* basis, const double* centre,
                 const double* half_extent, double crust, int material);

/* kind HOOKE: p = {kn, en, etat, muc, kr}; kind HERTZ: p = {Es, en, Gs, muc, kr};
   kind HOOKE_MEMORY: p = {kn, en, etat, muc, kt, mur, etarpf, epst};
   kind HERTZ_MEMORY: p = {Es, en, Gs, muc, mur, etarpf, epst} */
int orc_set_contact_model(orc_sys* s, int mat_a, int mat_b, int kind,
                          const double* p);
/* Imposed wall motions (ObstacleImposedVelocity); p as for the library:
   1 constant {vx,vy,vz}; 2 sin {dx,dy,dz,amp,period,phase_deg};
   3 crenel {dx,dy,dz,amp,period}; 4 multisin {dx,dy,dz,ax,ay,az,px,py,pz,
   phix,phiy,phiz}; 5 constant rotation {wx,wy,wz,cx,cy,cz,about_centre}. */
#define ORC_MOTION_CONSTANT_TRANSLATION 1
#define ORC_MOTION_SIN_TRANSLATION 2
#define ORC_MOTION_CRENEL_TRANSLATION 3
#define ORC_MOTION_MULTISIN_TRANSLATION 4
#define ORC_MOTION_CONSTANT_ROTATION 5
#define ORC_MAX_MOTIONS 64
int orc_add_wall_motion(orc_sys* s, int wall, int type, double t_start, double t_end,
                        const double* p);
/* composite obstacles and box-element cylindrical shells (as gdem.h) */
#define ORC_MAX_COMPOSITES 64
int orc_add_composite(orc_sys* s, int parent, int* id);
int orc_attach_wall(orc_sys* s, int wall, int composite);
int orc_add_composite_motion(orc_sys* s, int composite, int type, double t_start,
                             double t_end, const double* p);
int orc_get_composite(const orc_sys* s, int composite, double* centre, double* vel,
                      double* omega);
int orc_set_motion_restriction(orc_sys* s, int wall, int composite, int mask);
void orc_cylindrical_shell_boxes(const double* basis, const double* centre, double height,
                                 double radius, double width, int n, double* wall_basis,
                                 double* wall_centre, double* wall_half);
int orc_add_cylindrical_shell(orc_sys* s, const double* basis, const double* centre,
                              double height, double radius, double width, int n, double crust,
                              int material, int parent, int* id);
/* the shell with Rectangle elements: half = (height/2, ly/2, 0) */
void orc_cylindrical_shell_rectangles(const double* basis, const double* centre, double height,
                                      double radius, double width, int n, double* wall_basis,
                                      double* wall_centre, double* wall_half);
int orc_add_cylindrical_shell_rectangles(orc_sys* s, const double* basis, const double* centre,
                                         double height, double radius, double width, int n,
                                         double crust, int material, int parent, int* id);
/* time (orc_step advances it); walls moved over [time - dt, time] */
int orc_set_time(orc_sys* s, double time);
double orc_get_time(const orc_sys* s);
int orc_move_walls(orc_sys* s, double time, double dt);
int orc_get_wall(const orc_sys* s, int wall, double* centre, double* basis, double* vel,
                 double* omega);
/* imposed force on a wall, p = {ax, ay, az, mass (<= 0: automatic), vmax} */
int orc_add_wall_force(orc_sys* s, int wall, double t_start, double t_end, const double* p);
/* the same on composite comp's load (mass must be > 0) */
int orc_add_composite_force(orc_sys* s, int comp, double t_start, double t_end, const double* p);
/* wall torsors of the last force evaluation (3 * nwalls each) */
int orc_get_wall_loads(const orc_sys* s, double* force, double* torque);
/* dt of the force evaluations (memory models); orc_step sets it too. */
int orc_set_time_step(orc_sys* s, double dt);
/* Contact memory of particle p (its own map): up to cap entries of key
   (partner id, or -1 - wall), kdelta, previous normal, spring torque. */
int orc_get_memory(const orc_sys* s, int p, int cap, int32_t* key, double* kd,
                   double* nrm, double* spr);
/* Quaternion::setRotFromTwoVectors + rotateVector, for the pins. */
void orc_rotate_two_vectors(const double* u, const double* v, const double* x,
                            double* out);

/* Phases of one leap-frog step, as in Grains::Simulation. */
int orc_link(orc_sys* s);                    /* initial LinkedCell::Link */
int orc_link_update(orc_sys* s);             /* LinkedCell::LinkUpdate */
int orc_compute_forces(orc_sys* s);          /* InitializeForces + ComputeForces */
int orc_move(orc_sys* s, double dt_vel, double dt_disp);   /* AllComponents::Move */
int orc_advance_velocity(orc_sys* s, double dt_vel);
/* Initial binning + a0 (Grains::do_before_time_stepping), then nsteps KDK. */
int orc_prime(orc_sys* s);
int orc_step(orc_sys* s, int nsteps, double dt);
/* orc_step on nthreads threads (OpenMP): cell rows split into slabs swept
   in two colours, each pair evaluated once as in orc_step (the summation
   order of a particle's contacts differs).  Memoryless models and static
   walls only; no contact log, no wall loads.  The all-core CPU baseline. */
int orc_step_parallel(orc_sys* s, int nsteps, double dt, int nthreads);

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
/* contact statistics of the run: max overlap, mean overlap, time of the max,
   mean GJK iterations, overlaps counted, GJK reports counted */
void orc_get_contact_stats(const orc_sys* s, double* out);
/* ClosestPoint_ErreurHandling of two spheres (closed form of its GJK) */
int orc_sphere_sphere_retry(const double* a, double ra, double cra, const double* b, double rb,
                            double crb, double* pt, double* ov, double* dist, int* nbIter);

/* ---- Stand-alone kernels of the restatement (used by the unit tests) ---- */
/* returns 1 contact, 0 no contact, -1 ContactError */
int orc_sphere_sphere(const double* a, double ra, double cra, const double* b,
                      double rb, double crb, double* point, double* ovec,
                      double* dist);
/* sphere (row-major accumulated basis) against a rectangle: 1 contact, 0 none */
int orc_sphere_rectangle(const double* centre, double r, const double* sphere_basis,
                         const double* basis, const double* origin, double lx, double ly,
                         double* point, double* ovec, double* dist);
int orc_sphere_box(const double* centre, double r, double cr,
                   const double* basis, const double* origin,
                   const double* half_extent, double crb, double* point,
                   double* ovec, double* dist);
/* Force of one contact on body 0 (body 1 receives the opposite).  kind as above;
   body b: x, v, w (world), m, req.  moment (may be NULL) receives the
   rolling-resistance moment on body 0 (body 1 receives the opposite). */
void orc_contact_force(int kind, const double* p, const double* point,
                       const double* ovec, double dist, const double* x0,
                       const double* v0, const double* w0, double m0,
                       double req0, const double* x1, const double* v1,
                       const double* w1, double m1, double req1, double* f,
                       double* moment);

/* ---- Convex particles (boxes, polyhedra) with the Johnson GJK ---------- */
#define ORC_SHAPE_SPHERE 0
#define ORC_SHAPE_BOX 1
#define ORC_SHAPE_POLYHEDRON 2

typedef struct orc_csys orc_csys;

int orc_c_create(const orc_config* cfg, orc_csys** out);
void orc_c_destroy(orc_csys* s);
/* returns the shape index, or -1 */
int orc_c_add_box_shape(orc_csys* s, const double* half_extent);
int orc_c_add_sphere_shape(orc_csys* s, double r);
int orc_c_add_polyhedron_shape(orc_csys* s, int nv, const double* verts, int nf,
                               const int32_t* face_size, const int32_t* face_idx);
int orc_c_shape_info(const orc_csys* s, int shape, double* vol, double* inertia,
                     double* circ, double* bbox_ext);
/* basis: row-major rotation of each particle's transform */
int orc_c_set_particles(orc_csys* s, int n, const double* x, const double* v,
                        const double* w, const double* q, const double* basis,
                        const int32_t* shape, const double* density,
                        const double* crust, const int32_t* material);
int orc_c_add_rectangle_wall(orc_csys* s, const double* basis, const double* centre, double lx,
                             double ly, double crust, int material);
int orc_c_add_wall(orc_csys* s, const double* basis, const double* centre,
                   const double* half_extent, double crust, int material);
int orc_c_set_contact_model(orc_csys* s, int mat_a, int mat_b, int kind,
                            const double* p);
int orc_c_prime(orc_csys* s);
int orc_c_step(orc_csys* s, int nsteps, double dt);
int orc_c_get_state(const orc_csys* s, double* x, double* v, double* w,
                    double* q, double* basis, double* F, double* T);
void orc_c_enable_contact_log(orc_csys* s, int capacity);
int orc_c_get_contacts(const orc_csys* s, int capacity, int* n_out, int32_t* id0,
                       int32_t* id1, double* point, double* overlap_vec,
                       double* overlap, double* force);
int orc_c_last_error(const orc_csys* s, int* code, int32_t* id0, int32_t* id1);
int64_t orc_c_gjk_calls(const orc_csys* s);
/* AppCollision contact features of the run (as orc_get_contact_stats). */
void orc_c_get_contact_stats(const orc_csys* s, double* out);
/* one pair / one particle-wall contact on the current state, no accumulation:
   1 contact (point, overlap vector, overlap, force on the first body), 0 none,
   -1 ContactError */
int orc_c_pair(orc_csys* s, int p0, int p1, const double* shift, double* point,
               double* ovec, double* overlap, double* force);
int orc_c_wall_pair(orc_csys* s, int p, int wall, double* point, double* ovec,
                    double* overlap, double* force);
/* per-particle Polyhedron support history (m_curr_vertex) on / off (after
   orc_c_set_particles; off by default: every support call starts from the
   first vertex with neighbours) */
int orc_c_set_support_history(orc_csys* s, int on);
/* one support call of a shape from a given start vertex (-1: the default) */
int orc_c_support(const orc_csys* s, int shape, const double* dir, int start, double* point,
                  int* vertex);
/* narrow phase of the convex contacts: 0 Johnson GJK (default), 1 GJK_SV */
int orc_c_set_narrow_phase(orc_csys* s, int sv);
/* GJK between two shapes of this system under crust-scaled transforms:
   world witness points and distance */
int orc_c_closest_points(const orc_csys* s, int shape_a, const double* basis_a,
                         const double* origin_a, const double* scale_a, int shape_b,
                         const double* basis_b, const double* origin_b,
                         const double* scale_b, double* pa, double* pb, int* nbIter,
                         double* dist);

#ifdef __cplusplus
}
#endif
#endif
