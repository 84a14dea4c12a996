/*
 * gdem.h -- C ABI of the MI355X-native Grains3D DEM time step (libgdem.so).
 *
 * One context = one HIP device + one stream + the device-resident particle
 * state.  The reference is object-per-particle C++ (Grains3D); its plugin seams
 * are batched here, one call per phase of Grains::Simulation
 * (src/Grains3D/Grains/Base/src/Grains.cpp:344-364):
 *
 *   reference call (src/Grains3D/Grains/...)            C entry point
 *   --------------------------------------------------  ------------------------
 *   LinkedCell::set (Base/include/LinkedCell.hh:249)    gdem_create
 *   AllComponents / Particle construction               gdem_set_particles
 *   ConvexBuilderFactory::create (Geometry/src/ConvexBuilderFactory.cpp:20-50)
 *     for Box / Polyhedron                              gdem_add_box_shape,
 *                                                        gdem_add_polyhedron_shape
 *   Particle construction with a convex shape            gdem_set_convex_particles
 *   LinkedCell::Link(Obstacle*) (LinkedCell.cpp:1371)   gdem_add_wall
 *   ContactBuilderFactory::define + Hooke/Hertz ctor    gdem_set_contact_model
 *     (Component/src/ContactBuilderFactory.cpp:81-192)
 *   AppCollision::LinkUpdate (Base/include/AppCollision.hh:90)
 *                                                        gdem_link_update
 *   App::ComputeForces (Base/include/App.hh:38-39) with  gdem_compute_forces
 *     AllComponents::InitializeForces and every
 *     ContactForceModel::computeForces
 *     (Component/include/ContactForceModel.hh:66-68)
 *   TimeIntegrator::Move via AllComponents::Move         gdem_move
 *     (TimeIntegration/include/TimeIntegrator.hh:47-57)
 *   TimeIntegrator::advanceVelocity                      gdem_advance_velocity
 *     (TimeIntegrator.hh:59-65)
 *   Grains::do_before_time_stepping a0 (Grains.cpp:169-174)
 *                                                        gdem_prime
 *   Grains::Simulation loop body (Grains.cpp:286-452)   gdem_step
 *   Particle / kinematics getters                        gdem_get_state
 *   C-ABI precedent: InterfaceCBasilisk/include/InterfaceGrains3DBasilisk.h:11-33
 *
 * Conventions
 *   - Return 0 on success, else one of GDEM_ERR_*; the exception the reference
 *     would throw maps to CONTACT (ContactError, RigidBodyWithCrust.cpp:624-630,
 *     raised as the reference raises it: a sphere pair or sphere-wall overlap
 *     that reaches the crust sum is first retried as
 *     ClosestPoint_ErreurHandling does, RigidBodyWithCrust.cpp:534-588, with an
 *     imposed overlap equal to the crust sum, and only an overlap the bodies
 *     shrunk by ten crusts still have is an error, for spheres and convex
 *     bodies; the differentiated step raises it at the crust sum),
 *     MOTION (MotionError, Particle.cpp:447-457) or SIMULATION
 *     (SimulationError, LinkedCell.cpp:1506-1531).  gdem_last_error names the
 *     offending ids (smallest particle id among the offenders; a wall is -1-k).
 *   - Host pointers are borrowed for the duration of the call; the context owns
 *     all device memory.  Particle arrays are in particle-id order; vectors are
 *     xyz-interleaved (3 doubles per particle), quaternions (x, y, z, w).
 *   - Calls are ordered on the context's stream.  gdem_step, gdem_move,
 *     gdem_advance_velocity and gdem_compute_forces return without waiting;
 *     errors raised on the device are reported by the next synchronizing call
 *     (gdem_synchronize, gdem_get_*).  A context is not thread-safe.
 *   - Arithmetic is IEEE FP64 without FMA contraction; the result of a call
 *     does not depend on launch shape or run (no floating-point atomics).
 */
#ifndef GDEM_H
#define GDEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDEM_OK 0
#define GDEM_ERR_CONTACT 1
#define GDEM_ERR_MOTION 2
#define GDEM_ERR_SIMULATION 3
#define GDEM_ERR_HIP 4
#define GDEM_ERR_ARG 5
#define GDEM_ERR_CAPACITY 6

#define GDEM_MODEL_HOOKE 1
#define GDEM_MODEL_HERTZ 2
#define GDEM_MODEL_HOOKE_MEMORY 3
#define GDEM_MODEL_HERTZ_MEMORY 4

/* Imposed wall motions (ObstacleImposedVelocity types) */
#define GDEM_MOTION_CONSTANT_TRANSLATION 1
#define GDEM_MOTION_SIN_TRANSLATION 2
#define GDEM_MOTION_CRENEL_TRANSLATION 3
#define GDEM_MOTION_MULTISIN_TRANSLATION 4
#define GDEM_MOTION_CONSTANT_ROTATION 5

#define GDEM_MAX_MATERIALS 32
#define GDEM_MAX_WALLS 64
#define GDEM_MAX_SHAPES 16
#define GDEM_MAX_VERTICES 8

typedef struct gdem_ctx gdem_ctx;

typedef struct {
  double origin[3];     /* domain origin (Grains.cpp:545-579) */
  double size[3];       /* domain size */
  int32_t periodic[3];  /* 1 = periodic (needs >= 3 cells in that direction) */
  double cell_size;     /* requested cell edge = 2 * CellSizeFactor * max radius
                           (Grains.cpp:1013-1022, 2522-2526) */
  double gravity[3];    /* GrainsExec::m_vgravity */
  int32_t device;       /* HIP device ordinal */
  int32_t reserved;
} gdem_config;

/* LinkedCell::set (Base/src/LinkedCell.cpp:47-419) + the context. */
int gdem_create(const gdem_config* cfg, gdem_ctx** out);
int gdem_destroy(gdem_ctx* ctx);

/* Spheres (Particle(RigidBodyWithCrust*, density, mat, type), Particle.cpp:179-232;
   kinematics setters ParticleKinematics.cpp:170-222).  w is the world angular
   velocity; mass, inertia and equivalent radius are derived as Particle.cpp:158-164.
   Replaces any previous particle set. */
int gdem_set_particles(gdem_ctx* ctx, int64_t n, const double* x, const double* v,
                       const double* w, const double* q, const double* radius,
                       const double* density, const double* crust,
                       const int32_t* material);

/* Convex shapes, registered before gdem_set_convex_particles; shape ids count
   from 0 in registration order.
   Box(LX, LY, LZ) (Geometry/src/Box.cpp:27-33; half extents here): volume,
   inertia and circumscribed radius as Box.cpp:181-229, 273-279. */
int gdem_add_box_shape(gdem_ctx* ctx, const double* half_extent, int32_t* shape_id);

/* Polyhedron from vertices and faces (Geometry/src/Polyhedron.cpp:324-365):
   vertex adjacency from the face loops, volume and inertia from a fan of
   tetrahedra (GrainsExec::computeVolumeInertiaContrib,
   Base/src/GrainsExec.cpp:748-800), vertices about the centre of mass.
   nv <= GDEM_MAX_VERTICES; face k has face_size[k] vertex indices, listed
   consecutively in face_index. */
int gdem_add_polyhedron_shape(gdem_ctx* ctx, int32_t nv, const double* vertices,
                              int32_t nfaces, const int32_t* face_size,
                              const int32_t* face_index, int32_t* shape_id);

/* Convex particles (Particle(DOMNode*), Particle.cpp:179-232, with a
   RigidBodyWithCrust over a registered shape): basis is the row-major
   rotation of each particle's Transform (9 doubles), consistent with q.
   Mass = density * V, inertia = density * I_shape, equivalent radius from V,
   crust scaling per axis (extent - crust) / extent (RigidBodyWithCrust.cpp:51-80).
   Replaces any previous particle set; the step then runs the GJK narrow phase
   (Geometry/src/Convex.cpp:536-627) for every pair and wall. */
int gdem_set_convex_particles(gdem_ctx* ctx, int64_t n, const double* x, const double* v,
                              const double* w, const double* q, const double* basis,
                              const int32_t* shape, const double* density,
                              const double* crust, const int32_t* material);

/* Row-major basis of every convex particle in id order (RigidBody::getTransform,
   Geometry/include/RigidBody.hh:167). */
int gdem_get_basis(gdem_ctx* ctx, double* basis);

/* Static box obstacle (SimpleObstacle with a Box convex; Box(LX,LY,LZ) uses
   full edges, here half extents).  basis is the row-major rotation of its
   Transform, centre its origin. */
int gdem_add_wall(gdem_ctx* ctx, const double* basis, const double* centre,
                  const double* half_extent, double crust, int32_t material);

/* Static rectangle obstacle (SimpleObstacle with a Rectangle convex,
   Geometry/src/Rectangle.cpp:13-21: sides lx along the local x, ly along the
   local y, no thickness, normal along the local z).  Spheres meet it through
   ClosestPointRECTANGLE (Geometry/src/RigidBodyWithCrust.cpp:1107-1226): a
   contact only where the sphere's deepest point lies beyond the plane and its
   projection falls strictly inside the rectangle; no overlap limit.  Convex
   particles meet it the same way through their full shape's support point
   (the reference's branch for any convex first body). */
int gdem_add_rectangle_wall(gdem_ctx* ctx, const double* basis, const double* centre,
                            double lx, double ly, double crust, int32_t material);

/* HookeContactForceModel(map) (HookeContactForceModel.cpp:11-20):
     kind GDEM_MODEL_HOOKE, p = {kn, en, etat, muc, kr}            (5 values)
   HertzContactForceModel(map) (HertzContactForceModel.cpp:12-22):
     kind GDEM_MODEL_HERTZ, p = {Es, en, Gs, muc, kr}              (5 values)
   kr >= 0 is the rolling resistance of these two (the moment
   -(kr R* |FN| |w0 x (x0 - c) - w1 x (x1 - c)| / |w0 - w1|) (w0 - w1) on body
   0, HookeContactForceModel.cpp:95-111, HertzContactForceModel.cpp:99-113):
   sphere contexts; with kr != 0 gdem_step runs the canonical layout, and
   convex contexts and differentiated systems refuse it.
   HookeMemoryContactForceModel(map) (HookeMemoryContactForceModel.cpp:11-29):
     kind GDEM_MODEL_HOOKE_MEMORY,
     p = {kn, en, etat, muc, kt, mur, etarpf, epst}                 (8 values)
   HertzMemoryContactForceModel(map) (HertzMemoryContactForceModel.cpp:12-31):
     kind GDEM_MODEL_HERTZ_MEMORY,
     p = {Es, en, Gs, muc, mur, etarpf, epst}                       (7 values)
   The memory models keep a tangential spring, the previous contact normal and
   a rolling-friction spring torque per contact while it lasts (the contact
   maps of Component.cpp:567-689).  Their defaults as the reference's XML
   reader leaves them: etat -1, Gs Es/4, mur 0, etarpf 0, and epst 0 (the
   reader stores its 1e-10 default under another key,
   HookeMemoryContactForceModel.cpp:300-306).  Sphere contexts only. */
/* Imposed velocity of wall k over [t_start, t_end]
   (ObstacleImposedVelocity, Component/src/ObstacleImposedVelocity.cpp:54-257,
   moved as AllComponents::Move / SimpleObstacle::Move do,
   AllComponents.cpp:151-220, SimpleObstacle.cpp:119-175).  Parameters:
     CONSTANT_TRANSLATION  p = {vx, vy, vz}
     SIN_TRANSLATION       p = {dx, dy, dz, amplitude, period, phase_degrees}
     CRENEL_TRANSLATION    p = {dx, dy, dz, amplitude, period}
     MULTISIN_TRANSLATION  p = {dx, dy, dz, ax, ay, az, px, py, pz, phix, phiy, phiz}
     CONSTANT_ROTATION     p = {wx, wy, wz, cx, cy, cz, about_centre}
   (direction d normalised; rotation about the wall's centre when
   about_centre != 0, else about (cx, cy, cz)).  Each step moves the walls
   over [t - dt, t] and gives them the summed velocities, which the contact
   law sees (Obstacle::getVelocityAtPoint, Obstacle.cpp:128-133).  Sphere
   contexts; graph replay is off while walls move. */
int gdem_add_wall_motion(gdem_ctx* ctx, int32_t wall, int32_t type, double t_start,
                         double t_end, const double* p);
/* Composite obstacles (CompositeObstacle, Component/src/CompositeObstacle.cpp):
   a group of walls (and composites) moving as one.  gdem_add_composite
   creates an empty composite under parent (-1: the root) and returns its
   index; gdem_attach_wall makes wall k a member (once).  A composite's centre
   is the volume-weighted centre of its members (computeVolumeCenterOfMass,
   :228-255).  Its imposed motions (gdem_add_composite_motion, same types and
   parameters as gdem_add_wall_motion) move its centre and are composed onto
   every member at its lever from the centre (CompositeObstacle::Move,
   :135-220; ObstacleKinematicsVelocity::Compose, :62-77), on top of the
   members' own motions.  Sphere contexts; at most 64 composites. */
int gdem_add_composite(gdem_ctx* ctx, int32_t parent, int32_t* composite);
int gdem_attach_wall(gdem_ctx* ctx, int32_t wall, int32_t composite);
int gdem_add_composite_motion(gdem_ctx* ctx, int32_t composite, int32_t type, double t_start,
                              double t_end, const double* p);
/* Obstacle::setRestrictedGeomDirMotion (Component/src/Obstacle.cpp:613-617):
   the geometric translation of wall k (composite < 0) or of a composite
   (wall < 0) is zeroed along the directions of mask (bit 0 x, 1 y, 2 z); its
   velocity is kept (SimpleObstacle.cpp:131-136, 146-149;
   CompositeObstacle.cpp:153-160). */
int gdem_set_motion_restriction(gdem_ctx* ctx, int32_t wall, int32_t composite, int32_t mask);
/* Centre and velocities of a composite (any pointer may be NULL). */
int gdem_get_composite(gdem_ctx* ctx, int32_t composite, double* centre, double* vel,
                       double* omega);
/* CylindricalShell with box elements (Component/src/CylindricalShell.cpp:13-93):
   a composite of n walls, box i of edges (width, ly, height) centred at
   radius + width/2 from the shell axis at angle (i + 1/2) 2 pi / n, with
   ly = 2 (radius + width)(sin d - (1 - cos d)/tan d), d = 2 pi / n, placed by
   the shell's transform (row-major basis, centre: its axis is the basis' z).
   The walls are appended in order i; *composite receives the shell's index. */
int gdem_add_cylindrical_shell(gdem_ctx* ctx, const double* basis, const double* centre,
                               double height, double radius, double width, int32_t n,
                               double crust, int32_t material, int32_t parent,
                               int32_t* composite);
/* CylindricalShell with its default Rectangle elements (the same file
   without ElementaryObstacle="Box"): rectangle i of sides (height, ly),
   centred at the inner radius, its local x along the axis and its normal
   towards it. */
int gdem_add_cylindrical_shell_rectangles(gdem_ctx* ctx, const double* basis,
                                          const double* centre, double height, double radius,
                                          double width, int32_t n, double crust,
                                          int32_t material, int32_t parent, int32_t* composite);
/* The shell's rectangles alone (no context): basis (9n), centre (3n),
   half sides (3n: height/2, ly/2, 0). */
int gdem_cylindrical_shell_rectangles(const double* basis, const double* centre, double height,
                                      double radius, double width, int32_t n,
                                      double* wall_basis, double* wall_centre, double* wall_half);
/* The shell's boxes alone (no context): basis (9n), centre (3n), half edges (3n). */
int gdem_cylindrical_shell_boxes(const double* basis, const double* centre, double height,
                                 double radius, double width, int32_t n, double* wall_basis,
                                 double* wall_centre, double* wall_half);

/* Loads on the walls (AllComponents::computeObstaclesLoad,
   Component/src/AllComponents.cpp:1685-1693): with tracking on, every force
   evaluation also sums, per wall, the reactions of its contacts (force, and
   torque about the wall's centre, rolling torques included; Obstacle::addForce /
   addTorque, Component/src/Obstacle.cpp:575-597), in a fixed order without
   atomics.  gdem_get_wall_loads returns those of the last evaluation,
   3 * nwalls values each (either pointer may be NULL).  Sphere contexts.
   Switching on what is already on keeps the loads.  Switched on after
   gdem_prime and before the first step, the priming evaluation is run again
   to form its loads; later, with a memory contact model, or after state was
   set with F and T, that is refused (GDEM_ERR_ARG): switch loads on first. */
int gdem_set_wall_loads(gdem_ctx* ctx, int32_t on);
int gdem_get_wall_loads(gdem_ctx* ctx, double* force, double* torque);

/* Imposed force on wall k over [t_start, t_end] (ObstacleImposedForce,
   ConstantTranslation, Component/src/ObstacleImposedForce.cpp:30-90, 227-296):
   p = {ax, ay, az, mass, vmax}.  Each step the wall moves so that its load
   (the previous force evaluation's) tends to (ax, ay, az): a PID law on the
   load error along that direction with gains from mass (mass <= 0: chosen from
   dt, |a| and the wall's crust, as the reference's automatic mass); until
   the first nonzero load it moves at vmax against that direction.  As in the
   reference, the controller's memory is shared by all imposed forces.  Turns
   wall-load tracking on; gdem_step then advances one step per host round
   trip while the force is pending.  Sphere contexts. */
int gdem_add_wall_force(gdem_ctx* ctx, int32_t wall, double t_start, double t_end,
                        const double* p);
/* Imposed force on a composite (CompositeObstacle::LinkImposedMotion(
   ObstacleImposedForce*) and Move, Component/src/CompositeObstacle.cpp:
   112-120, 182-203): the same law on the composite's load, the sum of its
   members' (CompositeObstacle::getTorsor, :650-660, as left after the
   previous step: zero at the first step); its translation moves the
   composite's centre and is added to every member's.  mass must be > 0 (the
   automatic mass would divide by the composite's zero crust). */
int gdem_add_composite_force(gdem_ctx* ctx, int32_t composite, double t_start, double t_end,
                             const double* p);

/* Simulation time (Grains::m_time, Base/src/Grains.cpp:114, :290); gdem_step
   advances it by dt per step. */
int gdem_set_time(gdem_ctx* ctx, double time);
int gdem_get_time(gdem_ctx* ctx, double* time);
/* The obstacle part of AllComponents::Move (Component/src/AllComponents.cpp:
   151-220): sets the time and moves the walls over [time - dt, time], for
   callers that drive the phases themselves (gdem_move, then this, then
   gdem_compute_forces).  gdem_prime sets the initial wall velocities
   (AllComponents::setKinematicsObstacleWithoutMoving, :235-251) once the
   time step is known. */
int gdem_move_walls(gdem_ctx* ctx, double time, double dt);
/* Current transform and velocity of wall k (any pointer may be NULL). */
int gdem_get_wall(gdem_ctx* ctx, int32_t wall, double* centre, double* basis, double* vel,
                  double* omega);

/* Time step of the force evaluations for the memory models (the dt that
   Grains::computeParticlesForceAndTorque passes to App::ComputeForces,
   Grains.cpp:459-482).  gdem_step(dt) and gdem_move(., dt) set it; a force
   evaluation with a memory model and no time step set fails with
   GDEM_ERR_ARG. */
int gdem_set_time_step(gdem_ctx* ctx, double dt);

int gdem_set_contact_model(gdem_ctx* ctx, int32_t mat_a, int32_t mat_b, int32_t kind,
                           const double* p);

/* AppCollision::LinkUpdate -> LinkedCell::LinkUpdate (LinkedCell.cpp:1452-1547):
   counting sort of the particles by Morton cell key. */
int gdem_link_update(gdem_ctx* ctx);

/* Grains::computeParticlesForceAndTorque (Grains.cpp:459-482): weight, then
   LinkedCell::ComputeForces (LinkedCell.cpp:1019-1103) with every contact
   force (Hooke/Hertz + Coulomb) accumulated per particle in a fixed order.
   Re-bins first if positions changed since the last binning. */
int gdem_compute_forces(gdem_ctx* ctx);

/* AllComponents::Move (AllComponents.cpp:151-166) -> ParticleKinematics::Move
   (ParticleKinematics.cpp:62-104) -> SecondOrderLeapFrog::Move
   (SecondOrderLeapFrog.cpp:44-67). */
int gdem_move(gdem_ctx* ctx, double dt_vel, double dt_disp);

/* AllComponents::advanceParticlesVelocity (AllComponents.cpp:224-231) ->
   SecondOrderLeapFrog::advanceVelocity (SecondOrderLeapFrog.cpp:74-94). */
int gdem_advance_velocity(gdem_ctx* ctx, double dt_vel);

/* Link + a0 (Grains::do_before_time_stepping, Grains.cpp:169-174). */
int gdem_prime(gdem_ctx* ctx);

/* nsteps of the kick-drift-kick loop (Grains.cpp:344-364), fused: each step is
   one binning pass plus one force kernel whose tail applies the closing
   half-kick and the next step's kick and drift.  Primes first if needed. */
int gdem_step(gdem_ctx* ctx, int32_t nsteps, double dt);

/* Replay of gdem_step's fused steps from a hipGraph captured on first use
   (one graph per state-buffer parity and dt; default on).  Steps run
   eagerly while contact recording or kernel timing is on.  Results are
   bit-identical either way.  No counterpart in the reference (its loop is
   host code, Grains.cpp:286-452). */
int gdem_set_graph(gdem_ctx* ctx, int32_t on);

/* Skin of the sphere candidate lists (default: automatic, 0.2 r_max; 0:
   lists rebuilt at every force evaluation).  Binning and the half-stencil
   sweep run only when a particle has moved skin/2 since the last rebuild;
   every evaluation still tests each candidate for contact exactly, so the
   pair set of each step is the reference's (LinkedCell::ComputeForces,
   Base/src/LinkedCell.cpp:1019-1103).  The skin is limited so that
   2 r_max + skin fits in a cell edge, and the lists are off when that
   leaves less than 0.05 r_max.  Negative: automatic. */
int gdem_set_skin(gdem_ctx* ctx, double skin);
/* Effective skin (0: lists off) and the rebuilds of the candidate lists since
   the particles or the skin were last set.  No counterpart in the reference,
   whose lists are rebuilt at every step. */
int gdem_get_skin(gdem_ctx* ctx, double* skin, int64_t* rebuilds);
/* Capacity of the lists, which the reference (std::list, unbounded) does not
   have.  Each particle's lists hold 24 entries per direction (convex: 32
   candidates); a build that overflows raises GDEM_ERR_CAPACITY.  With skin
   lists on, gdem_step ends by measuring the longest list of the last build
   and, when it comes within 4 entries of the capacity, halves the skin
   (switching the lists off below 0.05 r_max) and rebuilds at the next
   evaluation, so a packing that densifies keeps running; the results do not
   depend on the skin.  The check synchronises the context's stream.
   gdem_set_skin_adapt(ctx, 0) switches it off (on by default).
   gdem_get_list_occupancy gives the longest list of the last build (skin
   lists: candidates within the skin; otherwise the outgoing contacts of the
   last evaluation; convex: circumscribed-sphere candidates), the capacity and
   the number of halvings so far. */
int gdem_set_skin_adapt(gdem_ctx* ctx, int32_t on);
int gdem_get_list_occupancy(gdem_ctx* ctx, int32_t* peak, int32_t* capacity,
                            int32_t* skin_halvings);

/* Waits for the stream; returns the first deferred device error. */
int gdem_synchronize(gdem_ctx* ctx);

/* Particle state in id order (any pointer may be NULL).  F and T are the
   torsor of the last force evaluation (Torsor, Component/src/Torsor.cpp). */
int gdem_get_state(gdem_ctx* ctx, double* x, double* v, double* w, double* q,
                   double* F, double* T);

/* Restores particle state in id order (any pointer may be NULL: kept), as
   the reference's reload does (Particle::read2014 / read2014_binary,
   Component/src/Particle.cpp:973-1081; ParticleKinematics::
   readParticleKinematics2014, ParticleKinematics.cpp:378-406: the body-frame
   angular velocity follows from w and q).  F and T go together: given, they
   become the leap-frog's a0 and the next gdem_step starts from them without a
   force evaluation (SecondOrderLeapFrog::readParticleKinematics2014,
   SecondOrderLeapFrog.cpp:129-150).  Sphere contexts. */
int gdem_set_state(gdem_ctx* ctx, const double* x, const double* v, const double* w,
                   const double* q, const double* F, const double* T);

/* Contact memory of the history models (Component::m_contactMap,
   Component.hh:505-508): every entry, sorted by (owner id, key); key is the
   partner's id, or -1 - k for wall k; kd, nrm, spr are the owner's view of the
   tangential spring, previous normal and rolling spring torque (3 doubles
   each).  Writes min(count, capacity) entries; *n_out = count. */
int gdem_get_memory(gdem_ctx* ctx, int64_t capacity, int64_t* n_out, int32_t* owner,
                    int32_t* key, double* kd, double* nrm, double* spr);
/* Replaces the contact memory by count entries as above
   (Component::copyContactInMap via readContactMap2014, Component.cpp:613-624,
   858-916).  Needs a memory model to be set. */
int gdem_set_memory(gdem_ctx* ctx, int64_t count, const int32_t* owner, const int32_t* key,
                    const double* kd, const double* nrm, const double* spr);

/* Particle count, walls, convex flag, memory models in use. */
int gdem_info(gdem_ctx* ctx, int64_t* n, int32_t* nwalls, int32_t* convex,
              int32_t* memory_models);

/* Particle records of a reload file in the reference's 2014 format
   (AllComponents::write, Component/src/AllComponents.cpp:1148-1181;
   Particle::write2014 / write2014_binary, Particle.cpp:1168-1199): one record
   per particle in id order with geometric type geom_type[id] (NULL: 0), id
   = id + id_offset, tag 0, activity 1, transform (origin = x, basis = R(q)
   for spheres or the context's basis for convex bodies, type), kinematics
   (v, q as w then x, y, z, angular velocity), the leap-frog's F and T of
   the last force evaluation and the contact memory (keys (1, partner id +
   id_offset, 1), walls (1, wall_id_offset + k, 1)).  binary != 0: the
   hybrid "_Particles.result.bin" file (each record preceded by its byte count
   as a size_t); else the text "_Particles.result" file. */
int gdem_write_particles_2014(gdem_ctx* ctx, const char* path, int32_t binary,
                              int32_t id_offset, int32_t wall_id_offset,
                              const int32_t* geom_type);
/* Reads such a file into the context (its particles already set, in the same
   number, with their radius, density, crust and material, e.g. from the
   reload file's reference particles): record ids map to particle id - id_offset,
   memory keys to particles or to walls (wall_id_offset + k); positions,
   orientations, velocities, F, T (the context is then primed) and contact
   memory are restored.  geom_type_out[id] receives the records' types (may be
   NULL).  Sphere contexts. */
int gdem_read_particles_2014(gdem_ctx* ctx, const char* path, int32_t binary,
                             int32_t id_offset, int32_t wall_id_offset,
                             int32_t* geom_type_out);

/* The particles' constants as gdem_set_particles received them, in id order
   (any pointer may be NULL).  Sphere contexts. */
int gdem_get_particles(gdem_ctx* ctx, double* radius, double* density, double* crust,
                       int32_t* material);
/* The contact model of material pair (a, b): kind (0: none) and its
   parameters as gdem_set_contact_model took them (p: 8 doubles, unused ones
   0). */
int gdem_get_contact_model(gdem_ctx* ctx, int32_t a, int32_t b, int32_t* kind, double* p);
/* Wall k as added (basis row-major, centre, half extents, crust, material,
   kind GDEM_WALL_BOX or GDEM_WALL_RECTANGLE); any pointer may be NULL. */
#define GDEM_WALL_BOX 0
#define GDEM_WALL_RECTANGLE 1
int gdem_get_wall_shape(gdem_ctx* ctx, int32_t k, double* basis, double* centre, double* half,
                        double* crust, int32_t* material, int32_t* kind);
/* Names of materials and walls as the reload files carry them (the
   reference's material names, ContactBuilderFactory::m_materials,
   Component/src/ContactBuilderFactory.cpp:174-187, and obstacle names,
   Obstacle::m_name).  A name not set reads as "".  get_*: GDEM_ERR_CAPACITY
   when cap bytes do not hold the name and its terminating zero. */
int gdem_set_material_name(gdem_ctx* ctx, int32_t k, const char* name);
int gdem_get_material_name(gdem_ctx* ctx, int32_t k, char* buf, int32_t cap);
int gdem_set_wall_name(gdem_ctx* ctx, int32_t k, const char* name);
int gdem_get_wall_name(gdem_ctx* ctx, int32_t k, char* buf, int32_t cap);

/* A reload file set in the reference's format, as Grains::saveReload writes
   it (Base/src/Grains.cpp:2543-2581) for a sphere context:
     <root><suffix>.result          main file: "__Format2024__", "#Time t",
                                    the contact-model file, the reference
                                    particle types (Particle::write,
                                    Component/src/Particle.cpp:1093-1149), the
                                    obstacle tree (a "__AllObstacles___"
                                    composite of box walls, SimpleObstacle::
                                    write, Component/src/SimpleObstacle.cpp:354-
                                    367), the particle-file count and mode
                                    (AllComponents::write, Component/src/
                                    AllComponents.cpp:1062-1146), "#EndTime";
     <root>_MatContact.xml          materials and contact models
                                    (ContactBuilderFactory::save, Component/src/
                                    ContactBuilderFactory.cpp:270-325);
     <root><suffix>_Particles.result[.bin]   the particle records
                                    (gdem_write_particles_2014);
     <root>_RFTable.txt             "t <root><suffix>" appended.
   The reference particle types are the distinct (radius, density, crust,
   material) of the particles in order of first appearance; type k has
   IDnumber k.  Materials are named by gdem_set_material_name ("mat<k>" when
   unset) and take the values 1, 2, 4, ... in index order; walls are named by
   gdem_set_wall_name ("Obstacle<k>" when unset), obstacle k has id
   wall_id_offset + k and a zero torsor.  binary != 0: hybrid mode (binary
   particle file). */
int gdem_write_reload(gdem_ctx* ctx, const char* root, const char* suffix, double time,
                      int32_t binary, int32_t id_offset, int32_t wall_id_offset);
/* Restart from such a file set (main file path; the contact-model file and
   the particle file beside it, as Grains::readReload finds them): creates a
   context on cfg (the domain is not part of a reload file) with the
   materials (index = log2 of the value, names set), the contact models, the
   box walls (names set) and the particles (constants from their types, state,
   F, T and contact memory from the records), primed as the reference's reload
   leaves it.  *time receives the file's time (may be NULL).  Sphere particles,
   box and rectangle obstacles, one particle file. */
int gdem_read_reload(const gdem_config* cfg, const char* main_path, int32_t id_offset,
                     int32_t wall_id_offset, gdem_ctx** out, double* time);

/* Contact list of the next force evaluations, kept for the last one.
   capacity_per_particle bounds the contacts a particle can be the first body
   of; 0 disables recording.  Records are in canonical order: sorted by (id0,
   id1) with id0 the first body of the pair as the reference orders it
   (LinkedCell.cpp:1034-1060: same cell -> lower id first here; otherwise the
   body whose half stencil holds the other's cell).  A wall partner is -1-k. */
int gdem_set_contact_recording(gdem_ctx* ctx, int32_t capacity_per_particle);
int gdem_get_contacts(gdem_ctx* ctx, int64_t capacity, int64_t* n_out, int32_t* id0,
                      int32_t* id1, double* point, double* overlap_vec,
                      double* overlap, double* force);

int gdem_last_error(gdem_ctx* ctx, int32_t* code, int32_t* id0, int32_t* id1);

/* Counters since the last reset: force evaluations, contacts (pairs with
   overlap < 0, walls included) summed over them, and the event-timed total of
   the force kernel in ms when timing is on (events on the context's stream).
   In the packed layout of gdem_step (spheres with skin lists and memoryless
   models) the force kernel is the integrator kernel, which sums each
   particle's contact records and applies the leap-frog tail. */
int gdem_set_kernel_timing(gdem_ctx* ctx, int32_t on);
int gdem_get_counters(gdem_ctx* ctx, int64_t* force_evals, int64_t* contacts,
                      double* force_kernel_ms, int64_t* timed_launches);
/* Contact statistics of the run (AppCollision::addToContactsFeatures,
   Base/src/AppCollision.cpp:160-190, reported at the end of Grains::Simulation,
   Base/src/Grains.cpp:225-240): switched on (and reset) by
   gdem_set_contact_stats; every force evaluation then adds its contacts'
   overlaps (fixed-order workgroup reductions, reproducible) and the GJK
   iterations the closest-point routines report (1 for a sphere contact, 2 for
   a two-sphere contact-error retry, the Johnson GJK's count for convex
   bodies, none for rectangles and sphere-wall retries).
   gdem_get_contact_stats fills out[6] = {largest overlap, mean overlap, time
   of the largest overlap, mean GJK iterations, overlaps counted, GJK reports
   counted}.  Off by default (one extra kernel per evaluation when on; a
   sphere context then steps in its canonical layout). */
int gdem_set_contact_stats(gdem_ctx* ctx, int32_t on);
/* Narrow phase of convex contacts, the <CollisionDetection Type="GJK"|"GJK_SV">
   choice of the input file (Base/src/Grains.cpp:655-666, GrainsExec::
   m_colDetGJK_SV): GDEM_GJK, the Johnson sub-algorithm (closest_points,
   Geometry/src/Convex.cpp:536-627; the default), or GDEM_GJK_SV, the
   signed-volume one (closest_points_GJK_SV, Geometry/src/GJK_SV.cpp:428-524),
   both at the default tolerance EPSILON and without acceleration.  The
   contact-error retry uses the Johnson GJK either way, as the reference's
   ClosestPoint_ErreurHandling does (RigidBodyWithCrust.cpp:545-546).  Sphere
   contacts are closed forms and are not affected. */
#define GDEM_GJK 0
#define GDEM_GJK_SV 1
int gdem_set_narrow_phase(gdem_ctx* ctx, int32_t alg);
int gdem_get_narrow_phase(gdem_ctx* ctx, int32_t* alg);
int gdem_get_contact_stats(gdem_ctx* ctx, double* out);

/* Event-timed totals (ms) and launch counts of two kernels since the last
   reset: [0] the force kernel (as gdem_get_counters), [1] the pair kernel of
   the packed layout (screening and evaluation of each candidate pair once,
   with the candidate sweep on a rebuild).  No reference counterpart: the
   reference's phase timers (Grains.cpp:246-258) time whole phases on the CPU. */
int gdem_get_kernel_times(gdem_ctx* ctx, double* ms, int64_t* launches);
int gdem_reset_counters(gdem_ctx* ctx);

/* Grid actually used: cells per direction and edge lengths. */
int gdem_get_grid(gdem_ctx* ctx, int32_t* nc, double* edge);

const char* gdem_version(void);

/* ------------------------------------------------------------------------
 * One force evaluation of a sphere system that allocates no device memory.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It evaluates what gdem_prime / gdem_compute_forces evaluate
 * on a context holding the same particles (material 0, one contact model,
 * no walls): binning, detection, Hooke/Hertz contact forces and torques, and
 * the weight.  All of its buffers live in one static device array of the
 * library (about 34 MB per device, held until the process ends).
 *   cfg       domain, grid request, gravity, device (as gdem_create)
 *   kind, params  GDEM_MODEL_HOOKE or GDEM_MODEL_HERTZ, as gdem_set_contact_model
 *   pos_r, vel_m, om_req   device arrays of n x 4 doubles: (x, y, z, radius),
 *             (v, mass), (angular velocity, equivalent-sphere radius); mass
 *             and equivalent radius as gdem_set_particles derives them from
 *             the density: m = rho 4 pi r^3 / 3, req = ((3 / 4 pi) vol)^(1/3)
 *   crust     crust thickness of every particle
 *   F, T      device arrays of n x 3 doubles (particle order), written
 *   stream    hipStream_t (NULL: the null stream); the call waits for it
 *   err_ids   optional host int32[2]: the ids an error names
 * Returns GDEM_ERR_CAPACITY when n, the cell count or the row count (cells
 * along y times cells along z) exceeds the maxima below, or when a particle
 * detects or receives more contacts than a list holds; CONTACT and
 * SIMULATION as gdem_last_error describes them.  Calls on one device share
 * the static array: calls that overlap in time on one device race. */
#define GDEM_NOALLOC_MAX_PARTICLES 32768
#define GDEM_NOALLOC_MAX_CELLS (1 << 19)
#define GDEM_NOALLOC_MAX_ROWS 16384
int gdem_sphere_forces_noalloc(const gdem_config* cfg, int32_t kind, const double* params,
                               int64_t n, const double* pos_r, const double* vel_m,
                               const double* om_req, double crust, double* F, double* T,
                               void* stream, int32_t* err_ids);

/* ------------------------------------------------------------------------
 * Batched systems: many independent small sphere systems (calibration
 * sweeps) advanced by one persistent launch per call.  There is no batch in
 * the reference, which runs one Grains application per system
 * (Base/src/Grains.cpp:286-452); each system here is that application's step
 * with its own domain, grid, walls and Hooke/Hertz parameters, and is
 * bit-identical to the same system advanced alone by gdem_step with its
 * lists rebuilt at every step (gdem_set_skin(ctx, 0), as the batch does).
 * Particle material is 0, wall material 1.
 * ------------------------------------------------------------------------ */
typedef struct gdem_batch gdem_batch;

typedef struct {
  gdem_config cfg;            /* domain, grid request, gravity (device unused) */
  int64_t n;                  /* particles */
  const double *x, *v, *w, *q, *radius, *density, *crust;  /* as gdem_set_particles */
  int32_t kind;               /* GDEM_MODEL_HOOKE or GDEM_MODEL_HERTZ */
  int32_t nwalls;             /* box walls (<= GDEM_MAX_WALLS) */
  double pp[5];               /* particle-particle parameters (gdem_set_contact_model) */
  double pw[5];               /* particle-wall parameters */
  const double* wall_basis;   /* nwalls x 9, row-major rotation */
  const double* wall_centre;  /* nwalls x 3 */
  const double* wall_half;    /* nwalls x 3, half extents */
  double wall_crust;
  int32_t substeps;           /* the system advances substeps steps of dt / substeps for
                                 each step of dt (0 or 1: one); the batch's policy for
                                 stiff contacts: gdem_step(nsteps * substeps,
                                 dt / substeps) advances the same system alone */
  int32_t reserved;
} gdem_batch_system;

int gdem_batch_create(int32_t device, gdem_batch** out);
int gdem_batch_destroy(gdem_batch* b);
/* Copies the system's arrays; ids count from 0 in call order. */
int gdem_batch_add_system(gdem_batch* b, const gdem_batch_system* sys, int32_t* sys_id);
/* Uploads every system; no system can be added afterwards. */
int gdem_batch_commit(gdem_batch* b);
/* Workgroup size (64, 128, 256 or 512) and workgroup count (0 = fill the
   device) of the persistent launch; results do not depend on either. */
int gdem_batch_set_launch(gdem_batch* b, int32_t block_size, int32_t workgroups);
/* Link + a0 of every system (gdem_prime). */
int gdem_batch_prime(gdem_batch* b);
/* nsteps of every system (gdem_step semantics), one persistent launch. */
int gdem_batch_step(gdem_batch* b, int32_t nsteps, double dt);
int gdem_batch_synchronize(gdem_batch* b);
/* State of one system in particle-id order (any pointer may be NULL). */
int gdem_batch_get_state(gdem_batch* b, int32_t sys, double* x, double* v, double* w,
                         double* q, double* F, double* T);
/* First failing system and its error (code, ids as gdem_last_error). */
int gdem_batch_last_error(gdem_batch* b, int32_t* sys, int32_t* code, int32_t* id0,
                          int32_t* id1);
int gdem_batch_set_kernel_timing(gdem_batch* b, int32_t on);
int gdem_batch_get_counters(gdem_batch* b, int64_t* steps, int64_t* contacts,
                            double* kernel_ms, int64_t* launches);
int gdem_batch_info(gdem_batch* b, int32_t* nsys, int64_t* nparticles);

/* ------------------------------------------------------------------------
 * Differentiated calibration: loss = mean square error of particle positions
 * at observation steps against target trajectories, and its gradient with
 * respect to each system's Hooke parameters (kn, en, mu; etat = -1, the same
 * set for particle-particle and particle-wall contacts) and to the initial
 * velocities, by a hand-written adjoint of the contact-force and leap-frog
 * kernels (checkpoints every K steps, recomputation per window, the contact
 * set of each step held as the forward found it).  There is no counterpart in
 * the reference, which has no derivatives; the forward is its step
 * (Base/src/Grains.cpp:344-364) for spheres.  Systems are described as for
 * gdem_batch_add_system (pw must equal pp, kind Hooke, pp[2] = -1).  Arrays
 * over particles concatenate the systems in add order, ids within.
 * ------------------------------------------------------------------------ */
typedef struct gdem_diff gdem_diff;

int gdem_diff_create(int32_t device, gdem_diff** out);
int gdem_diff_destroy(gdem_diff* d);
int gdem_diff_add_system(gdem_diff* d, const gdem_batch_system* sys, int32_t* sys_id);
/* nsteps of dt, checkpoints every checkpoint_every steps, observation steps
   (strictly increasing, in 1..nsteps).  A system added with substeps > 1
   advances substeps steps of dt / substeps per step (its forward equals
   gdem_step(nsteps * substeps, dt / substeps)); the substep count is held
   when the parameters change, as the contact set is. */
int gdem_diff_commit(gdem_diff* d, int32_t nsteps, double dt, int32_t checkpoint_every,
                     int32_t nobs, const int32_t* obs_steps);
int gdem_diff_set_launch(gdem_diff* d, int32_t block_size, int32_t workgroups);
/* theta: nsys x (kn, en, mu); v0: total particles x 3.  NULL keeps a value. */
int gdem_diff_set_parameters(gdem_diff* d, const double* theta, const double* v0);
/* Target positions, nobs x total particles x 3. */
int gdem_diff_set_targets(gdem_diff* d, const double* target);
/* Positions at the observation steps of the current parameters (nobs x total x 3). */
int gdem_diff_record(gdem_diff* d, double* xobs);
/* Loss of every system (nsys). */
int gdem_diff_loss(gdem_diff* d, double* loss);
/* Loss, d loss / d theta (nsys x 3) and d loss / d v0 (total x 3); any
   pointer may be NULL. */
int gdem_diff_gradient(gdem_diff* d, double* loss, double* grad_theta, double* grad_v0);
/* Forward-mode directional derivative of every system's loss along
   (dtheta: nsys x 3 in (kn, en, mu), dv0: total x 3), the contact set held
   as in the gradient: an independent check of gdem_diff_gradient. */
int gdem_diff_directional(gdem_diff* d, const double* dtheta, const double* dv0, double* loss,
                          double* dloss);
int gdem_diff_last_error(gdem_diff* d, int32_t* sys, int32_t* code, int32_t* id0, int32_t* id1);
int gdem_diff_info(gdem_diff* d, int32_t* nsys, int64_t* nparticles);

#ifdef __cplusplus
}
#endif
#endif
