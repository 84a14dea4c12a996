// gdem_grains.hpp -- C++ host classes with the plugin seams of Grains3D's step,
// forwarding to the C ABI of include/gdem.h (libgdem.so).
//
// A Grains3D user keeps the vocabulary of the reference: contact force models
// built from a map of parameters and registered per material pair, a time
// integrator chosen by name, a LinkedCell application that links particles and
// obstacles and computes forces, and the kick-drift-kick loop of
// Grains::Simulation.  The work is batched: one call per phase for all
// particles, on the GPU.  Reference paths are relative to src/Grains3D/Grains/.
#pragma once

#include <array>
#include <list>
#include <map>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gdem.h"

namespace gdem {
namespace grains {

// Base/include/Error.hh:19 -- overlap reached the sum of the crust thicknesses
class ContactError : public std::runtime_error {
 public:
  ContactError(int id0, int id1, double time);
  void Message(std::ostream& fileOut) const;
  int id0, id1;
  double time;
};

// Error.hh:75 -- displacement over one step exceeded the crust thickness
class MotionError : public std::runtime_error {
 public:
  MotionError(int id, double time);
  void Message(std::ostream& fileOut) const;
  int id;
  double time;
};

// Error.hh:129 -- particle outside the linked-cell grid
class SimulationError : public std::runtime_error {
 public:
  SimulationError(int id, double time);
  void Message(std::ostream& fileOut) const;
  int id;
  double time;
};

// Global settings of the step (Base/include/GrainsExec.hh: m_vgravity,
// m_TIScheme).
struct GrainsExec {
  static std::array<double, 3> m_vgravity;
  static std::string m_TIScheme;
};

// Component/include/ContactForceModel.hh:26-77.  The per-pair computeForces of
// the reference runs on the device for all pairs at once; a model here carries
// its parameters and the analytic estimate of a binary collision.
class ContactForceModel {
 public:
  explicit ContactForceModel(std::map<std::string, double>& parameters);
  virtual ~ContactForceModel();
  virtual std::string name() const = 0;
  // Gravityless binary collision of two spheres of masses m0, m1 and
  // equivalent radii req0, req1 at relative velocity v0.
  virtual void computeAndWriteEstimates(double m0, double m1, double crust0, double crust1,
                                        double req0, double req1, double v0, double dt,
                                        std::ostream& OUT) const = 0;
  virtual int32_t kind() const = 0;  // GDEM_MODEL_*
  const double* parameters() const { return m_p; }

 protected:
  double m_p[8] = {0., 0., 0., 0., 0., 0., 0., 0.};  // as in include/gdem.h, per kind
};

// Component/src/HookeContactForceModel.cpp:11-20; keys kn, en, etat, muc, kr
// (etat defaults to -1, gamma_t = gamma_n, as defineParameters does, :186-189)
class HookeContactForceModel : public ContactForceModel {
 public:
  explicit HookeContactForceModel(std::map<std::string, double>& parameters);
  std::string name() const override { return "HookeContactForceModel"; }
  void computeAndWriteEstimates(double m0, double m1, double crust0, double crust1, double req0,
                                double req1, double v0, double dt,
                                std::ostream& OUT) const override;
  int32_t kind() const override { return GDEM_MODEL_HOOKE; }
  double contactTime(double m0, double m1) const;
  double maxOverlap(double m0, double m1, double v0) const;
};

// Component/src/HertzContactForceModel.cpp:12-22; keys Es, en, Gs, muc, kr
// (Gs defaults to Es / 4 as defineParameters does, :193-196)
class HertzContactForceModel : public ContactForceModel {
 public:
  explicit HertzContactForceModel(std::map<std::string, double>& parameters);
  std::string name() const override { return "HertzContactForceModel"; }
  void computeAndWriteEstimates(double m0, double m1, double crust0, double crust1, double req0,
                                double req1, double v0, double dt,
                                std::ostream& OUT) const override;
  int32_t kind() const override { return GDEM_MODEL_HERTZ; }
};

// Component/src/HookeMemoryContactForceModel.cpp:11-29; keys kn, en, muc, kt;
// etat defaults to -1, mur and etarpf to 0, and epst to 0: defineParameters
// (:256-307) stores its 1e-10 default under the key "eps", so the constructor
// reads 0.  Tangential spring, contact-plane rotation and rolling friction
// with memory of each contact (gdem_set_contact_model, kind
// GDEM_MODEL_HOOKE_MEMORY); spheres only.
class HookeMemoryContactForceModel : public ContactForceModel {
 public:
  explicit HookeMemoryContactForceModel(std::map<std::string, double>& parameters);
  std::string name() const override { return "HookeMemoryContactForceModel"; }
  void computeAndWriteEstimates(double m0, double m1, double crust0, double crust1, double req0,
                                double req1, double v0, double dt,
                                std::ostream& OUT) const override;
  int32_t kind() const override { return GDEM_MODEL_HOOKE_MEMORY; }
};

// Component/src/HertzMemoryContactForceModel.cpp:12-31; keys Es, en, muc; Gs
// defaults to Es / 4, mur and etarpf to 0, epst to 0 (as for HookeMemory).
class HertzMemoryContactForceModel : public ContactForceModel {
 public:
  explicit HertzMemoryContactForceModel(std::map<std::string, double>& parameters);
  std::string name() const override { return "HertzMemoryContactForceModel"; }
  void computeAndWriteEstimates(double m0, double m1, double crust0, double crust1, double req0,
                                double req1, double v0, double dt,
                                std::ostream& OUT) const override;
  int32_t kind() const override { return GDEM_MODEL_HERTZ_MEMORY; }
};

// Component/src/ContactBuilderFactory.cpp:22-192: materials and the model of
// each unordered pair of materials.
class ContactBuilderFactory {
 public:
  static void defineMaterial(std::string const& material, bool matForObstacle);
  static void define(std::string const& matA, std::string const& matB,
                     std::unique_ptr<ContactForceModel> model);
  static ContactForceModel* contactForceModel(std::string const& matA, std::string const& matB);
  static int materialIndex(std::string const& material);
  static bool checkContactForceModelsExist(std::string& matA, std::string& matB);
  static void eraseAllContactForceModels();
  static void reset();
};

// TimeIntegration/include/TimeIntegrator.hh:47-65
class TimeIntegrator {
 public:
  virtual ~TimeIntegrator();
  virtual TimeIntegrator* clone() const = 0;
  virtual std::string name() const = 0;
};

// TimeIntegration/src/SecondOrderLeapFrog.cpp: kick-drift-kick on the device
class SecondOrderLeapFrog : public TimeIntegrator {
 public:
  TimeIntegrator* clone() const override { return new SecondOrderLeapFrog(*this); }
  std::string name() const override { return "SecondOrderLeapFrog"; }
};

// TimeIntegration/src/TimeIntegratorBuilderFactory.cpp:13-38; only the
// leap-frog scheme has a device implementation.
class TimeIntegratorBuilderFactory {
 public:
  static TimeIntegrator* create(std::string const& scheme = GrainsExec::m_TIScheme);
};

// A sphere as inserted by Grains::insertParticle (Base/src/Grains.cpp:2070-2143)
struct SphereParticle {
  double position[3];
  double velocity[3];
  double angular_velocity[3];
  double quaternion[4];  // (x, y, z, w)
  double radius;
  double density;
  double crust;
  std::string material;
};

// A static SimpleObstacle with a Box convex (Box(LX, LY, LZ): full edges)
struct BoxObstacle {
  std::string name;
  double centre[3];
  double edges[3];
  double basis[9];
  double crust;
  std::string material;
};

// A static SimpleObstacle with a Rectangle convex (Geometry/include/
// Rectangle.hh: sides LX along the local x, LY along the local y, no
// thickness, normal along the local z).  Sphere and convex particles.
struct RectangleObstacle {
  std::string name;
  double centre[3];
  double LX, LY;
  double basis[9];
  double crust;
  std::string material;
};

// CylindricalShell (Component/include/CylindricalShell.hh): n elements
// (named name0 .. name{n-1}) around the axis of the shell's transform
// (centre, row-major basis; the axis is the basis' z), a composite obstacle.
// As the XML attribute ElementaryObstacle: "Box" gives box elements, any
// other value (the default, empty) the reference's default Rectangle
// elements (CylindricalShell.cpp:31-35).
struct CylindricalShell {
  std::string name;
  double centre[3];
  double basis[9];
  double height, radius, width;
  int n;
  double crust;
  std::string material;
  std::string elementary_obstacle;
};

// ObstacleImposedVelocity (Component/include/ObstacleImposedVelocity.hh): an
// imposed motion of a named obstacle over [tstart, tend].  type and the
// parameter names follow the reference's XML node (ObstacleImposedVelocity.cpp:
// 54-257):
//   "ConstantTranslation"  VX VY VZ
//   "SinTranslation"       VX VY VZ (direction) Amplitude Period PhaseShift (degrees)
//   "CrenelTranslation"    VX VY VZ (direction) Amplitude Period
//   "MultiSinTranslation"  VX VY VZ (direction) AX AY AZ PX PY PZ PhiX PhiY PhiZ
//   "ConstantRotation"     WX WY WZ, and CX CY CZ for a rotation about that
//                          point instead of the obstacle's centre
struct ObstacleImposedVelocity {
  std::string obstacle;
  std::string type;
  double tstart = 0., tend = 0.;
  std::map<std::string, double> parameters;
};

// ObstacleImposedForce (Component/include/ObstacleImposedForce.hh), type
// "ConstantTranslation": parameters AX AY AZ (the force amplitude), Vmax, and
// Mass (absent: chosen automatically, as the reference does without it).
struct ObstacleImposedForce {
  std::string obstacle;
  std::string type;
  double tstart = 0., tend = 0.;
  std::map<std::string, double> parameters;
};

// Base/include/LinkedCell.hh (an AppCollision application): the grid, the
// particle/obstacle links, LinkUpdate and ComputeForces.
class LinkedCell {
 public:
  LinkedCell(const double origin[3], const double size[3], const bool periodic[3], int device = 0);
  ~LinkedCell();
  LinkedCell(const LinkedCell&) = delete;
  LinkedCell& operator=(const LinkedCell&) = delete;

  // LinkedCell::set (Base/src/LinkedCell.cpp:47-419); returns an error count
  size_t set(double cellsize, std::string const& oshift = "");
  // LinkedCell::Link(Particle*) for every inserted particle
  void Link(std::vector<SphereParticle> const& particles);
  // LinkedCell::Link(Obstacle*) (LinkedCell.cpp:1371-1445)
  void Link(BoxObstacle const& obstacle);
  void Link(RectangleObstacle const& obstacle);
  // AppCollision::LinkUpdate (LinkedCell.cpp:1452-1547)
  void LinkUpdate(double time, double dt);
  // App::ComputeForces (Base/include/App.hh:38-39)
  void ComputeForces(double time, double dt);

  void getCells(int32_t nc[3], double edge[3]) const;
  size_t numberOfParticles() const { return m_n; }
  gdem_ctx* context() const { return m_ctx; }
  // CompositeObstacle (Component/src/CompositeObstacle.cpp): an empty
  // composite, member of `parent` ("" for the root)
  void LinkComposite(std::string const& name, std::string const& parent = "");
  // CompositeObstacle::append (CompositeObstacle.cpp:77-80) of a linked box
  void Append(std::string const& obstacle, std::string const& composite);
  // CylindricalShell: a composite of its boxes, member of `parent`
  void Link(CylindricalShell const& shell, std::string const& parent = "");
  // Obstacle::setRestrictedGeomDirMotion (Component/src/Obstacle.cpp:613-617)
  // of a linked box obstacle or composite: directions 0 (x), 1 (y), 2 (z)
  // whose geometric translation is suppressed
  void setRestrictedGeomDirMotion(std::string const& obstacle, std::list<size_t> const& dirs);
  // index of the linked box obstacle / composite of that name (-1 if none)
  int obstacleIndex(std::string const& name) const;
  int compositeIndex(std::string const& name) const;

 private:
  void push_models();
  std::map<std::string, int> m_obstacles;
  std::map<std::string, int> m_composites;
  int m_nobstacles = 0;
  double m_origin[3], m_size[3];
  bool m_periodic[3];
  int m_device;
  gdem_ctx* m_ctx = nullptr;
  size_t m_n = 0;
};

// The step of Grains::Simulation (Base/src/Grains.cpp:271-452) for a
// LinkedCell and a time integrator; errors come back as the reference's
// exceptions.
class Grains {
 public:
  Grains(LinkedCell& lc, TimeIntegrator const& ti);
  // a0 of the leap-frog (Grains.cpp:169-174)
  void do_before_time_stepping();
  // Kick-drift-kick loop from the current time over time_interval
  void Simulation(double time_interval, double dt);
  // AllComponents::LinkImposedMotion (Component/src/AllComponents.cpp:130-146)
  void LinkImposedMotion(ObstacleImposedVelocity const& motion);
  void LinkImposedMotion(ObstacleImposedForce const& force);
  // AllComponents::computeObstaclesLoad (AllComponents.cpp:1685-1693): track
  // the obstacles' loads, and read a box obstacle's torsor (force, torque
  // about its centre) of the last force evaluation
  void trackObstaclesLoad(bool on);
  void getObstacleLoad(std::string const& name, double force[3], double torque[3]) const;
  // Grains::m_time (the Simulation loop adds dt before each step's move)
  void setTime(double time);
  // Phases (Grains.cpp:459-527), for callers that interleave their own work;
  // obstacles move over [time - dt_obstacle, time] when dt_obstacle > 0
  void moveParticlesAndObstacles(double dt_particle_vel, double dt_particle_disp,
                                 double dt_obstacle = 0.);
  void computeParticlesForceAndTorque();
  void advanceParticlesVelocity(double dt_particle_vel);
  // Particle state in insertion order (6 vectors of 3 per particle + quaternion)
  void getState(std::vector<double>& x, std::vector<double>& v, std::vector<double>& w,
                std::vector<double>& q, std::vector<double>& F, std::vector<double>& T) const;
  double time() const { return m_time; }
  // Grains::m_dt (read from the input's time step in the reference); the
  // memory contact models use it in every force evaluation
  void setTimeStep(double dt) { m_dt = dt; }
  // The particle records of a reload file (the "_Particles.result[.bin]" part
  // of Grains::saveReload, Base/src/Grains.cpp:2543-2581, and of the reload
  // read, AllComponents::read): write the current state, or restore it so that
  // Simulation continues from the stored F, T without recomputing a0.
  void saveParticles2014(std::string const& path, bool binary, int id_offset = 0,
                         int wall_id_offset = 0);
  void readParticles2014(std::string const& path, bool binary, int id_offset = 0,
                         int wall_id_offset = 0);

 private:
  void check(int rc) const;
  LinkedCell& m_lc;
  double m_time = 0.;
  double m_dt = 0.;
  bool m_primed = false;
};

}  // namespace grains
}  // namespace gdem
