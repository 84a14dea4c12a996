// gdem_wall_motion.h -- imposed motions of the box walls (host side): the
// velocity and displacement of each wall over one time step, as the
// reference's obstacle kinematics computes them, and the moved wall.
// Reference paths are relative to src/Grains3D/Grains/.
#pragma once
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_convex.h"
#include "gdem_host.h"

#pragma clang fp contract(off)

namespace gdem {

// ObstacleImposedVelocity (Component/src/ObstacleImposedVelocity.cpp): one
// imposed motion of a wall over [t0, t1].
struct WallMotion {
  int32_t wall = 0, type = 0;
  double t0 = 0., t1 = 0.;
  v3 vel = {0., 0., 0.};    // m_translationalVelocity
  v3 omega = {0., 0., 0.};  // m_angularVelocity
  v3 unit = {0., 0., 0.};   // m_unit_vitRef (normalised)
  double amplitude = 0., period = 0., phase = 0.;
  v3 ms_amp = {0., 0., 0.}, ms_period = {0., 0., 0.}, ms_phase = {0., 0., 0.};
  bool about_centre = true;  // m_rotationCenterIsCenterOfMass
  v3 centre = {0., 0., 0.};
};

// Constructor with the XML values (ObstacleImposedVelocity.cpp:54-257),
// parameters as include/gdem.h lists them per type.  Returns false for an
// unknown type.
inline bool make_wall_motion(int32_t wall, int32_t type, double t0, double t1, const double* p,
                             WallMotion& m) {
  m = WallMotion{};
  m.wall = wall;
  m.type = type;
  m.t0 = t0;
  m.t1 = t1;
  auto unit = [](double x, double y, double z) {
    v3 u = mk(x, y, z);
    return dvd(u, norm(u));  // Vector3::normalize (Geometry/src/Vector3.cpp:67-70)
  };
  switch (type) {
    case GDEM_MOTION_CONSTANT_TRANSLATION:
      m.vel = mk(p[0], p[1], p[2]);
      return true;
    case GDEM_MOTION_SIN_TRANSLATION:
      m.unit = unit(p[0], p[1], p[2]);
      m.amplitude = p[3];
      m.period = p[4];
      m.phase = p[5] * kPi / 180.;
      return true;
    case GDEM_MOTION_CRENEL_TRANSLATION:
      m.unit = unit(p[0], p[1], p[2]);
      m.amplitude = p[3];
      m.period = p[4];
      return true;
    case GDEM_MOTION_MULTISIN_TRANSLATION:
      m.unit = unit(p[0], p[1], p[2]);
      m.ms_period = mk(p[6], p[7], p[8]);
      m.ms_phase = mk(p[9] * kPi / 180., p[10] * kPi / 180., p[11] * kPi / 180.);
      // the amplitude is stored already multiplied by the direction
      m.ms_amp = mk(p[3] * m.unit.x, p[4] * m.unit.y, p[5] * m.unit.z);
      return true;
    case GDEM_MOTION_CONSTANT_ROTATION:
      m.omega = mk(p[0], p[1], p[2]);
      m.about_centre = p[6] != 0.;
      m.centre = mk(p[3], p[4], p[5]);
      return true;
    default:
      return false;
  }
}

// isActif / isCompleted (ObstacleImposedVelocity.cpp:293-313)
inline bool motion_active(const WallMotion& m, double td, double te, double dt, double& subint) {
  subint = te - td;
  if (td < m.t0) subint -= (m.t0 - td);
  if (m.t1 < te) subint -= (te - m.t1);
  return subint > dt * 1.e-5;
}
inline bool motion_completed_at(double t1, double t, double dt) { return t > t1 + dt * 1.e-5; }
inline bool motion_completed(const WallMotion& m, double t, double dt) {
  return motion_completed_at(m.t1, t, dt);
}
inline bool interval_active(double t0, double t1, double td, double te, double dt, double& subint) {
  subint = te - td;
  if (td < t0) subint -= (t0 - td);
  if (t1 < te) subint -= (te - t1);
  return subint > dt * 1.e-5;
}

inline double comp(v3 a, int i) { return i == 0 ? a.x : i == 1 ? a.y : a.z; }
inline void set_comp(v3& a, int i, double v) { (i == 0 ? a.x : i == 1 ? a.y : a.z) = v; }

// translationalVelocity (ObstacleImposedVelocity.cpp:317-355), updating the
// stored velocity as the reference does
inline v3 motion_velocity(WallMotion& m, double time, double dt, v3 cg) {
  if (motion_completed(m, time, dt)) {
    m.vel = mk(0., 0., 0.);
  } else if (m.type == GDEM_MOTION_SIN_TRANSLATION) {
    m.vel = scl(m.unit, m.amplitude * sin(2. * kPi * (time - m.t0) / m.period + m.phase));
  } else if (m.type == GDEM_MOTION_CRENEL_TRANSLATION) {
    m.vel = scl(m.unit, m.amplitude * ((int)((time - m.t0) / m.period) % 2 == 0 ? 1 : -1));
  } else if (m.type == GDEM_MOTION_MULTISIN_TRANSLATION) {
    for (int i = 0; i < 3; ++i)
      set_comp(m.vel, i,
               comp(m.ms_amp, i) *
                   sin(2. * kPi * (time - m.t0) / comp(m.ms_period, i) + comp(m.ms_phase, i)) *
                   comp(m.unit, i));
  } else if (m.type == GDEM_MOTION_CONSTANT_ROTATION && !m.about_centre) {
    m.vel = cross(m.omega, sub(cg, m.centre));
  }
  return m.vel;
}

// angularVelocity (ObstacleImposedVelocity.cpp:359-366)
inline v3 motion_omega(WallMotion& m, double time, double dt) {
  if (motion_completed(m, time, dt)) m.omega = mk(0., 0., 0.);
  return m.omega;
}

// angularMotion (ObstacleImposedVelocity.cpp:432-438)
inline v3 motion_rotation(const WallMotion& m, double subint) { return scl(m.omega, subint); }

// Quaternion of a rotation vector: (sin(d/2)/d) rota, cos(d/2); identity for 0
// (ObstacleKinematicsVelocity.cpp:87-93, ObstacleImposedVelocity.cpp:411-418)
inline quat rotation_quat(v3 rota) {
  const double d = norm(rota);
  if (d != 0.) return qmk(scl(rota, sin(d / 2.) / d), cos(d / 2.));
  return qmk(mk(0., 0., 0.), 1.);
}

// Quaternion::rotateVector (Geometry/src/Quaternion.cpp:293-298)
inline v3 rotate_vector(quat q, v3 v) {
  const double nq = norm(q.v);
  return add(add(scl(v, q.w * q.w - nq * nq), scl(q.v, 2. * dot(v, q.v))),
             scl(cross(q.v, v), 2. * q.w));
}

// translationalMotion (ObstacleImposedVelocity.cpp:370-427)
inline v3 motion_translation(WallMotion& m, double time, double dt, double subint, v3 cg) {
  const double ti = time - dt < m.t0 ? m.t0 : time - dt;
  const double te = time > m.t1 ? m.t1 : time;
  switch (m.type) {
    case GDEM_MOTION_CONSTANT_TRANSLATION:
      return scl(m.vel, subint);
    case GDEM_MOTION_CRENEL_TRANSLATION:
      return scl(m.unit, m.amplitude * ((int)((0.5 * (ti + te) - m.t0) / m.period) % 2 == 0 ? 1 : -1) *
                             subint);
    case GDEM_MOTION_SIN_TRANSLATION:
      return scl(m.unit, (m.amplitude * m.period / (2. * kPi)) *
                             (cos(2. * kPi * (ti - m.t0) / m.period + m.phase) -
                              cos(2. * kPi * (te - m.t0) / m.period + m.phase)));
    case GDEM_MOTION_MULTISIN_TRANSLATION: {
      v3 tr = mk(0., 0., 0.);
      for (int i = 0; i < 3; ++i) {
        const double a = comp(m.ms_amp, i), per = comp(m.ms_period, i), ph = comp(m.ms_phase, i);
        set_comp(tr, i, (a * per / (2. * kPi)) *
                            (cos(2. * kPi * (ti - m.t0) / per + ph) -
                             cos(2. * kPi * (te - m.t0) / per + ph)) *
                            comp(m.unit, i));
      }
      return tr;
    }
    case GDEM_MOTION_CONSTANT_ROTATION:
      if (!m.about_centre) {
        const quat q = rotation_quat(motion_rotation(m, subint));
        const v3 lever = sub(cg, m.centre);
        return sub(rotate_vector(q, lever), lever);
      }
      return mk(0., 0., 0.);
    default:
      return mk(0., 0., 0.);
  }
}

// Per-step state of an obstacle's ObstacleKinematicsVelocity
// (ObstacleKinematicsVelocity.cpp:60-177): translation and rotation over the
// step, rotation quaternion over the step, velocities.
struct Kin {
  v3 trans, rot, vt, vr;
  quat q;
};
inline Kin kin_zero() {
  Kin k;
  k.trans = k.rot = k.vt = k.vr = mk(0., 0., 0.);
  k.q = qmk(mk(0., 0., 0.), 1.);
  return k;
}

// Compose (ObstacleKinematicsVelocity.cpp:62-77): the motion of the composite
// an obstacle belongs to, with the obstacle at lever from its centre
inline void kin_compose(Kin& k, const Kin& o, v3 lever) {
  k.rot = add(k.rot, o.rot);
  k.q = qmul(o.q, k.q);
  k.trans = add(o.trans, sub(rotate_vector(o.q, lever), lever));
  k.vr = add(k.vr, o.vr);
  k.vt = add(k.vt, add(o.vt, cross(o.vr, lever)));
}

// ImposedMotion (ObstacleKinematicsVelocity.cpp:86-122) of obstacle `target`
// (wall k >= 0, composite c as -1 - c) over [time - dt, time]: its active
// motions' displacement, rotation and velocities added to its kinematics;
// returns whether its own motions moved it.
inline bool kin_imposed(Kin& k, std::vector<WallMotion>& motions, int32_t target, double time,
                        double dt, v3 cg) {
  v3 depl = mk(0., 0., 0.), rota = mk(0., 0., 0.), vt = mk(0., 0., 0.), vr = mk(0., 0., 0.);
  double subint = 0.;
  for (WallMotion& m : motions) {
    if (m.wall != target || !motion_active(m, time - dt, time, dt, subint)) continue;
    depl = add(depl, motion_translation(m, time, dt, subint, cg));
    rota = add(rota, motion_rotation(m, subint));
    vt = add(vt, motion_velocity(m, time, dt, cg));
    vr = add(vr, motion_omega(m, time, dt));
  }
  const quat qrot = rotation_quat(rota);
  k.trans = add(k.trans, depl);
  k.rot = add(k.rot, rota);
  k.q = qmul(qrot, k.q);
  k.vr = add(k.vr, vr);
  k.vt = add(k.vt, vt);
  return norm(depl) != 0. || norm(rota) != 0.;
}

// ObstacleImposedForce, type ConstantTranslation (Component/src/
// ObstacleImposedForce.cpp:30-90): a wall driven over [t0, t1] so that its
// contact load tends to amp (a proportional-integral-derivative law on the
// load error along amp's direction), or moving at vmax against that direction
// while it has never felt a load.  mass <= 0: chosen from the time step, the
// amplitude and the crust (m_automass).
struct WallForce {
  int32_t wall = 0;  // wall k, or composite c as -1 - c
  double t0 = 0., t1 = 0.;
  v3 amp = {0., 0., 0.}, dir = {0., 0., 0.};
  double mass = 0., vmax = 0.;
  bool automass = false;
  v3 vt = {0., 0., 0.};  // m_translationalVelocity
};

// The controller's memory: function-level statics of
// ObstacleImposedForce::translationalVelocity (:252-257), so shared by every
// imposed force of the run.
struct ForcePid {
  v3 dforce_nm1 = {0., 0., 0.}, dforce_nm2 = {0., 0., 0.}, depl_nm1 = {0., 0., 0.};
  bool nonzeroforce = false;
};

// Per-step state of a wall's ObstacleKinematicsForce (ObstacleKinematicsForce.cpp)
struct Conf {
  v3 vt = {0., 0., 0.}, trans = {0., 0., 0.};
};

// ObstacleImposedForce::translationalVelocity (ObstacleImposedForce.cpp:227-296)
// for ConstantTranslation: the wall's load from the last force evaluation in,
// the velocity and the translation over the active subinterval out.
inline void force_velocity(WallForce& m, ForcePid& pid, double time, double dt, double subint,
                           v3 force, double crust, v3& vt, v3& trans) {
  if (m.automass) m.mass = dt * dt * norm(m.amp) * 1.e7 / (2. * crust);
  const double Kp = 0.5 * dt * dt / m.mass;
  const double Ki = 2. * Kp / dt;
  const double Kd = 3. * Kp * dt / 24.;
  if (!pid.nonzeroforce) pid.nonzeroforce = norm(force) > kEps;
  const v3 d0 = sub(force, m.amp);
  const v3 dforce = mk(d0.x * m.dir.x, d0.y * m.dir.y, d0.z * m.dir.z);
  if (pid.nonzeroforce) {
    const v3 depl = add(sub(add(pid.depl_nm1, scl(dforce, Kp + Ki * dt + Kd / dt)),
                            scl(pid.dforce_nm1, Kp + 2. * Kd / dt)),
                        scl(pid.dforce_nm2, Kd / dt));
    m.vt = dvd(depl, dt);
    trans = scl(depl, subint / dt);
    pid.dforce_nm2 = pid.dforce_nm1;
    pid.dforce_nm1 = dforce;
    pid.depl_nm1 = depl;
  } else {
    m.vt = scl(m.dir, -m.vmax);
    trans = scl(m.vt, subint);
  }
  if (motion_completed_at(m.t1, time, dt)) m.vt = mk(0., 0., 0.);
  vt = m.vt;
}

// Transform::composeLeftByRotation (Geometry/src/Transform.cpp:140-159):
// basis <- R(q) basis, one column at a time
inline void compose_left_rotation(quat q, double* b) {
  const double qx = q.v.x, qy = q.v.y, qz = q.v.z, qw = q.w;
  for (int i = 0; i < 3; ++i) {
    const double mx = b[i], my = b[3 + i], mz = b[6 + i];
    const double px = qy * mz - qz * my + qw * mx;
    const double py = qz * mx - qx * mz + qw * my;
    const double pz = qx * my - qy * mx + qw * mz;
    const double pw = -qx * mx - qy * my - qz * mz;
    b[i] = qy * pz - qz * py - pw * qx + qw * px;
    b[3 + i] = qz * px - qx * pz - pw * qy + qw * py;
    b[6 + i] = qx * py - qy * px - pw * qz + qw * pz;
  }
}

// A CompositeObstacle (Component/src/CompositeObstacle.cpp): a group of walls
// and composites that moves as one; its position is the volume-weighted centre
// of its members (computeVolumeCenterOfMass, :228-255).
struct Composite {
  int32_t parent = -1;  // -1: the root obstacle
  uint32_t restrict_mask = 0;  // directions whose geometric translation is zeroed
  // members in the order they were attached: wall k, or composite c as -1 - c
  std::vector<int32_t> members;
  v3 centre = {0., 0., 0.};
  double basis[9] = {1., 0., 0., 0., 1., 0., 0., 0., 1.};
  double volume = 0.;
  Kin kin = kin_zero();
  Conf conf;  // its ObstacleKinematicsForce
  v3 vel = {0., 0., 0.}, omega = {0., 0., 0.};
};

// The imposed motions of all obstacles, the composite tree and the
// step-to-step state of the obstacle part of AllComponents::Move
// (Component/src/AllComponents.cpp:151-220).
struct WallKinematics {
  // every motion, each obstacle's in the order of ObstacleKinematicsVelocity::append
  // (ObstacleKinematicsVelocity.cpp:30-41: before the first one that does
  // not start earlier, operator< of ObstacleImposedVelocity.cpp:474-478);
  // WallMotion::wall is the obstacle (wall k, or composite c as -1 - c)
  std::vector<WallMotion> motions;
  // still in the list of motions not completed (m_AllImposedVelocitiesOnObstacles)
  std::vector<char> live;
  bool active_prev = false;  // anyactive_previousdt
  std::vector<Composite> comps;
  std::vector<int32_t> wall_parent;  // composite of each wall, -1 for the root
  std::vector<Kin> wall_kin;
  // imposed forces (ObstacleKinematicsForce::append pushes back, :28-31), the
  // shared controller memory, each wall's confinement state, and the wall
  // loads of the last force evaluation (3 per wall) the controller reads
  std::vector<WallForce> forces;
  std::vector<char> force_live;
  ForcePid pid;
  std::vector<Conf> wall_conf;
  const double* loads = nullptr;
  // a composite's load is the sum its getTorsor leaves after each step
  // (Grains.cpp:381, CompositeObstacle.cpp:650-660): none before the first
  bool comp_loads = false;
  // Obstacle::setRestrictedGeomDirMotion (Component/src/Obstacle.cpp:613-617):
  // bit d set = no geometric translation along d (the velocity is kept)
  std::vector<uint32_t> wall_restrict;

  void insert_motion(const WallMotion& m) {
    size_t pos = motions.size();
    for (size_t i = 0; i < motions.size(); ++i)
      if (motions[i].wall == m.wall && !(motions[i].t0 < m.t0)) {
        pos = i;
        break;
      }
    motions.insert(motions.begin() + (long)pos, m);
    live.insert(live.begin() + (long)pos, 1);
  }
  void insert_force(const WallForce& f) {
    forces.push_back(f);
    force_live.push_back(1);
  }
  bool any_force_live() const {
    for (char l : force_live)
      if (l) return true;
    return false;
  }
  bool any_live() const {
    for (char l : live)
      if (l) return true;
    return any_force_live();
  }
  void drop_completed(double time, double dt) {
    for (size_t i = 0; i < motions.size(); ++i)
      if (live[i] && motion_completed(motions[i], time, dt)) live[i] = 0;
    for (size_t i = 0; i < forces.size(); ++i)
      if (force_live[i] && motion_completed_at(forces[i].t1, time, dt)) force_live[i] = 0;
  }
  void fit(int nwalls) {
    if ((int)wall_parent.size() < nwalls) {
      wall_parent.resize((size_t)nwalls, -1);
      wall_kin.resize((size_t)nwalls, kin_zero());
      wall_conf.resize((size_t)nwalls, Conf{});
      wall_restrict.resize((size_t)nwalls, 0u);
    }
  }
  static v3 restricted(v3 t, uint32_t mask) {
    if (mask & 1u) t.x = 0.;
    if (mask & 2u) t.y = 0.;
    if (mask & 4u) t.z = 0.;
    return t;
  }
  // computeVolumeCenterOfMass of composite c (CompositeObstacle.cpp:228-255,
  // SimpleObstacle.cpp:732-735 with Box::getVolume, Box.cpp:273-279):
  // members in the order they were attached (walls by index here)
  std::pair<v3, double> volume_centre(int32_t c, const Wall* walls) {
    v3 centre = mk(0., 0., 0.);
    double vol = 0.;
    for (int32_t mbr : comps[c].members) {
      std::pair<v3, double> pp;
      if (mbr < 0) {
        pp = volume_centre(-1 - mbr, walls);
      } else {
        const Wall& wl = walls[mbr];
        pp = {mk(wl.origin[0], wl.origin[1], wl.origin[2]), 8.0 * wl.ext[0] * wl.ext[1] * wl.ext[2]};
      }
      centre = add(centre, scl(pp.first, pp.second));
      vol += pp.second;
    }
    if (vol > 0.) centre = dvd(centre, vol);
    comps[c].centre = centre;
    comps[c].volume = vol;
    return {centre, vol};
  }
  void reset(Wall* walls, int nwalls) {
    // Obstacle::resetKinematics (Component/src/Obstacle.cpp:160-167),
    // recursively from the root (CompositeObstacle.cpp:449-460)
    for (int w = 0; w < nwalls; ++w) {
      for (int d = 0; d < 3; ++d) walls[w].vel[d] = walls[w].omega[d] = 0.;
      wall_kin[w] = kin_zero();
      wall_conf[w] = Conf{};  // ObstacleKinematicsForce::reset (:118-124)
    }
    for (Composite& c : comps) {
      c.kin = kin_zero();
      c.conf = Conf{};
      c.vel = c.omega = mk(0., 0., 0.);
    }
  }
  // Force of composite c's torsor (CompositeObstacle::getTorsor,
  // Component/src/CompositeObstacle.cpp:650-660): zero, then += each member's
  // in the order they were attached, a member composite's own sum in its place
  v3 composite_load(int32_t c) const {
    v3 f = mk(0., 0., 0.);
    for (int32_t mbr : comps[c].members)
      f = add(f, mbr < 0 ? composite_load(-1 - mbr)
                         : (loads ? mk(loads[3 * mbr], loads[3 * mbr + 1], loads[3 * mbr + 2])
                                  : mk(0., 0., 0.)));
    return f;
  }
  // SimpleObstacle::Move (Component/src/SimpleObstacle.cpp:116-175) of wall w:
  // translate, rotate about its centre, new bounding box; velocities of its
  // kinematics when it moved
  void move_wall(Wall* walls, int w, double time, double dt, bool mother, bool mother_force,
                 bool displace) {
    Wall& wl = walls[w];
    Kin& k = wall_kin[w];
    Conf& cf = wall_conf[w];
    const v3 cg = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
    const bool moving = kin_imposed(k, motions, w, time, dt, cg) || mother;
    if (moving && displace) {
      // SimpleObstacle.cpp:131-136: restricted directions do not move
      const v3 tr = restricted(k.trans, wall_restrict[w]);
      const double origin[3] = {wl.origin[0] + tr.x, wl.origin[1] + tr.y, wl.origin[2] + tr.z};
      double basis[9], ext[3];
      memcpy(basis, wl.basis, sizeof(basis));
      memcpy(ext, wl.ext, sizeof(ext));
      compose_left_rotation(k.q, basis);
      wl = make_wall(basis, origin, ext, wl.crust, wl.material, wl.kind);
    }
    // the imposed forces (ObstacleKinematicsForce::ImposedMotion,
    // ObstacleKinematicsForce.cpp:40-63: every active one in turn, the last
    // one's velocity and translation kept), then the translation
    // (SimpleObstacle.cpp:142-152)
    const v3 load = loads ? mk(loads[3 * w], loads[3 * w + 1], loads[3 * w + 2]) : mk(0., 0., 0.);
    for (WallForce& m : forces) {
      double sub_dt = 0.;
      if (m.wall != w || !interval_active(m.t0, m.t1, time - dt, time, dt, sub_dt)) continue;
      force_velocity(m, pid, time, dt, sub_dt, load, wl.crust, cf.vt, cf.trans);
    }
    const bool move_force = norm(cf.vt) != 0. || mother_force;
    if (move_force && displace) {
      const v3 tr = restricted(cf.trans, wall_restrict[w]);  // SimpleObstacle.cpp:146-149
      const double origin[3] = {wl.origin[0] + tr.x, wl.origin[1] + tr.y, wl.origin[2] + tr.z};
      double basis[9], ext[3];
      memcpy(basis, wl.basis, sizeof(basis));
      memcpy(ext, wl.ext, sizeof(ext));
      wl = make_wall(basis, origin, ext, wl.crust, wl.material, wl.kind);
    }
    if (!moving && !move_force) return;
    // Obstacle::setVelocity (Obstacle.cpp:635-640): both kinematics
    const v3 vel = add(k.vt, cf.vt);
    wl.vel[0] = vel.x;
    wl.vel[1] = vel.y;
    wl.vel[2] = vel.z;
    wl.omega[0] = k.vr.x;
    wl.omega[1] = k.vr.y;
    wl.omega[2] = k.vr.z;
  }
  // CompositeObstacle::Move (Component/src/CompositeObstacle.cpp:135-220):
  // the composite's own motion, composed onto each member at its lever from
  // the (moved) centre, then the members move
  void move_composite(Wall* walls, int nwalls, int32_t c, double time, double dt, bool mother,
                      bool mother_force, bool displace) {
    Composite& C = comps[c];
    const bool moving = kin_imposed(C.kin, motions, -1 - c, time, dt, C.centre) || mother;
    if (moving && displace) {
      // CompositeObstacle.cpp:153-160: the composite's own translation only
      C.centre = add(C.centre, restricted(C.kin.trans, C.restrict_mask));
      compose_left_rotation(C.kin.q, C.basis);
    }
    if (moving)
      for (int32_t mbr : C.members) {
        if (mbr < 0) {
          Composite& M = comps[-1 - mbr];
          kin_compose(M.kin, C.kin, sub(M.centre, C.centre));
        } else {
          const v3 pos = mk(walls[mbr].origin[0], walls[mbr].origin[1], walls[mbr].origin[2]);
          kin_compose(wall_kin[mbr], C.kin, sub(pos, C.centre));
        }
      }
    // the composite's imposed forces on its torsor (CompositeObstacle.cpp:
    // 182-203): its translation, then composed onto every member
    // (ObstacleKinematicsForce::Compose, ObstacleKinematicsForce.cpp:51-55)
    const v3 load = comp_loads ? composite_load(c) : mk(0., 0., 0.);
    for (WallForce& m : forces) {
      double sub_dt = 0.;
      if (m.wall != -1 - c || !interval_active(m.t0, m.t1, time - dt, time, dt, sub_dt)) continue;
      force_velocity(m, pid, time, dt, sub_dt, load, 0., C.conf.vt, C.conf.trans);
    }
    const bool move_force = norm(C.conf.vt) != 0. || mother_force;
    if (move_force && displace) C.centre = add(C.centre, restricted(C.conf.trans, C.restrict_mask));
    if (move_force)
      for (int32_t mbr : C.members) {
        Conf& m = mbr < 0 ? comps[-1 - mbr].conf : wall_conf[mbr];
        m.vt = add(m.vt, C.conf.vt);
        m.trans = add(m.trans, C.conf.trans);
      }
    move_members(walls, nwalls, c, time, dt, moving, move_force, displace);
    // Obstacle::setVelocity (Obstacle.cpp:635-640): both kinematics
    if (moving || move_force) {
      C.vel = add(C.kin.vt, C.conf.vt);
      C.omega = C.kin.vr;
    }
  }
  void move_members(Wall* walls, int nwalls, int32_t parent, double time, double dt, bool mother,
                    bool mother_force, bool displace) {
    for (int32_t k = 0; k < (int32_t)comps.size(); ++k)
      if (comps[k].parent == parent)
        move_composite(walls, nwalls, k, time, dt, mother, mother_force, displace);
    for (int w = 0; w < nwalls; ++w)
      if (wall_parent[w] == parent) move_wall(walls, w, time, dt, mother, mother_force, displace);
  }
  // Obstacle part of AllComponents::Move at `time` (over [time - dt, time]);
  // returns whether any wall changed.
  bool move(Wall* walls, int nwalls, double time, double dt) {
    if (!any_live()) return false;
    fit(nwalls);
    bool anyactive = false;
    double subint = 0.;
    for (size_t i = 0; i < motions.size(); ++i)
      if (live[i] && motion_active(motions[i], time - dt, time, dt, subint)) anyactive = true;
    for (size_t i = 0; i < forces.size(); ++i)
      if (force_live[i] &&
          interval_active(forces[i].t0, forces[i].t1, time - dt, time, dt, subint))
        anyactive = true;
    const bool changed = anyactive || active_prev;
    if (changed) reset(walls, nwalls);
    active_prev = anyactive;
    if (anyactive) move_members(walls, nwalls, -1, time, dt, false, false, true);
    drop_completed(time, dt);
    comp_loads = true;
    return changed;
  }
  // AllComponents::setKinematicsObstacleWithoutMoving (AllComponents.cpp:235-251):
  // the velocities at the initial time, obstacles left in place.
  bool init(Wall* walls, int nwalls, double time, double dt) {
    if (!any_live()) return false;
    fit(nwalls);
    reset(walls, nwalls);
    move_members(walls, nwalls, -1, time, dt, false, false, false);
    drop_completed(time, dt);
    return true;
  }
};

// CylindricalShell with box elements (Component/src/CylindricalShell.cpp:13-93):
// n boxes of edges (width, ly, height) on a circle of radius radius + width/2
// about the shell's z axis, box i at angle (i + 1/2) 2 pi / n and rotated by
// that angle, then placed by the shell's transform (basis, centre) with
// Transform::composeLeftByTransform (Geometry/src/Transform.cpp:355-360) and a
// zero translation.  Writes box i's basis (row-major), centre and half edges.
inline void cylindrical_shell_boxes(const double* sb, const double* sc, double height,
                                    double radius, double width, int n, double* basis,
                                    double* centre, double* half) {
  const double ext_rad = radius + width / 2.;
  const double delta = 2. * kPi / double(n);
  const double ly = 2. * (radius + width) * (sin(delta) - (1. - cos(delta)) / tan(delta));
  for (int i = 0; i < n; ++i) {
    const double angle = (0.5 + double(i)) * delta;
    const double ca = cos(angle), sa = sin(angle);
    const double cg[3] = {ext_rad * ca, ext_rad * sa, 0.};
    const double m[9] = {ca, -sa, 0., sa, ca, 0., 0., 0., 1.};
    double* B = basis + 9 * i;
    double* C = centre + 3 * i;
    for (int r = 0; r < 3; ++r) {
      // Matrix * Vector3 (Matrix.cpp:517-520) and Matrix * Matrix (:537-549)
      C[r] = sc[r] + (sb[3 * r] * cg[0] + sb[3 * r + 1] * cg[1] + sb[3 * r + 2] * cg[2]);
      for (int col = 0; col < 3; ++col)
        B[3 * r + col] = sb[3 * r] * m[col] + sb[3 * r + 1] * m[3 + col] + sb[3 * r + 2] * m[6 + col];
      C[r] = C[r] + 0.;  // SimpleObstacle::Translate by the zero vector
    }
    // Box(width, ly, height) (Geometry/src/Box.cpp:27-33)
    half[3 * i] = fabs(width) / 2.;
    half[3 * i + 1] = fabs(ly) / 2.;
    half[3 * i + 2] = fabs(height) / 2.;
  }
}

// CylindricalShell with its default Rectangle elements
// (Component/src/CylindricalShell.cpp:55-87 without ElementaryObstacle="Box"):
// rectangle i of sides (LX, LY) = (height, ly) centred at the inner radius,
// at angle (i + 1/2) 2 pi / n, its basis R(angle) M with M the rows
// (0, 0, -1), (0, 1, 0), (1, 0, 0) (Matrix::operator*=, Matrix.cpp:447-460),
// so that its local x is the shell axis and its normal (local z) points at
// the axis; then the shell's transform.  half receives (LX/2, LY/2, 0).
inline void cylindrical_shell_rectangles(const double* sb, const double* sc, double height,
                                         double radius, double width, int n, double* basis,
                                         double* centre, double* half) {
  const double ext_rad = radius;
  const double delta = 2. * kPi / double(n);
  const double ly = 2. * (radius + width) * (sin(delta) - (1. - cos(delta)) / tan(delta));
  const double mm[9] = {0., 0., -1., 0., 1., 0., 1., 0., 0.};
  for (int i = 0; i < n; ++i) {
    const double angle = (0.5 + double(i)) * delta;
    const double ca = cos(angle), sa = sin(angle);
    const double cg[3] = {ext_rad * ca, ext_rad * sa, 0.};
    const double m0[9] = {ca, -sa, 0., sa, ca, 0., 0., 0., 1.};
    double m[9];
    for (int r = 0; r < 3; ++r)
      for (int col = 0; col < 3; ++col)
        m[3 * r + col] = m0[3 * r] * mm[col] + m0[3 * r + 1] * mm[3 + col] + m0[3 * r + 2] * mm[6 + col];
    double* B = basis + 9 * i;
    double* C = centre + 3 * i;
    for (int r = 0; r < 3; ++r) {
      C[r] = sc[r] + (sb[3 * r] * cg[0] + sb[3 * r + 1] * cg[1] + sb[3 * r + 2] * cg[2]);
      for (int col = 0; col < 3; ++col)
        B[3 * r + col] = sb[3 * r] * m[col] + sb[3 * r + 1] * m[3 + col] + sb[3 * r + 2] * m[6 + col];
      C[r] = C[r] + 0.;
    }
    // Rectangle(height, ly) (Geometry/src/Rectangle.cpp:13-21)
    half[3 * i] = height / 2.;
    half[3 * i + 1] = ly / 2.;
    half[3 * i + 2] = 0.;
  }
}

}  // namespace gdem
