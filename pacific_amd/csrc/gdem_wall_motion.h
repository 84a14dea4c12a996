// gdem_wall_motion.h -- imposed motions of the box walls (host side): the
// velocity and displacement of each wall over one time step, as the
// reference's obstacle kinematics computes them, and the moved wall.
// Reference paths are relative to src/Grains3D/Grains/.
#pragma once
#include <cmath>
#include <cstring>
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
inline bool motion_completed(const WallMotion& m, double t, double dt) { return t > m.t1 + dt * 1.e-5; }

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
        // Quaternion::rotateVector (Geometry/src/Quaternion.cpp:293-298)
        const double nq = norm(q.v);
        const v3 r = add(add(scl(lever, q.w * q.w - nq * nq), scl(q.v, 2. * dot(lever, q.v))),
                         scl(cross(q.v, lever), 2. * q.w));
        return sub(r, lever);
      }
      return mk(0., 0., 0.);
    default:
      return mk(0., 0., 0.);
  }
}

// One wall's kinematics over [time - dt, time]
// (ObstacleKinematicsVelocity::ImposedMotion, Component/src/ObstacleKinematicsVelocity.cpp:72-103):
// the summed displacement, rotation quaternion and velocities of its active
// motions; returns whether it moved.
inline bool wall_imposed_motion(std::vector<WallMotion>& motions, int32_t wall, double time,
                                double dt, v3 cg, v3& depl, quat& qrot, v3& vt, v3& vr) {
  depl = mk(0., 0., 0.);
  v3 rota = mk(0., 0., 0.);
  vt = mk(0., 0., 0.);
  vr = mk(0., 0., 0.);
  double subint = 0.;
  for (WallMotion& m : motions) {
    if (m.wall != wall || !motion_active(m, time - dt, time, dt, subint)) continue;
    depl = add(depl, motion_translation(m, time, dt, subint, cg));
    rota = add(rota, motion_rotation(m, subint));
    vt = add(vt, motion_velocity(m, time, dt, cg));
    vr = add(vr, motion_omega(m, time, dt));
  }
  qrot = rotation_quat(rota);
  return norm(depl) != 0. || norm(rota) != 0.;
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

// The imposed motions of all walls and the step-to-step state of the
// obstacle part of AllComponents::Move (Component/src/AllComponents.cpp:151-220).
struct WallKinematics {
  // every motion, each wall's in the order of ObstacleKinematicsVelocity::append
  // (ObstacleKinematicsVelocity.cpp:30-41: before the first one that does
  // not start earlier, operator< of ObstacleImposedVelocity.cpp:474-478)
  std::vector<WallMotion> motions;
  // still in the list of motions not completed (m_AllImposedVelocitiesOnObstacles)
  std::vector<char> live;
  bool active_prev = false;  // anyactive_previousdt

  void add(const WallMotion& m) {
    size_t pos = motions.size();
    for (size_t i = 0; i < motions.size(); ++i)
      if (motions[i].wall == m.wall && !(motions[i].t0 < m.t0)) {
        pos = i;
        break;
      }
    motions.insert(motions.begin() + (long)pos, m);
    live.insert(live.begin() + (long)pos, 1);
  }
  bool any_live() const {
    for (char l : live)
      if (l) return true;
    return false;
  }
  void drop_completed(double time, double dt) {
    for (size_t i = 0; i < motions.size(); ++i)
      if (live[i] && motion_completed(motions[i], time, dt)) live[i] = 0;
  }
  static void reset(Wall* walls, int nwalls) {
    // Obstacle::resetKinematics (Component/src/Obstacle.cpp:160-167)
    for (int w = 0; w < nwalls; ++w)
      for (int d = 0; d < 3; ++d) walls[w].vel[d] = walls[w].omega[d] = 0.;
  }
  // SimpleObstacle::Move (Component/src/SimpleObstacle.cpp:116-175) of wall w:
  // translate, rotate about its centre, new bounding box; velocities of its
  // kinematics when it moved
  void move_wall(Wall* walls, int w, double time, double dt, bool displace) {
    Wall& wl = walls[w];
    v3 depl, vt, vr;
    quat qrot;
    const v3 cg = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
    if (!wall_imposed_motion(motions, w, time, dt, cg, depl, qrot, vt, vr)) return;
    if (displace) {
      // m_QuaternionRotationOverDt = qrot * identity after the reset
      // (ObstacleKinematicsVelocity.cpp:115-117, :168-177)
      const quat q = qmul(qrot, qmk(mk(0., 0., 0.), 1.));
      const double origin[3] = {wl.origin[0] + depl.x, wl.origin[1] + depl.y,
                                wl.origin[2] + depl.z};
      double basis[9], ext[3];
      memcpy(basis, wl.basis, sizeof(basis));
      memcpy(ext, wl.ext, sizeof(ext));
      compose_left_rotation(q, basis);
      wl = make_wall(basis, origin, ext, wl.crust, wl.material);
    }
    // Obstacle::setVelocity (Obstacle.cpp:635-640)
    wl.vel[0] = vt.x;
    wl.vel[1] = vt.y;
    wl.vel[2] = vt.z;
    wl.omega[0] = vr.x;
    wl.omega[1] = vr.y;
    wl.omega[2] = vr.z;
  }
  // Obstacle part of AllComponents::Move at `time` (over [time - dt, time]);
  // returns whether any wall changed.
  bool move(Wall* walls, int nwalls, double time, double dt) {
    if (!any_live()) return false;
    bool anyactive = false;
    double subint = 0.;
    for (size_t i = 0; i < motions.size(); ++i)
      if (live[i] && motion_active(motions[i], time - dt, time, dt, subint)) anyactive = true;
    const bool changed = anyactive || active_prev;
    if (changed) reset(walls, nwalls);
    active_prev = anyactive;
    if (anyactive)
      for (int w = 0; w < nwalls; ++w) move_wall(walls, w, time, dt, true);
    drop_completed(time, dt);
    return changed;
  }
  // AllComponents::setKinematicsObstacleWithoutMoving (AllComponents.cpp:235-251):
  // the velocities at the initial time, walls left in place.
  bool init(Wall* walls, int nwalls, double time, double dt) {
    if (!any_live()) return false;
    reset(walls, nwalls);
    for (int w = 0; w < nwalls; ++w) move_wall(walls, w, time, dt, false);
    drop_completed(time, dt);
    return true;
  }
};

}  // namespace gdem
