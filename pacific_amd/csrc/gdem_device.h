// gdem_device.h -- FP64 device arithmetic of the Grains3D DEM step.
//
// Every helper keeps the operand order of the reference C++ so that results
// round exactly as the CPU path does (the library is built with
// -ffp-contract=off; no FMA is formed).  Paths are relative to
// src/Grains3D/Grains/ of the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace gdem {

constexpr double kPi = 3.14159265358979323846;  // Base/include/Basic.hh:22
constexpr double kEps = 1.0e-10;                // Basic.hh:26 (EPSILON)
constexpr int kMaxMaterials = 32;
constexpr int kMaxWalls = 64;

struct v3 {
  double x, y, z;
};

__host__ __device__ __forceinline__ v3 mk(double x, double y, double z) {
  v3 r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
// Geometry/src/Group3.cpp:149-196
__host__ __device__ __forceinline__ v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ v3 scl(v3 a, double d) { return mk(a.x * d, a.y * d, a.z * d); }
__host__ __device__ __forceinline__ v3 dvd(v3 a, double d) { return mk(a.x / d, a.y / d, a.z / d); }
__host__ __device__ __forceinline__ v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Geometry/src/Vector3.cpp:114-119
__host__ __device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
// Vector3.cpp:147-151 (IEEE sqrt, correctly rounded)
__host__ __device__ __forceinline__ double norm(v3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// Group3.cpp:390-395
__host__ __device__ __forceinline__ v3 round3(v3 a) {
  return mk(fabs(a.x) < kEps ? 0. : a.x, fabs(a.y) < kEps ? 0. : a.y,
            fabs(a.z) < kEps ? 0. : a.z);
}

struct quat {
  v3 v;
  double w;
};
__host__ __device__ __forceinline__ quat qmk(v3 v, double w) {
  quat q;
  q.v = v;
  q.w = w;
  return q;
}
// Geometry/src/Quaternion.cpp:74-77
__host__ __device__ __forceinline__ quat qconj(quat q) { return qmk(neg(q.v), q.w); }
// Quaternion.cpp:95-100: [0, lhs] x q (operator,(Vector3, Quaternion), :468-472)
__host__ __device__ __forceinline__ quat qleftvec(quat q, v3 lhs) {
  double tmp = dot(neg(lhs), q.v);
  v3 vt = add(cross(lhs, q.v), scl(lhs, q.w));
  return qmk(vt, tmp);
}
// Quaternion.cpp:108-113
__host__ __device__ __forceinline__ v3 qmulv(quat p, quat r) {
  return add(add(cross(p.v, r.v), scl(r.v, p.w)), scl(p.v, r.w));
}
// Quaternion.cpp:442-448
__host__ __device__ __forceinline__ quat qmul(quat p, quat r) {
  double tmp = (p.w * r.w) - dot(p.v, r.v);
  v3 vt = add(add(cross(p.v, r.v), scl(r.v, p.w)), scl(p.v, r.w));
  return qmk(vt, tmp);
}

// Contact model parameters (one entry per unordered material pair).
struct Model {
  int32_t kind;  // 0 none, 1 Hooke, 2 Hertz, 3 HookeMemory, 4 HertzMemory
  int32_t pad;
  // Hooke {kn, en, etat, muc, kr}; Hertz {Es, en, Gs, muc, kr};
  // HookeMemory {kn, en, etat, muc, kt, mur, etarpf, epst};
  // HertzMemory {Es, en, Gs, muc, mur, etarpf, epst}
  double p[8];
  double beta;     // HookeContactForceModel.cpp:19
  double m2sqrt56; // HertzContactForceModel.cpp:21
};

// Number of parameters of a contact model kind (0 if unknown).
__host__ __device__ __forceinline__ int model_nparams(int32_t kind) {
  return kind == 1 || kind == 2 ? 5 : kind == 3 ? 8 : kind == 4 ? 7 : 0;
}

// Static box obstacle.
// kind of a wall's convex: a Box (ext = half edges) or a Rectangle
// (Geometry/src/Rectangle.cpp: ext = (LX/2, LY/2, 0), normal along local z)
enum { WALL_BOX = 0, WALL_RECTANGLE = 1 };

struct Wall {
  double basis[9], origin[3], ext[3];
  double inv_basis[9], inv_origin[3];
  double bb_centre[3], bb_ext[3];
  double crust, req;
  // velocity of an imposed motion (Obstacle::getVelocityAtPoint,
  // Component/src/Obstacle.cpp:128-133: vel + omega x (p - origin))
  double vel[3], omega[3];
  double circ;  // circumscribed radius (Rectangle::computeCircumscribedRadius)
  int32_t material, kind;
};

// Velocity of a rigid body at a point: ParticleKinematics::Velocity
// (Component/src/ParticleKinematics.cpp:229-232).
__host__ __device__ __forceinline__ v3 vel_at(v3 x, v3 v, v3 w, v3 pt) { return add(v, cross(w, sub(pt, x))); }

// performForcesCalculus of HookeContactForceModel (HookeContactForceModel.cpp:44-114)
// and HertzContactForceModel (HertzContactForceModel.cpp:46-114).  Returns
// delFN + delFT, the force on body 0; delM receives the rolling-resistance
// moment on body 0 (body 1 takes -delM), zero when kr = 0.
__host__ __device__ __forceinline__ v3 contact_force(const Model& m, v3 pt, v3 pen, double odist, v3 x0,
                                            v3 v0, v3 w0, double m0, double req0, v3 x1, v3 v1,
                                            v3 w1, double m1, double req1, v3& delM) {
  v3 normal = dvd(pen, norm(pen));
  normal = round3(normal);
  v3 tmpV = sub(vel_at(x0, v0, w0, pt), vel_at(x1, v1, w1, pt));
  v3 v_n = scl(normal, dot(tmpV, normal));
  v3 v_t = sub(tmpV, v_n);
  v3 delFN, delFT;
  double normFN, muc;
  if (m.kind == 1) {
    const double kn = m.p[0], etat_in = m.p[2];
    muc = m.p[3];
    delFN = scl(pen, kn);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double gamman = -2. * m.beta * sqrt(avmass * kn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double etat = etat_in == -1. ? -m.beta * sqrt(kn / avmass) : etat_in;
    double gammat = 2. * etat * avmass;
    delFT = scl(v_t, -gammat);
  } else {
    const double Es = m.p[0], Gs = m.p[2];
    muc = m.p[3];
    double deltan = -odist;
    double Req = 1. / (1. / req0 + 1. / req1);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double sq = sqrt(Req * deltan);
    double Sn = 2. * Es * sq;
    double St = 8. * Gs * sq;
    double kn = (4. / 3.) * Es * sq;
    delFN = scl(pen, kn);
    double gamman = m.m2sqrt56 * m.beta * sqrt(avmass * Sn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double gammat = m.m2sqrt56 * m.beta * sqrt(avmass * St);
    delFT = scl(v_t, -gammat);
  }
  double fn = muc * normFN;
  double ft = norm(delFT);
  if (fn < ft) {
    double nvt = norm(v_t);
    v3 tangent = mk(0., 0., 0.);
    if (nvt > kEps) tangent = dvd(neg(v_t), nvt);
    delFT = scl(tangent, fn);
  }
  // rolling resistance moment (HookeContactForceModel.cpp:95-111,
  // HertzContactForceModel.cpp:99-113): against the relative angular
  // velocity, scaled by the relative rolling speed at the contact point
  delM = mk(0., 0., 0.);
  const double kr = m.p[4];
  if ((m.kind == 1 || m.kind == 2) && kr != 0.) {
    const v3 wrel = sub(w0, w1);
    const double normwrel = norm(wrel);
    const double normwtrel = norm(sub(cross(w0, sub(x0, pt)), cross(w1, sub(x1, pt))));
    if (normwrel > kEps) {
      const double Req = 1. / (1. / req0 + 1. / req1);
      delM = scl(wrel, -(kr * Req * normFN * normwtrel / normwrel));
    }
  }
  return add(delFN, delFT);
}

// The force alone (callers whose models have kr = 0).
__host__ __device__ __forceinline__ v3 contact_force(const Model& m, v3 pt, v3 pen, double odist, v3 x0,
                                            v3 v0, v3 w0, double m0, double req0, v3 x1, v3 v1,
                                            v3 w1, double m1, double req1) {
  v3 delM;
  return contact_force(m, pt, pen, odist, x0, v0, w0, m0, req0, x1, v1, w1, m1, req1, delM);
}

// ---------------------------------------------------------------- memory models
// Quaternion::setRotFromTwoVectors (Geometry/src/Quaternion.cpp:263-286): the
// unit quaternion (qv, qw) turning u onto v; a half turn about an axis normal
// to u when they are opposite.
__host__ __device__ __forceinline__ void rot_two_vectors(v3 u, v3 v, v3& qv, double& qw) {
  const double nn = sqrt(dot(u, u) * dot(v, v));
  double real = nn + dot(u, v);
  v3 vect;
  if (real < 1.e-6 * nn) {
    real = 0.;
    vect = fabs(u.x) > fabs(u.z) ? mk(-u.y, u.x, 0.) : mk(0., -u.z, u.y);
  } else {
    vect = cross(u, v);
  }
  // qq * (1 / Norm(qq)), Norm as Quaternion.cpp:558-562
  const double d = 1. / sqrt(vect.x * vect.x + vect.y * vect.y + vect.z * vect.z + real * real);
  qv = scl(vect, d);
  qw = d * real;
}

// Quaternion::rotateVector (Quaternion.cpp:293-298)
__host__ __device__ __forceinline__ v3 rotate_by(v3 qv, double qw, v3 v) {
  const double nq = norm(qv);
  const double a = qw * qw - nq * nq;
  const double b = 2. * dot(v, qv);
  const double c = 2. * qw;
  return add(add(scl(v, a), scl(qv, b)), scl(cross(qv, v), c));
}

// computeTangentialVector of the memory models (HookeMemoryContactForceModel.cpp:54-65,
// HertzMemoryContactForceModel.cpp:55-66)
__host__ __device__ __forceinline__ v3 tangent_vector(double n_t, v3 ut, v3 kd, double epst) {
  if (norm(ut) > 0.001) return dvd(neg(ut), norm(ut));
  const v3 nut = scl(ut, n_t);
  const v3 s = add(kd, nut);
  if ((norm(kd) > epst || norm(nut) > epst) && norm(s) > epst) return dvd(neg(s), norm(s));
  return mk(0., 0., 0.);
}

// Memory of one contact seen from its first body: cumulative tangential
// spring (kdelta), contact normal of the previous evaluation, rolling-friction
// spring torque (Component.hh:505-508).  The second body keeps the negation.
struct ContactMemory {
  v3 kd, nrm, spr;
};

// performForcesCalculus of HookeMemoryContactForceModel (HookeMemoryContactForceModel.cpp:72-193)
// and HertzMemoryContactForceModel (HertzMemoryContactForceModel.cpp:73-195): the
// force on body 0 (delFN + delFT) is returned, the rolling torque on body 0 in
// delM; mem (body 0's view, found = the contact existed) is updated in place.
__host__ __device__ __forceinline__ v3 memory_contact_force(const Model& m, v3 pt, v3 pen,
                                                   double odist, v3 x0, v3 v0, v3 w0, double m0,
                                                   double req0, v3 x1, v3 v1, v3 w1, double m1,
                                                   double req1, double dt, bool found,
                                                   ContactMemory& mem, v3& delM) {
  const bool hooke = m.kind == 3;
  v3 normal = dvd(pen, norm(pen));
  normal = round3(normal);
  const v3 tmpV = sub(vel_at(x0, v0, w0, pt), vel_at(x1, v1, w1, pt));
  const v3 v_n = scl(normal, dot(tmpV, normal));
  const v3 v_t = sub(tmpV, v_n);
  const double muc = m.p[3];
  v3 delFN;
  double normFN, gamman, gammat, kn, kt, mur, etarpf, epst;
  if (hooke) {
    kn = m.p[0];
    kt = m.p[4];
    mur = m.p[5];
    etarpf = m.p[6];
    epst = m.p[7];
    delFN = scl(pen, kn);
    const double avmass = 1. / (1. / m0 + 1. / m1);
    gamman = -2. * m.beta * sqrt(avmass * kn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    const double etat = m.p[2] == -1. ? -m.beta * sqrt(kn / avmass) : m.p[2];
    gammat = 2. * etat * avmass;
  } else {
    const double Es = m.p[0], Gs = m.p[2];
    mur = m.p[4];
    etarpf = m.p[5];
    epst = m.p[6];
    const double deltan = -odist;
    const double Req = 1. / (1. / req0 + 1. / req1);
    const double avmass = 1. / (1. / m0 + 1. / m1);
    const double sq = sqrt(Req * deltan);
    const double Sn = 2. * Es * sq;
    const double St = 8. * Gs * sq;
    kn = (4. / 3.) * Es * sq;
    delFN = scl(pen, kn);
    gamman = m.m2sqrt56 * m.beta * sqrt(avmass * Sn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    gammat = m.m2sqrt56 * m.beta * sqrt(avmass * St);
    kt = 8. * Gs * sq;
  }
  const v3 viscousFT = scl(v_t, -gammat);
  // rotate the stored spring into the current contact plane
  v3 qv = mk(0., 0., 0.);
  double qw = 0.;
  if (found) {
    rot_two_vectors(mem.nrm, normal, qv, qw);
    mem.kd = rotate_by(qv, qw, mem.kd);
  }
  mem.kd = add(mem.kd, scl(v_t, dt * kt));
  mem.nrm = normal;
  const v3 tij = tangent_vector(gammat, v_t, mem.kd, epst);
  const double normFT = norm(add(neg(mem.kd), viscousFT));
  v3 delFT;
  if (normFT <= muc * normFN) {
    delFT = scl(tij, norm(add(neg(mem.kd), viscousFT)));
  } else {
    // Coulomb limit: the spring is reset to what the limit force needs
    if (!hooke || kt != 0.) mem.kd = dvd(add(scl(tij, -muc * normFN), viscousFT), kt);
    delFT = scl(tij, muc * normFN);
  }
  if (mur != 0.) {
    const double Req = 1. / (1. / req0 + 1. / req1);
    const double kr = 3. * kn * mur * mur * Req * Req;
    const double etar = 3. * gamman * mur * mur * Req * Req;
    const double max_normMk = mur * Req * normFN;
    const v3 wrel = sub(w0, w1);
    if (found) mem.spr = rotate_by(qv, qw, mem.spr);
    mem.spr = sub(mem.spr, scl(wrel, kr * dt));
    const double normMk = norm(mem.spr);
    if (normMk > max_normMk) mem.spr = scl(mem.spr, max_normMk / normMk);
    delM = sub(mem.spr, scl(wrel, etarpf * etar));
  } else {
    delM = mk(0., 0., 0.);
  }
  return add(delFN, delFT);
}

// ClosestPointSPHERESPHERE (Geometry/src/RigidBodyWithCrust.cpp:596-640).
// 1 = contact candidate (dist <= 0), 0 = apart, -1 = ContactError.
__host__ __device__ __forceinline__ int sphere_sphere(v3 A, double rA, double crA, v3 B, double rB,
                                             double crB, v3& c, v3& ov, double& dist) {
  v3 AB = sub(B, A);
  double nAB = norm(AB);
  dist = nAB - (rA + rB);
  if (dist > 0.) return 0;
  if (-dist >= crA + crB) return -1;
  c = add(A, dvd(scl(AB, rA + 0.5 * dist), nAB));
  ov = scl(dvd(AB, nAB), dist);
  return 1;
}

// Transform::operator() (Geometry/src/Transform.cpp:312-317)
__host__ __device__ __forceinline__ v3 xform(const double* b, const double* o, v3 p) {
  return mk(dot(mk(b[0], b[1], b[2]), p) + o[0], dot(mk(b[3], b[4], b[5]), p) + o[1],
            dot(mk(b[6], b[7], b[8]), p) + o[2]);
}

// Box corners, Box::setCornersFaces (Geometry/src/Box.cpp:92-111)
__host__ __device__ __forceinline__ v3 box_corner_pt(const Wall& wl, int k) {
  const int sx = (k == 1 || k == 2 || k == 5 || k == 6) ? 1 : -1;
  const int sy = (k >= 4) ? 1 : -1;
  const int sz = (k == 2 || k == 3 || k == 6 || k == 7) ? 1 : -1;
  return mk(sx < 0 ? -wl.ext[0] : wl.ext[0], sy < 0 ? -wl.ext[1] : wl.ext[1],
            sz < 0 ? -wl.ext[2] : wl.ext[2]);
}

// Box::ContactCornerSPHERE (Box.cpp:725-742)
__host__ __device__ __forceinline__ v3 box_corner(const Wall& wl, v3 g, double r, int k, double& overlap) {
  v3 cp = mk(0., 0., 0.);
  v3 d = sub(g, box_corner_pt(wl, k));
  double nd = norm(d);
  if (nd < r) {
    overlap = nd - r;
    cp = sub(g, scl(d, 1. + 0.5 * overlap / nd));
  }
  return cp;
}

// Box::ContactEdgeSPHERE (Box.cpp:750-769)
__host__ __device__ __forceinline__ v3 box_edge(const Wall& wl, v3 g, double r, int k, int dir,
                                       double& overlap) {
  v3 cp = mk(0., 0., 0.);
  v3 c = box_corner_pt(wl, k);
  v3 pr = g;
  if (dir == 0) pr.x = c.x;
  else if (dir == 1) pr.y = c.y;
  else pr.z = c.z;
  v3 d = sub(pr, c);
  double nd = norm(d);
  if (nd < r) {
    overlap = nd - r;
    cp = sub(g, scl(d, 1. + 0.5 * overlap / nd));
  }
  return cp;
}

// Box::IntersectionPointSPHERE (Box.cpp:466-716)
__host__ __device__ inline v3 box_sphere(const Wall& wl, v3 g, double r, double& overlap) {
  const double ex = wl.ext[0], ey = wl.ext[1], ez = wl.ext[2];
  v3 cp = mk(0., 0., 0.);
  double nd;
  overlap = 1.;
  if (g.x > ex) {
    if (g.y > ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 6, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 5, overlap);
      else cp = box_edge(wl, g, r, 5, 2, overlap);
    } else if (g.y < -ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 2, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 1, overlap);
      else cp = box_edge(wl, g, r, 1, 2, overlap);
    } else {
      if (g.z > ez) cp = box_edge(wl, g, r, 6, 1, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 1, 1, overlap);
      else {
        nd = g.x - ex;
        if (nd < r) { overlap = nd - r; cp = mk(ex + 0.5 * overlap, g.y, g.z); }
      }
    }
  } else if (g.x < -ex) {
    if (g.y > ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 7, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 4, overlap);
      else cp = box_edge(wl, g, r, 4, 2, overlap);
    } else if (g.y < -ey) {
      if (g.z > ez) cp = box_corner(wl, g, r, 3, overlap);
      else if (g.z < -ez) cp = box_corner(wl, g, r, 0, overlap);
      else cp = box_edge(wl, g, r, 0, 2, overlap);
    } else {
      if (g.z > ez) cp = box_edge(wl, g, r, 3, 1, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 0, 1, overlap);
      else {
        nd = -ex - g.x;
        if (nd < r) { overlap = nd - r; cp = mk(-ex - 0.5 * overlap, g.y, g.z); }
      }
    }
  } else {
    if (g.y > ey) {
      if (g.z > ez) cp = box_edge(wl, g, r, 6, 0, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 4, 0, overlap);
      else {
        nd = g.y - ey;
        if (nd < r) { overlap = nd - r; cp = mk(g.x, ey + 0.5 * overlap, g.z); }
      }
    } else if (g.y < -ey) {
      if (g.z > ez) cp = box_edge(wl, g, r, 3, 0, overlap);
      else if (g.z < -ez) cp = box_edge(wl, g, r, 0, 0, overlap);
      else {
        nd = -ey - g.y;
        if (nd < r) { overlap = nd - r; cp = mk(g.x, -ey - 0.5 * overlap, g.z); }
      }
    } else {
      if (g.z > ez) {
        nd = g.z - ez;
        if (nd < r) { overlap = nd - r; cp = mk(g.x, g.y, ez + 0.5 * overlap); }
      } else if (g.z < -ez) {
        nd = -ez - g.z;
        if (nd < r) { overlap = nd - r; cp = mk(g.x, g.y, -ez - 0.5 * overlap); }
      } else {
        overlap = -10000.;
      }
    }
  }
  return cp;
}

// ClosestPointSPHEREBOX with the sphere first (RigidBodyWithCrust.cpp:647-683).
__host__ __device__ __forceinline__ int sphere_wall(const Wall& wl, v3 A, double r, double cr, v3& point,
                                           v3& ov, double& dist) {
  double overlap = 0.;
  v3 local = xform(wl.inv_basis, wl.inv_origin, A);
  v3 cp = box_sphere(wl, local, r, overlap);
  dist = overlap;
  if (overlap < 0.) {
    if (-overlap >= cr + wl.crust) return -1;
    v3 contact = xform(wl.basis, wl.origin, cp);
    v3 AB = sub(contact, A);
    ov = scl(AB, overlap / norm(AB));
    point = contact;
    return 1;
  }
  return 0;
}

// ClosestPointRECTANGLE with the sphere first and the rectangle second
// (Geometry/src/RigidBodyWithCrust.cpp:1107-1226, branch :1172-1218, no
// overlap check), behind ClosestPoint's circumscribed-sphere test
// (:282-296; the bounding-volume option is off by default).  The rectangle's
// normal is its basis' z column, not renormalised (the reference's
// rNorm.normalized() discards its result), turned towards the sphere.  The
// sphere's support point is r d / |d| for d = -n: the reference runs it
// through the sphere's accumulated rotation B as B (r (B^T d) / |B^T d|),
// which is the same point to rounding and exactly it while B is the identity.
// A contact exists only where that point lies beyond the plane and its
// projection falls strictly inside the rectangle (Rectangle::isIn,
// Rectangle.cpp:365-370).
__host__ __device__ __forceinline__ int sphere_rect(const Wall& wl, v3 A, double r, v3& point, v3& ov,
                                           double& dist) {
  const v3 rp = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
  if (!(norm(sub(A, rp)) < r + wl.circ)) return 0;
  v3 n = mk(wl.basis[2], wl.basis[5], wl.basis[8]);
  n = scl(n, copysign(1., dot(n, sub(A, rp))));
  // Sphere::support (Geometry/src/Sphere.cpp:99-109) of -n, then the
  // sphere's transform
  const v3 d = neg(n);
  const double s = norm(d);
  v3 pa = A;
  if (s > kEps) {
    const double rr = r / s;
    pa = add(mk(d.x * rr, d.y * rr, d.z * rr), A);
  }
  if (!(dot(n, sub(pa, rp)) < 0.)) return 0;
  const v3 pb = add(scl(n, dot(sub(rp, pa), n)), pa);
  const v3 l = xform(wl.inv_basis, wl.inv_origin, pb);
  if (!(fabs(l.z) < kEps && fabs(l.y) < wl.ext[1] && fabs(l.x) < wl.ext[0])) return 0;
  point = add(dvd(pa, 2.0), dvd(pb, 2.0));
  ov = sub(pb, pa);
  dist = -norm(ov);
  return 1;
}

// SSPRK3 update of the body-frame angular velocity for a sphere
// (TimeIntegration/src/SecondOrderLeapFrog.cpp:56-66 with
// ParticleKinematicsSphere::computeAngularAccelerationBodyFixed,
// Component/src/ParticleKinematicsSphere.cpp:47-55; coupling factor 1).
__host__ __device__ __forceinline__ v3 ssprk3_sphere(v3 tq, v3 wbf, double I, double dt) {
  v3 a = mk(tq.x / (1. * I), tq.y / (1. * I), tq.z / (1. * I));
  v3 K1 = scl(a, dt);
  v3 K2 = scl(a, dt);
  v3 K3 = scl(a, dt);
  return add(wbf, dvd(add(add(K1, K2), scl(K3, 4.)), 6.));
}

}  // namespace gdem
