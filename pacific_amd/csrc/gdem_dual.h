// gdem_dual.h -- the sphere contact and Hooke + Coulomb force written over a
// scalar type S, for the differentiated step (gdem_diff.hip).
//
// With S = double the functions compute what sphere_sphere, sphere_wall and
// contact_force (gdem_device.h) compute, in the same operation order.  With
// S = Dual (a value and one directional derivative) the same code carries a
// tangent, so the adjoint kernels form each contact's Jacobian column by
// column with exactly the branch decisions (contact / no contact, wall
// region, Coulomb sliding) of the forward evaluation.  Paths are relative to
// src/Grains3D/Grains/ of the reference.
#pragma once
#include <hip/hip_runtime.h>

#include "gdem_device.h"

#pragma clang fp contract(off)

namespace gdem {

// ---------------------------------------------------------------- scalars
struct Dual {
  double v, d;
};

__host__ __device__ __forceinline__ double val(double a) { return a; }
__host__ __device__ __forceinline__ double val(Dual a) { return a.v; }

__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
}
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__host__ __device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__host__ __device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__host__ __device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__host__ __device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__host__ __device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__host__ __device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__host__ __device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__host__ __device__ __forceinline__ Dual operator/(double a, Dual b) {
  return {a / b.v, -a * b.d / (b.v * b.v)};
}

__host__ __device__ __forceinline__ double sqrt_(double a) { return sqrt(a); }
// d sqrt(a) = da / (2 sqrt(a)); at a = 0 (a zero vector's norm) the
// one-sided derivative is taken as 0
__host__ __device__ __forceinline__ Dual sqrt_(Dual a) {
  const double s = sqrt(a.v);
  return {s, s > 0. ? a.d / (2. * s) : 0.};
}

template <class S>
__host__ __device__ __forceinline__ S lift(double a) {
  return S(a);
}
template <>
__host__ __device__ __forceinline__ Dual lift<Dual>(double a) {
  return {a, 0.};
}

// ---------------------------------------------------------------- vectors
template <class S>
struct V3 {
  S x, y, z;
};

template <class S>
__host__ __device__ __forceinline__ V3<S> mk3(S x, S y, S z) {
  V3<S> r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
template <class S>
__host__ __device__ __forceinline__ V3<S> lift3(v3 a) {
  return mk3(lift<S>(a.x), lift<S>(a.y), lift<S>(a.z));
}
template <class S>
__host__ __device__ __forceinline__ v3 val3(V3<S> a) {
  return mk(val(a.x), val(a.y), val(a.z));
}
template <class S>
__host__ __device__ __forceinline__ V3<S> add(V3<S> a, V3<S> b) {
  return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class S>
__host__ __device__ __forceinline__ V3<S> sub(V3<S> a, V3<S> b) {
  return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class S, class T>
__host__ __device__ __forceinline__ V3<S> scl(V3<S> a, T d) {
  return mk3(a.x * d, a.y * d, a.z * d);
}
template <class S, class T>
__host__ __device__ __forceinline__ V3<S> dvd(V3<S> a, T d) {
  return mk3(a.x / d, a.y / d, a.z / d);
}
template <class S>
__host__ __device__ __forceinline__ V3<S> neg(V3<S> a) {
  return mk3(-a.x, -a.y, -a.z);
}
template <class S>
__host__ __device__ __forceinline__ S dot(V3<S> a, V3<S> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class S>
__host__ __device__ __forceinline__ V3<S> cross(V3<S> a, V3<S> b) {
  return mk3(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
}
template <class S>
__host__ __device__ __forceinline__ S norm(V3<S> a) {
  return sqrt_(a.x * a.x + a.y * a.y + a.z * a.z);
}
// Group3::round (Geometry/src/Group3.cpp:390-395): a component below EPSILON is
// set to zero, its derivative with it
template <class S>
__host__ __device__ __forceinline__ V3<S> round3(V3<S> a) {
  const S z = lift<S>(0.);
  return mk3(fabs(val(a.x)) < kEps ? z : a.x, fabs(val(a.y)) < kEps ? z : a.y,
             fabs(val(a.z)) < kEps ? z : a.z);
}

// ---------------------------------------------------------------- geometry
// ClosestPointSPHERESPHERE (Geometry/src/RigidBodyWithCrust.cpp:596-640)
template <class S>
__host__ __device__ __forceinline__ int ss_contact(V3<S> A, double rA, double crA, V3<S> B,
                                                   double rB, double crB, V3<S>& c, V3<S>& ov,
                                                   S& dist) {
  V3<S> AB = sub(B, A);
  S nAB = norm(AB);
  dist = nAB - (rA + rB);
  if (val(dist) > 0.) return 0;
  if (-val(dist) >= crA + crB) return -1;
  c = add(A, dvd(scl(AB, rA + 0.5 * dist), nAB));
  ov = scl(dvd(AB, nAB), dist);
  return 1;
}

template <class S>
__host__ __device__ __forceinline__ V3<S> xform_t(const double* b, const double* o, V3<S> p) {
  return mk3(dot(lift3<S>(mk(b[0], b[1], b[2])), p) + o[0],
             dot(lift3<S>(mk(b[3], b[4], b[5])), p) + o[1],
             dot(lift3<S>(mk(b[6], b[7], b[8])), p) + o[2]);
}

// Box::ContactCornerSPHERE (Box.cpp:725-742)
template <class S>
__host__ __device__ __forceinline__ V3<S> box_corner_t(const Wall& wl, V3<S> g, double r, int k,
                                                       S& overlap) {
  V3<S> cp = lift3<S>(mk(0., 0., 0.));
  V3<S> d = sub(g, lift3<S>(box_corner_pt(wl, k)));
  S nd = norm(d);
  if (val(nd) < r) {
    overlap = nd - r;
    cp = sub(g, scl(d, 1. + 0.5 * overlap / nd));
  }
  return cp;
}

// Box::ContactEdgeSPHERE (Box.cpp:750-769)
template <class S>
__host__ __device__ __forceinline__ V3<S> box_edge_t(const Wall& wl, V3<S> g, double r, int k, int dir,
                                                     S& overlap) {
  V3<S> cp = lift3<S>(mk(0., 0., 0.));
  const V3<S> c = lift3<S>(box_corner_pt(wl, k));
  V3<S> pr = g;
  if (dir == 0) pr.x = c.x;
  else if (dir == 1) pr.y = c.y;
  else pr.z = c.z;
  V3<S> d = sub(pr, c);
  S nd = norm(d);
  if (val(nd) < r) {
    overlap = nd - r;
    cp = sub(g, scl(d, 1. + 0.5 * overlap / nd));
  }
  return cp;
}

// Box::IntersectionPointSPHERE (Box.cpp:466-716)
template <class S>
__host__ __device__ inline V3<S> box_sphere_t(const Wall& wl, V3<S> g, double r, S& overlap) {
  const double ex = wl.ext[0], ey = wl.ext[1], ez = wl.ext[2];
  const double gx = val(g.x), gy = val(g.y), gz = val(g.z);
  V3<S> cp = lift3<S>(mk(0., 0., 0.));
  S nd;
  overlap = lift<S>(1.);
  if (gx > ex) {
    if (gy > ey) {
      if (gz > ez) cp = box_corner_t(wl, g, r, 6, overlap);
      else if (gz < -ez) cp = box_corner_t(wl, g, r, 5, overlap);
      else cp = box_edge_t(wl, g, r, 5, 2, overlap);
    } else if (gy < -ey) {
      if (gz > ez) cp = box_corner_t(wl, g, r, 2, overlap);
      else if (gz < -ez) cp = box_corner_t(wl, g, r, 1, overlap);
      else cp = box_edge_t(wl, g, r, 1, 2, overlap);
    } else {
      if (gz > ez) cp = box_edge_t(wl, g, r, 6, 1, overlap);
      else if (gz < -ez) cp = box_edge_t(wl, g, r, 1, 1, overlap);
      else {
        nd = g.x - ex;
        if (val(nd) < r) { overlap = nd - r; cp = mk3(ex + 0.5 * overlap, g.y, g.z); }
      }
    }
  } else if (gx < -ex) {
    if (gy > ey) {
      if (gz > ez) cp = box_corner_t(wl, g, r, 7, overlap);
      else if (gz < -ez) cp = box_corner_t(wl, g, r, 4, overlap);
      else cp = box_edge_t(wl, g, r, 4, 2, overlap);
    } else if (gy < -ey) {
      if (gz > ez) cp = box_corner_t(wl, g, r, 3, overlap);
      else if (gz < -ez) cp = box_corner_t(wl, g, r, 0, overlap);
      else cp = box_edge_t(wl, g, r, 0, 2, overlap);
    } else {
      if (gz > ez) cp = box_edge_t(wl, g, r, 3, 1, overlap);
      else if (gz < -ez) cp = box_edge_t(wl, g, r, 0, 1, overlap);
      else {
        nd = (-ex) - g.x;
        if (val(nd) < r) { overlap = nd - r; cp = mk3(-ex - 0.5 * overlap, g.y, g.z); }
      }
    }
  } else {
    if (gy > ey) {
      if (gz > ez) cp = box_edge_t(wl, g, r, 6, 0, overlap);
      else if (gz < -ez) cp = box_edge_t(wl, g, r, 4, 0, overlap);
      else {
        nd = g.y - ey;
        if (val(nd) < r) { overlap = nd - r; cp = mk3(g.x, ey + 0.5 * overlap, g.z); }
      }
    } else if (gy < -ey) {
      if (gz > ez) cp = box_edge_t(wl, g, r, 3, 0, overlap);
      else if (gz < -ez) cp = box_edge_t(wl, g, r, 0, 0, overlap);
      else {
        nd = (-ey) - g.y;
        if (val(nd) < r) { overlap = nd - r; cp = mk3(g.x, -ey - 0.5 * overlap, g.z); }
      }
    } else {
      if (gz > ez) {
        nd = g.z - ez;
        if (val(nd) < r) { overlap = nd - r; cp = mk3(g.x, g.y, ez + 0.5 * overlap); }
      } else if (gz < -ez) {
        nd = (-ez) - g.z;
        if (val(nd) < r) { overlap = nd - r; cp = mk3(g.x, g.y, -ez - 0.5 * overlap); }
      } else {
        overlap = lift<S>(-10000.);
      }
    }
  }
  return cp;
}

// ClosestPointSPHEREBOX with the sphere first (RigidBodyWithCrust.cpp:647-683)
template <class S>
__host__ __device__ __forceinline__ int sw_contact(const Wall& wl, V3<S> A, double r, double cr,
                                                   V3<S>& point, V3<S>& ov, S& dist) {
  S overlap = lift<S>(0.);
  V3<S> local = xform_t(wl.inv_basis, wl.inv_origin, A);
  V3<S> cp = box_sphere_t(wl, local, r, overlap);
  dist = overlap;
  if (val(overlap) < 0.) {
    if (-val(overlap) >= cr + wl.crust) return -1;
    V3<S> contact = xform_t(wl.basis, wl.origin, cp);
    V3<S> AB = sub(contact, A);
    ov = scl(AB, overlap / norm(AB));
    point = contact;
    return 1;
  }
  return 0;
}

// ---------------------------------------------------------------- force
// HookeContactForceModel::performForcesCalculus (HookeContactForceModel.cpp:44-114)
// with etat = -1 (gamma_t from beta) and kr = 0; beta carries its derivative
// with respect to en.  Returns the force on body 0.
template <class S>
__host__ __device__ __forceinline__ V3<S> hooke_t(V3<S> pt, V3<S> pen, V3<S> x0, V3<S> v0, V3<S> w0,
                                                  double m0, V3<S> x1, V3<S> v1, V3<S> w1,
                                                  double m1, S kn, S beta, S muc) {
  V3<S> normal = dvd(pen, norm(pen));
  normal = round3(normal);
  V3<S> tmpV = sub(add(v0, cross(w0, sub(pt, x0))), add(v1, cross(w1, sub(pt, x1))));
  V3<S> v_n = scl(normal, dot(tmpV, normal));
  V3<S> v_t = sub(tmpV, v_n);
  V3<S> delFN = scl(pen, kn);
  const double avmass = 1. / (1. / m0 + 1. / m1);
  S gamman = -2. * beta * sqrt_(avmass * kn);
  delFN = sub(delFN, scl(v_n, gamman));
  S normFN = norm(delFN);
  S etat = -beta * sqrt_(kn / avmass);
  S gammat = 2. * etat * avmass;
  V3<S> delFT = scl(v_t, -gammat);
  S fn = muc * normFN;
  S ft = norm(delFT);
  if (val(fn) < val(ft)) {
    S nvt = norm(v_t);
    V3<S> tangent = lift3<S>(mk(0., 0., 0.));
    if (val(nvt) > kEps) tangent = dvd(neg(v_t), nvt);
    delFT = scl(tangent, fn);
  }
  return add(delFN, delFT);
}

// Contact parameters of one system: kn, beta(en), mu.
template <class S>
struct HookeParams {
  S kn, beta, mu;
};

// Particle-particle contact, A = first body, B = second body moved to A's
// image: forces and torques on both.  False when not in contact.
template <class S>
__host__ __device__ __forceinline__ bool pair_torsors(V3<S> xA, V3<S> vA, V3<S> wA, double rA,
                                                      double crA, double mA, V3<S> xB, V3<S> vB,
                                                      V3<S> wB, double rB, double crB, double mB,
                                                      const HookeParams<S>& p, V3<S>& FA,
                                                      V3<S>& TA, V3<S>& FB, V3<S>& TB) {
  V3<S> c, ov;
  S dist;
  const int rc = ss_contact(xA, rA, crA, xB, rB, crB, c, ov, dist);
  if (rc <= 0 || !(val(dist) < 0.)) return false;
  const V3<S> f = hooke_t(c, ov, xA, vA, wA, mA, xB, vB, wB, mB, p.kn, p.beta, p.mu);
  FA = f;
  TA = cross(sub(c, xA), f);
  const V3<S> fn = neg(f);
  FB = fn;
  TB = cross(sub(c, xB), fn);
  return true;
}

// Particle-wall contact (SimpleObstacle::InterAction, SimpleObstacle.cpp:222-271)
template <class S>
__host__ __device__ __forceinline__ bool wall_torsor(const Wall& wl, V3<S> xA, V3<S> vA, V3<S> wA,
                                                     double r, double cr, double m,
                                                     const HookeParams<S>& p, V3<S>& FA,
                                                     V3<S>& TA) {
  V3<S> c, ov;
  S dist;
  const int rc = sw_contact(wl, xA, r, cr, c, ov, dist);
  if (rc <= 0 || !(val(dist) < 0.)) return false;
  const V3<S> z = lift3<S>(mk(0., 0., 0.));
  const V3<S> f = hooke_t(c, ov, xA, vA, wA, m, lift3<S>(mk(wl.origin[0], wl.origin[1], wl.origin[2])),
                          z, z, 1.e20, p.kn, p.beta, p.mu);
  FA = f;
  TA = cross(sub(c, xA), f);
  return true;
}

// ---------------------------------------------------------------- reverse mode
// Inputs of one particle-particle contact (A first, B second moved to A's image).
struct PairIn {
  v3 xA, vA, wA, xB, vB, wB;
  double rA, crA, mA, rB, crB, mB;
};
// Adjoints of the contact's force and torque on each body.
struct PairAdj {
  v3 FA, TA, FB, TB;
};
// Gradient of <adjoint, (F_A, T_A, F_B, T_B)> with respect to every input
// and to (kn, en, mu).
struct PairGrad {
  v3 xA, vA, wA, xB, vB, wB, th;
};

// Forward and reverse of hooke_t (kr = 0) at contact point c and overlap
// vector ov, followed by the torsors F_A = f, T_A = (c - xA) x f, F_B = -f,
// T_B = (c - xB) x (-f), for the adjoint adj.  Writes the velocity and angular
// velocity gradients of both bodies, the position gradients through the lever
// arms (g.xA, g.xB) and (kn, beta, mu) gradients in g.th; returns those of c
// and ov in gc, gov.  A wall is body B at rest with mass 1e20 and a zero
// adjoint (wall_torsor).
__host__ __device__ inline void hooke_torsor_vjp(v3 c, v3 ov, const PairIn& in, const PairAdj& adj,
                                                 double kn, double beta, double mu, PairGrad& g,
                                                 v3& gc, v3& gov) {
  const double np = norm(ov);
  const v3 nr = dvd(ov, np);
  const bool kx = !(fabs(nr.x) < kEps), ky = !(fabs(nr.y) < kEps), kz = !(fabs(nr.z) < kEps);
  const v3 normal = mk(kx ? nr.x : 0., ky ? nr.y : 0., kz ? nr.z : 0.);
  const v3 rA = sub(c, in.xA), rB = sub(c, in.xB);
  const v3 tmpV = sub(add(in.vA, cross(in.wA, rA)), add(in.vB, cross(in.wB, rB)));
  const double vns = dot(tmpV, normal);
  const v3 vn = scl(normal, vns);
  const v3 vt = sub(tmpV, vn);
  const double avm = 1. / (1. / in.mA + 1. / in.mB);
  const double sq1 = sqrt(avm * kn);
  const double gamman = -2. * beta * sq1;
  const v3 dFN = sub(scl(ov, kn), scl(vn, gamman));
  const double nFN = norm(dFN);
  const double sq2 = sqrt(kn / avm);
  const double etat = -beta * sq2;
  const double gammat = 2. * etat * avm;
  const v3 dFT0 = scl(vt, -gammat);
  const double fn = mu * nFN;
  const bool slip = fn < norm(dFT0);
  const double nvt = norm(vt);
  const bool has_t = nvt > kEps;
  const v3 tangent = slip && has_t ? dvd(neg(vt), nvt) : mk(0., 0., 0.);
  const v3 f = add(dFN, slip ? scl(tangent, fn) : dFT0);
  // torsors
  const v3 gf = sub(add(sub(adj.FA, adj.FB), cross(adj.TA, rA)), cross(adj.TB, rB));
  v3 grA = cross(f, adj.TA), grB = cross(adj.TB, f);
  v3 gvt = mk(0., 0., 0.), gdFN = gf;
  double gkn = 0., gbeta = 0., gmu = 0.;
  if (slip) {
    // f = dFN + tangent * mu |dFN|
    const double gfn = dot(gf, tangent);
    if (has_t) {
      const v3 gt = scl(gf, fn);
      gvt = add(scl(gt, -1. / nvt), scl(vt, dot(gt, vt) / (nvt * nvt * nvt)));
    }
    gmu = gfn * nFN;
    if (nFN > 0.) gdFN = add(gdFN, scl(dFN, gfn * mu / nFN));
  } else {
    // f = dFN + vt * (-2 avm (-beta sqrt(kn / avm)))
    gvt = scl(gf, -gammat);
    const double getat = 2. * avm * -dot(gf, vt);
    gbeta = -sq2 * getat;
    if (sq2 > 0.) gkn = -beta * getat * (0.5 / sq2) / avm;
  }
  // dFN = ov kn - vn (-2 beta sqrt(avm kn))
  gkn += dot(gdFN, ov);
  const double ggamman = -dot(gdFN, vn);
  gbeta += -2. * sq1 * ggamman;
  if (sq1 > 0.) gkn += -2. * beta * ggamman * (0.5 / sq1) * avm;
  gov = scl(gdFN, kn);
  const v3 gvn = sub(scl(gdFN, -gamman), gvt);
  // vt = tmpV - vn, vn = normal (tmpV . normal)
  const double gvns = dot(gvn, normal);
  const v3 gtmpV = add(gvt, scl(normal, gvns));
  v3 gnormal = add(scl(gvn, vns), scl(tmpV, gvns));
  gnormal = mk(kx ? gnormal.x : 0., ky ? gnormal.y : 0., kz ? gnormal.z : 0.);
  // normal = ov / |ov|
  gov = add(gov, sub(dvd(gnormal, np), scl(ov, dot(gnormal, ov) / (np * np * np))));
  // tmpV = vA + wA x rA - vB - wB x rB
  g.vA = gtmpV;
  g.vB = neg(gtmpV);
  g.wA = cross(rA, gtmpV);
  g.wB = neg(cross(rB, gtmpV));
  grA = add(grA, cross(gtmpV, in.wA));
  grB = sub(grB, cross(gtmpV, in.wB));
  // rA = c - xA, rB = c - xB
  gc = add(grA, grB);
  g.xA = neg(grA);
  g.xB = neg(grB);
  g.th = mk(gkn, gbeta, gmu);
}

// Reverse-mode derivative of one pair contact (pair_torsors over ss_contact
// and hooke_t, kr = 0), written out by hand: the forward values once, then
// each operation's transpose in reverse order, with the branch decisions
// (contact, Coulomb sliding, zero tangential velocity, rounded normal
// components) of the forward evaluation.  One call gives all 21 partial
// derivatives that pair_torsors over Dual gives one at a time.  dbeta is
// d beta / d en.  False when not in contact.
__host__ __device__ inline bool pair_vjp(const PairIn& in, const PairAdj& adj, double kn,
                                         double beta, double dbeta, double mu, PairGrad& g) {
  // ss_contact
  const v3 AB = sub(in.xB, in.xA);
  const double nAB = norm(AB);
  const double dist = nAB - (in.rA + in.rB);
  if (dist > 0. || -dist >= in.crA + in.crB || !(dist < 0.)) return false;
  const double sc = in.rA + 0.5 * dist;
  const v3 c = add(in.xA, dvd(scl(AB, sc), nAB));
  const v3 ov = scl(dvd(AB, nAB), dist);
  v3 gc, gov;
  hooke_torsor_vjp(c, ov, in, adj, kn, beta, mu, g, gc, gov);
  // c = xA + AB sc / |AB|, ov = AB dist / |AB|, sc = rA + dist / 2,
  // dist = |AB| - (rA + rB)
  const double n2 = nAB * nAB, n3 = n2 * nAB;
  const double gcAB = dot(gc, AB), govAB = dot(gov, AB);
  const v3 gAB = add(add(scl(gc, sc / nAB), scl(AB, gcAB * 0.5 / n2 - gcAB * sc / n3)),
                     add(scl(gov, dist / nAB), scl(AB, govAB / n2 - govAB * dist / n3)));
  g.xA = sub(add(g.xA, gc), gAB);
  g.xB = add(g.xB, gAB);
  g.th.y *= dbeta;
  return true;
}

// Transpose of a row-major 3x3 matrix applied to a vector.
__host__ __device__ inline v3 mat_t_vec(const double* b, v3 u) {
  return mk(b[0] * u.x + b[3] * u.y + b[6] * u.z, b[1] * u.x + b[4] * u.y + b[7] * u.z,
            b[2] * u.x + b[5] * u.y + b[8] * u.z);
}

// Reverse-mode derivative of a particle-wall contact (wall_torsor over
// sw_contact, box_sphere_t and hooke_t): gradients of <(aF, aT), (F, T)>
// with respect to the particle's x, v, w (g.xA, g.vA, g.wA) and (kn, en, mu)
// (g.th).  False when not in contact.
__host__ __device__ inline bool wall_vjp(const Wall& wl, v3 x, v3 v, v3 w, double r, double cr,
                                         double m, v3 aF, v3 aT, double kn, double beta,
                                         double dbeta, double mu, PairGrad& g) {
  // sw_contact with box_sphere_t, values only, remembering the region
  const v3 lg = xform(wl.inv_basis, wl.inv_origin, x);
  const double ex = wl.ext[0], ey = wl.ext[1], ez = wl.ext[2];
  const int sx = lg.x > ex ? 1 : (lg.x < -ex ? -1 : 0);
  const int sy = lg.y > ey ? 1 : (lg.y < -ey ? -1 : 0);
  const int sz = lg.z > ez ? 1 : (lg.z < -ez ? -1 : 0);
  const int nout = (sx != 0) + (sy != 0) + (sz != 0);
  if (nout == 0) return false;  // centre inside the box: overlap -10000, a contact error
  // corner of the region (the box corner, or the edge's corner point)
  const v3 corner = mk(sx < 0 ? -ex : ex, sy < 0 ? -ey : ey, sz < 0 ? -ez : ez);
  double overlap, nd = 0., kk = 0.;
  v3 cp, d = mk(0., 0., 0.);
  int face = -1;
  if (nout == 1) {
    face = sx ? 0 : (sy ? 1 : 2);
    const int sg = sx + sy + sz;
    const double gc = face == 0 ? lg.x : (face == 1 ? lg.y : lg.z);
    const double e = face == 0 ? ex : (face == 1 ? ey : ez);
    nd = sg > 0 ? gc - e : (-e) - gc;
    if (!(nd < r)) return false;
    overlap = nd - r;
    const double cf = sg > 0 ? e + 0.5 * overlap : -e - 0.5 * overlap;
    cp = mk(face == 0 ? cf : lg.x, face == 1 ? cf : lg.y, face == 2 ? cf : lg.z);
  } else {
    // an edge keeps the coordinate inside the box's range, a corner none
    const v3 pr = mk(sx ? lg.x : corner.x, sy ? lg.y : corner.y, sz ? lg.z : corner.z);
    d = sub(nout == 3 ? lg : pr, corner);
    nd = norm(d);
    if (!(nd < r)) return false;
    overlap = nd - r;
    kk = 1. + 0.5 * overlap / nd;
    cp = sub(lg, scl(d, kk));
  }
  if (!(overlap < 0.) || -overlap >= cr + wl.crust) return false;
  const v3 contact = xform(wl.basis, wl.origin, cp);
  const v3 AB = sub(contact, x);
  const double nAB = norm(AB);
  const v3 ov = scl(AB, overlap / nAB);
  PairIn in;
  in.xA = x; in.vA = v; in.wA = w; in.rA = r; in.crA = cr; in.mA = m;
  in.xB = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
  in.vB = in.wB = mk(0., 0., 0.);
  in.rB = 0.; in.crB = wl.crust; in.mB = 1.e20;
  PairAdj adj;
  adj.FA = aF; adj.TA = aT;
  adj.FB = adj.TB = mk(0., 0., 0.);
  v3 gc, gov;
  hooke_torsor_vjp(contact, ov, in, adj, kn, beta, mu, g, gc, gov);
  // ov = AB overlap / |AB|, AB = contact - x
  const double govAB = dot(gov, AB);
  const double n3 = nAB * nAB * nAB;
  const v3 gAB = sub(scl(gov, overlap / nAB), scl(AB, govAB * overlap / n3));
  double gover = govAB / nAB;
  const v3 gcontact = add(gc, gAB);
  v3 gx = sub(g.xA, gAB);
  // contact = basis cp + origin
  const v3 gcp = mat_t_vec(wl.basis, gcontact);
  v3 glg;
  if (face >= 0) {
    // cp: the face coordinate e + overlap / 2 (or -e - overlap / 2), the
    // others the centre's; overlap = nd - r with nd = +-(g - e)
    const int sg = sx + sy + sz;
    const double gcf = face == 0 ? gcp.x : (face == 1 ? gcp.y : gcp.z);
    gover += sg > 0 ? 0.5 * gcf : -0.5 * gcf;
    const double gnd = sg > 0 ? gover : -gover;
    glg = mk(face == 0 ? gnd : gcp.x, face == 1 ? gnd : gcp.y, face == 2 ? gnd : gcp.z);
  } else {
    // cp = g - d (1 + overlap / (2 nd)), overlap = nd - r, nd = |d|
    const double gk = -dot(gcp, d);
    gover += gk * 0.5 / nd;
    const double gnd = gover + gk * (-0.5 * overlap / (nd * nd));
    const v3 gd = add(scl(gcp, -kk), scl(d, gnd / nd));
    glg = add(gcp, nout == 3 ? gd : mk(sx ? gd.x : 0., sy ? gd.y : 0., sz ? gd.z : 0.));
  }
  gx = add(gx, mat_t_vec(wl.inv_basis, glg));
  g.xA = gx;
  g.th.y *= dbeta;
  return true;
}

}  // namespace gdem
