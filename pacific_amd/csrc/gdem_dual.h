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

}  // namespace gdem
