// gdem_convex.h -- convex shapes, crust-shrunk transforms and the Johnson and
// signed-volume GJK distances (FP64, reference operand order) for device and
// host.
//
// Paths are relative to src/Grains3D/Grains/.  Polyhedron::support hill-climbs
// from a vertex it keeps between calls (Polyhedron.cpp:265-300); here each call
// starts from the shape's first vertex with neighbours (the value after
// construction or a restart), so results do not depend on call history.  They
// differ only when two vertices tie within 1e-16 along the direction, and then
// only in which tied vertex comes back.
#pragma once
#include <type_traits>

#include "gdem_device.h"

#pragma clang fp contract(off)

namespace gdem {

constexpr int kMaxShapes = 16;
constexpr int kMaxVerts = 8;
enum { SHAPE_SPHERE = 0, SHAPE_BOX = 1, SHAPE_POLYHEDRON = 2 };

struct CShape {
  int32_t type, nv, first_vertex, pad;
  double ext[3];  // box half extents
  double v[kMaxVerts][3];
  int32_t cob_n[kMaxVerts];
  int32_t cob[kMaxVerts][kMaxVerts];
  double circ;         // circumscribed radius
  double bbox_ext[3];  // half extents of the bounding box at identity
  double vol;
  double inertia[6];
};

// Box::support (Geometry/src/Box.cpp:307-316)
__host__ __device__ inline v3 support_box(const CShape& sh, v3 v) {
  double nrm = norm(v);
  if (nrm < kEps) return mk(0., 0., 0.);
  return mk(v.x < 0. ? -sh.ext[0] : sh.ext[0], v.y < 0. ? -sh.ext[1] : sh.ext[1],
            v.z < 0. ? -sh.ext[2] : sh.ext[2]);
}

// Polyhedron::support (Polyhedron.cpp:265-300)
__host__ __device__ inline v3 support_poly(const CShape& sh, v3 v) {
  double nrm = norm(v);
  if (!(nrm > kEps)) return mk(0., 0., 0.);
  int curr = sh.first_vertex, last = -1;
  double h = dot(mk(sh.v[curr][0], sh.v[curr][1], sh.v[curr][2]), v), d = 0.;
  for (int guard = 0; guard < 64; ++guard) {
    int i = 0;
    const int n = sh.cob_n[curr];
    while (i != n) {
      const int c = sh.cob[curr][i];
      if (c != last) {
        d = dot(mk(sh.v[c][0], sh.v[c][1], sh.v[c][2]), v);
        if (!(d - h < 1.e-16)) break;
      }
      ++i;
    }
    if (i == n) break;
    last = curr;
    curr = sh.cob[curr][i];
    h = d;
  }
  return mk(sh.v[curr][0], sh.v[curr][1], sh.v[curr][2]);
}

// Sphere::support (Sphere.cpp:99-109), radius in ext[0]
__host__ __device__ inline v3 support_sphere(const CShape& sh, v3 v) {
  double s = norm(v);
  if (s > kEps) {
    double r = sh.ext[0] / s;
    return mk(v.x * r, v.y * r, v.z * r);
  }
  return mk(0., 0., 0.);
}

__host__ __device__ inline v3 support(const CShape& sh, v3 v) {
  if (sh.type == SHAPE_BOX) return support_box(sh, v);
  if (sh.type == SHAPE_POLYHEDRON) return support_poly(sh, v);
  return support_sphere(sh, v);
}

// Transform: row-major basis and origin (Transform.cpp:312-317)
struct Xf {
  double b[9];
  double o[3];
};

__host__ __device__ inline v3 xf_apply(const Xf& t, v3 p) {
  return mk(dot(mk(t.b[0], t.b[1], t.b[2]), p) + t.o[0],
            dot(mk(t.b[3], t.b[4], t.b[5]), p) + t.o[1],
            dot(mk(t.b[6], t.b[7], t.b[8]), p) + t.o[2]);
}

// Vector3 * Matrix (Matrix.cpp:527-530, Matrix::tdot :238-241)
__host__ __device__ inline v3 vec_mat(v3 v, const double* m) {
  return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
            m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

// getTransformWithCrust (RigidBodyWithCrust.cpp:749-760) with
// Matrix::multiplyByScalingMatrix (Matrix.cpp:347-358)
__host__ __device__ inline Xf crust_xf(const double* basis, v3 origin, const double* sc) {
  Xf t;
  for (int i = 0; i < 3; ++i) {
    t.b[3 * i + 0] = basis[3 * i + 0] * sc[0];
    t.b[3 * i + 1] = basis[3 * i + 1] * sc[1];
    t.b[3 * i + 2] = basis[3 * i + 2] * sc[2];
  }
  t.o[0] = origin.x;
  t.o[1] = origin.y;
  t.o[2] = origin.z;
  return t;
}

// ------------------------------------------------------------ Johnson GJK
// Geometry/src/Convex.cpp:219-430 (sub-algorithm) and :536-627 (driver).
// The tables are indexed by vertex slots and subset masks.  Every index here
// is a compile-time constant: the slot of the new vertex is a template
// argument (one instantiation per slot, picked by a switch) and the subsets
// are enumerated by unrolled loops in the order the reference visits them,
// so the tables live in registers instead of per-lane scratch memory.  The
// operations on each table entry, and their order, are the reference's.
struct Gjk {
  v3 y[4];
  double dp[4][4];
  double det[16][4];
};

// f(integral_constant<unsigned, K>) for K = B, B + D, ... while K != E.
template <unsigned B, unsigned E, int D, class F>
__host__ __device__ inline void gjk_for(F&& f) {
  if constexpr (B != E) {
    f(std::integral_constant<unsigned, B>{});
    gjk_for<(unsigned)((int)B + D), E, D>(f);
  }
}

template <unsigned LAST>
__host__ __device__ inline void gjk_compute_det(unsigned bits, unsigned all_bits, Gjk& g) {
  constexpr unsigned LB = 1u << LAST;
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    if (bits & (1u << i)) g.dp[i][LAST] = g.dp[LAST][i] = dot(g.y[i], g.y[LAST]);
  });
  g.dp[LAST][LAST] = dot(g.y[LAST], g.y[LAST]);
  g.det[LB][LAST] = 1.;
  gjk_for<0, 4, 1>([&](auto J) {
    constexpr unsigned j = decltype(J)::value, sj = 1u << j;
    if constexpr (j != LAST) {
      if (bits & sj) {
        constexpr unsigned s2 = sj | LB;
        g.det[s2][j] = g.dp[LAST][LAST] - g.dp[LAST][j];
        g.det[s2][LAST] = g.dp[j][j] - g.dp[j][LAST];
        gjk_for<0, j, 1>([&](auto K) {
          constexpr unsigned k = decltype(K)::value, sk = 1u << k;
          if constexpr (k != LAST) {
            if (bits & sk) {
              constexpr unsigned s3 = sk | s2;
              g.det[s3][k] = g.det[s2][j] * (g.dp[j][j] - g.dp[j][k]) +
                             g.det[s2][LAST] * (g.dp[LAST][j] - g.dp[LAST][k]);
              g.det[s3][j] = g.det[sk | LB][k] * (g.dp[k][k] - g.dp[k][j]) +
                             g.det[sk | LB][LAST] * (g.dp[LAST][k] - g.dp[LAST][j]);
              g.det[s3][LAST] = g.det[sk | sj][k] * (g.dp[k][k] - g.dp[k][LAST]) +
                                g.det[sk | sj][j] * (g.dp[j][k] - g.dp[j][LAST]);
            }
          }
        });
      }
    }
  });
  if (all_bits == 15) {
    g.det[15][0] = g.det[14][1] * (g.dp[1][1] - g.dp[1][0]) + g.det[14][2] * (g.dp[2][1] - g.dp[2][0]) +
                   g.det[14][3] * (g.dp[3][1] - g.dp[3][0]);
    g.det[15][1] = g.det[13][0] * (g.dp[0][0] - g.dp[0][1]) + g.det[13][2] * (g.dp[2][0] - g.dp[2][1]) +
                   g.det[13][3] * (g.dp[3][0] - g.dp[3][1]);
    g.det[15][2] = g.det[11][0] * (g.dp[0][0] - g.dp[0][2]) + g.det[11][1] * (g.dp[1][0] - g.dp[1][2]) +
                   g.det[11][3] * (g.dp[3][0] - g.dp[3][2]);
    g.det[15][3] = g.det[7][0] * (g.dp[0][0] - g.dp[0][3]) + g.det[7][1] * (g.dp[1][0] - g.dp[1][3]) +
                   g.det[7][2] * (g.dp[2][0] - g.dp[2][3]);
  }
}

template <unsigned S>
__host__ __device__ inline bool gjk_valid(unsigned all_bits, const Gjk& g) {
  bool ok = true;
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value, bit = 1u << i;
    if (all_bits & bit) {
      if constexpr ((S & bit) != 0) {
        if (g.det[S][i] <= 1.0e-20) ok = false;  // EPSILON3
      } else {
        if (g.det[S | bit][i] > 0) ok = false;
      }
    }
  });
  return ok;
}

template <unsigned S>
__host__ __device__ inline v3 gjk_vector(const Gjk& g) {
  double sum = 0.;
  v3 v = mk(0., 0., 0.);
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    if constexpr ((S & (1u << i)) != 0) {
      sum += g.det[S][i];
      v = add(v, scl(g.y[i], g.det[S][i]));
    }
  });
  return scl(v, 1. / sum);
}

template <unsigned S>
__host__ __device__ inline bool gjk_proper(const Gjk& g) {
  bool ok = true;
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    if constexpr ((S & (1u << i)) != 0)
      if (g.det[S][i] <= 1.0e-20) ok = false;
  });
  return ok;
}

// gjk_vector of a run-time subset.
__host__ __device__ inline v3 gjk_vector_of(unsigned bits, const Gjk& g) {
  v3 v = mk(0., 0., 0.);
  gjk_for<1, 16, 1>([&](auto S) {
    if (bits == decltype(S)::value) v = gjk_vector<decltype(S)::value>(g);
  });
  return v;
}

template <unsigned LAST>
__host__ __device__ inline bool gjk_closest(unsigned& bits, unsigned all_bits, Gjk& g, v3& v) {
  constexpr unsigned LB = 1u << LAST;
  gjk_compute_det<LAST>(bits, all_bits, g);
  // the subsets s of bits in decreasing order: the first with s | last valid
  bool found = false;
  unsigned pick = 0;
  gjk_for<15, 0, -1>([&](auto S) {
    constexpr unsigned s = decltype(S)::value;
    if constexpr ((s & LB) == 0) {
      if (!found && (s & bits) == s && gjk_valid<s | LB>(all_bits, g)) {
        found = true;
        pick = s | LB;
      }
    }
  });
  if (found) {
    bits = pick;
    v = gjk_vector_of(bits, g);
    return true;
  }
  if (gjk_valid<LB>(all_bits, g)) {
    bits = LB;
    v = g.y[LAST];
    return true;
  }
  double min_dist2 = __builtin_huge_val();
  gjk_for<15, 0, -1>([&](auto S) {
    constexpr unsigned s = decltype(S)::value;
    if ((s & all_bits) == s && gjk_proper<s>(g)) {
      const v3 u = gjk_vector<s>(g);
      const double dist2 = dot(u, u);
      if (dist2 < min_dist2) {
        min_dist2 = dist2;
        bits = s;
        v = u;
      }
    }
  });
  return false;
}

__host__ __device__ inline bool gjk_degenerate(unsigned all_bits, const v3* y, v3 w) {
  bool deg = false;
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    if ((all_bits & (1u << i)) && y[i].x == w.x && y[i].y == w.y && y[i].z == w.z) deg = true;
  });
  return deg;
}

// New vertex in slot LAST (the lowest free one), then the sub-algorithm.
template <unsigned LAST>
__host__ __device__ inline bool gjk_add(unsigned& bits, unsigned all_bits, Gjk& g, v3* p, v3* q,
                                        v3 w, v3 pl, v3 ql, v3& v) {
  g.y[LAST] = w;
  p[LAST] = pl;
  q[LAST] = ql;
  return gjk_closest<LAST>(bits, all_bits, g, v);
}

// closest_points (Convex.cpp:536-627): tolerance EPSILON (GrainsExec.cpp:51),
// no acceleration.  Returns the distance; pa, pb are the local witness points.
__host__ __device__ inline double closest_points(const CShape& a, const CShape& b, const Xf& a2w,
                                                 const Xf& b2w, v3& pa, v3& pb, int& nbIter) {
  unsigned bits = 0, all_bits = 0;
  v3 p[4], q[4];
  Gjk g;
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    p[i] = q[i] = g.y[i] = mk(0., 0., 0.);
    gjk_for<0, 4, 1>([&](auto J) { g.dp[i][decltype(J)::value] = 0.; });
  });
  gjk_for<0, 16, 1>([&](auto S) {
    gjk_for<0, 4, 1>([&](auto J) { g.det[decltype(S)::value][decltype(J)::value] = 0.; });
  });
  double mu = 0.;
  int numIterations = 0;
  const double relError = kEps;
  const double absError = 1.e-4 * relError;
  const v3 zero = mk(0., 0., 0.);
  v3 v = sub(xf_apply(a2w, support(a, zero)), xf_apply(b2w, support(b, zero)));
  v3 w = v;
  double dist = norm(v);
  while (bits < 15 && dist > 1.0e-15 && numIterations < 1000) {
    ++numIterations;
    const unsigned last = (unsigned)__builtin_ctz(~bits);  // the lowest free slot
    const v3 d = v;
    // the support points of the reference's p[last] and q[last] (stored in
    // gjk_add once the vertex is kept: a vertex that ends the loop is not)
    const v3 pl = support(a, vec_mat(neg(d), a2w.b));
    const v3 ql = support(b, vec_mat(d, b2w.b));
    w = sub(xf_apply(a2w, pl), xf_apply(b2w, ql));
    mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) break;
    if (gjk_degenerate(all_bits, g.y, w)) break;
    all_bits = bits | (1u << last);
    bool ok;
    switch (last) {
      case 0: ok = gjk_add<0>(bits, all_bits, g, p, q, w, pl, ql, v); break;
      case 1: ok = gjk_add<1>(bits, all_bits, g, p, q, w, pl, ql, v); break;
      case 2: ok = gjk_add<2>(bits, all_bits, g, p, q, w, pl, ql, v); break;
      default: ok = gjk_add<3>(bits, all_bits, g, p, q, w, pl, ql, v); break;
    }
    if (!ok) break;
    dist = norm(v);
  }
  // computePoints (Convex.cpp:314-337)
  double sum = 0.;
  pa = mk(0., 0., 0.);
  pb = mk(0., 0., 0.);
  gjk_for<1, 16, 1>([&](auto S) {
    constexpr unsigned s = decltype(S)::value;
    if (bits == s) {
      gjk_for<0, 4, 1>([&](auto I) {
        constexpr unsigned i = decltype(I)::value;
        if constexpr ((s & (1u << i)) != 0) {
          sum += g.det[s][i];
          pa = add(pa, scl(p[i], g.det[s][i]));
          pb = add(pb, scl(q[i], g.det[s][i]));
        }
      });
    }
  });
  const double s = 1. / sum;
  pa = scl(pa, s);
  pb = scl(pb, s);
  nbIter = numIterations;
  return dist;
}

// ------------------------------------------------------------ signed-volume GJK
// closest_points_GJK_SV (Geometry/src/GJK_SV.cpp:428-524): Montanari,
// Petrinic and Barbieri's signed-volume sub-algorithm.  The simplex is reduced
// by comparing the signs of the sub-simplex volumes (areas, lengths) with the
// whole's, on the axis / plane where the simplex's projection is largest; the
// weights are those volume ratios.  Vertex slots are compile-time template
// arguments (the sub-simplices of a simplex are known at compile time; a switch
// on the simplex mask picks the instantiation), so y, the weights and the
// support points stay in registers.  Terms are summed in the reference's order.

__host__ __device__ inline bool sv_same_sign(double a, double b) { return (a > 0) == (b > 0); }

__host__ __device__ inline double sv_comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// sum of lam[i] y[i] over the set bits of `bits` (among those of M), from zero
template <unsigned M>
__host__ __device__ inline v3 sv_point(unsigned bits, const v3* y, const double* lam) {
  v3 r = mk(0., 0., 0.);
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    if constexpr ((M & (1u << i)) != 0)
      if (bits & (1u << i)) r = add(r, scl(y[i], lam[i]));
  });
  return r;
}

// s1d (GJK_SV.cpp:18-89): the segment I1 I2, I1 < I2
template <unsigned I1, unsigned I2>
__host__ __device__ inline void sv_s1d(const v3* y, unsigned& bits, double* lam) {
  const v3 t = sub(y[I2], y[I1]);
  int I = 0;
  double neg_tI = -t.x;
  if (fabs(t.y) > fabs(neg_tI)) { I = 1; neg_tI = -t.y; }
  if (fabs(t.z) > fabs(neg_tI)) { I = 2; neg_tI = -t.z; }
  const double pI = dot(y[I2], t) / dot(t, t) * neg_tI + sv_comp(y[I2], I);
  const double C0 = -sv_comp(y[I2], I) + pI, C1 = sv_comp(y[I1], I) - pI;
  const bool s0 = sv_same_sign(neg_tI, C0), s1 = sv_same_sign(neg_tI, C1);
  if (s0 && s1) {
    lam[I1] = C0 / neg_tI;
    lam[I2] = C1 / neg_tI;
  } else if (s0) {
    bits &= ~(1u << I2);
    lam[I1] = 1.;
  } else {
    bits &= ~(1u << I1);
    lam[I2] = 1.;
  }
}

// One candidate face of the fallback of s2d / s3d: the sub-simplex M's
// closest point, kept if nearer than the best so far.
template <unsigned M, class Sub>
__host__ __device__ inline void sv_try(const v3* y, Sub sub_alg, double& d, unsigned& new_bits,
                                       double* lam) {
  unsigned used = M;
  double nl[4] = {0., 0., 0., 0.};
  sub_alg(used, nl);
  const v3 np = sv_point<M>(used, y, nl);
  const double d_star = dot(np, np);
  if (d_star < d) {
    new_bits = used;
    d = d_star;
    gjk_for<0, 4, 1>([&](auto I) { lam[decltype(I)::value] = nl[decltype(I)::value]; });
  }
}

// s2d (GJK_SV.cpp:94-235): the triangle P0 P1 P2, P0 < P1 < P2
template <unsigned P0, unsigned P1, unsigned P2>
__host__ __device__ inline void sv_s2d(const v3* y, unsigned& bits, double* lam) {
  const v3 a = y[P0], b = y[P1], c = y[P2];
  const v3 n = cross(sub(b, a), sub(c, a));
  const v3 p0 = scl(n, dot(a, n) / dot(n, n));
  // the projection plane of largest signed area
  int ix = 1, iy = 2;
  double mu_max = b.y * c.z + a.y * b.z + c.y * a.z - b.y * a.z - c.y * b.z - a.y * c.z;
  double mu = b.z * a.x + c.z * b.x + a.z * c.x - b.z * c.x - a.z * b.x - c.z * a.x;
  if (fabs(mu) > fabs(mu_max)) { mu_max = mu; ix = 0; }
  mu = b.x * c.y + a.x * b.y + c.x * a.y - b.x * a.y - c.x * b.y - a.x * c.y;
  if (fabs(mu) > fabs(mu_max)) { mu_max = mu; ix = 0; iy = 1; }
  // the areas with one vertex replaced by the projected origin
  const double px = sv_comp(p0, ix), py = sv_comp(p0, iy);
  auto area = [&](v3 u, v3 w) {
    const double ux = sv_comp(u, ix), uy = sv_comp(u, iy), wx = sv_comp(w, ix), wy = sv_comp(w, iy);
    return px * uy + py * wx + ux * wy - px * wy - py * ux - wx * uy;
  };
  const double C0 = area(b, c), C1 = area(c, a), C2 = area(a, b);
  const bool s0 = sv_same_sign(mu_max, C0), s1 = sv_same_sign(mu_max, C1),
             s2 = sv_same_sign(mu_max, C2);
  if (s0 && s1 && s2) {
    lam[P0] = C0 / mu_max;
    lam[P1] = C1 / mu_max;
    lam[P2] = C2 / mu_max;
    return;
  }
  double d = 1.e9;
  unsigned new_bits = 0;
  if (!s0)
    sv_try<(1u << P1) | (1u << P2)>(
        y, [&](unsigned& u, double* l) { sv_s1d<P1, P2>(y, u, l); }, d, new_bits, lam);
  if (!s1)
    sv_try<(1u << P0) | (1u << P2)>(
        y, [&](unsigned& u, double* l) { sv_s1d<P0, P2>(y, u, l); }, d, new_bits, lam);
  if (!s2)
    sv_try<(1u << P0) | (1u << P1)>(
        y, [&](unsigned& u, double* l) { sv_s1d<P0, P1>(y, u, l); }, d, new_bits, lam);
  bits = new_bits;
}

// s3d (GJK_SV.cpp:240-325): the tetrahedron y[0..3]
__host__ __device__ inline void sv_s3d(const v3* y, unsigned& bits, double* lam) {
  const v3 a = y[0], b = y[1], c = y[2], e = y[3];
  const double C0 = e.x * c.y * b.z + c.x * b.y * e.z + b.x * e.y * c.z - b.x * c.y * e.z -
                    c.x * e.y * b.z - e.x * b.y * c.z;
  const double C1 = a.x * c.y * e.z + c.x * e.y * a.z + e.x * a.y * c.z - e.x * c.y * a.z -
                    c.x * a.y * e.z - a.x * e.y * c.z;
  const double C2 = e.x * b.y * a.z + b.x * a.y * e.z + a.x * e.y * b.z - a.x * b.y * e.z -
                    b.x * e.y * a.z - e.x * a.y * b.z;
  const double C3 = a.x * b.y * c.z + b.x * c.y * a.z + c.x * a.y * b.z - c.x * b.y * a.z -
                    b.x * a.y * c.z - a.x * c.y * b.z;
  const double dM = C0 + C1 + C2 + C3;
  const bool s0 = sv_same_sign(dM, C0), s1 = sv_same_sign(dM, C1), s2 = sv_same_sign(dM, C2),
             s3 = sv_same_sign(dM, C3);
  if (s0 && s1 && s2 && s3) {
    lam[0] = C0 / dM;
    lam[1] = C1 / dM;
    lam[2] = C2 / dM;
    lam[3] = C3 / dM;
    return;
  }
  double d = 1.e9;
  unsigned new_bits = 0;
  if (!s0) sv_try<14>(y, [&](unsigned& u, double* l) { sv_s2d<1, 2, 3>(y, u, l); }, d, new_bits, lam);
  if (!s1) sv_try<13>(y, [&](unsigned& u, double* l) { sv_s2d<0, 2, 3>(y, u, l); }, d, new_bits, lam);
  if (!s2) sv_try<11>(y, [&](unsigned& u, double* l) { sv_s2d<0, 1, 3>(y, u, l); }, d, new_bits, lam);
  if (!s3) sv_try<7>(y, [&](unsigned& u, double* l) { sv_s2d<0, 1, 2>(y, u, l); }, d, new_bits, lam);
  bits = new_bits;
}

// sv_subalgorithm (GJK_SV.cpp:370-396): bits holds the simplex with its new
// vertex; returns the closest point of the reduced simplex.
__host__ __device__ inline v3 sv_subalgorithm(const v3* y, unsigned& bits, double* lam) {
  switch (bits) {
    case 1: lam[0] = 1.; break;
    case 2: lam[1] = 1.; break;
    case 4: lam[2] = 1.; break;
    case 8: lam[3] = 1.; break;
    case 3: sv_s1d<0, 1>(y, bits, lam); break;
    case 5: sv_s1d<0, 2>(y, bits, lam); break;
    case 6: sv_s1d<1, 2>(y, bits, lam); break;
    case 9: sv_s1d<0, 3>(y, bits, lam); break;
    case 10: sv_s1d<1, 3>(y, bits, lam); break;
    case 12: sv_s1d<2, 3>(y, bits, lam); break;
    case 7: sv_s2d<0, 1, 2>(y, bits, lam); break;
    case 11: sv_s2d<0, 1, 3>(y, bits, lam); break;
    case 13: sv_s2d<0, 2, 3>(y, bits, lam); break;
    case 14: sv_s2d<1, 2, 3>(y, bits, lam); break;
    default: sv_s3d(y, bits, lam); break;
  }
  return sv_point<15>(bits, y, lam);
}

template <unsigned LAST>
__host__ __device__ inline void sv_put(v3* y, v3* p, v3* q, v3 w, v3 pl, v3 ql) {
  y[LAST] = w;
  p[LAST] = pl;
  q[LAST] = ql;
}

// closest_points_GJK_SV (GJK_SV.cpp:428-524): tolerance EPSILON, no
// acceleration.  Against the Johnson driver: the degenerate test looks at the
// current simplex only, and the witness points are the weighted sums (the
// weights already sum to one).
__host__ __device__ inline double closest_points_sv(const CShape& a, const CShape& b,
                                                    const Xf& a2w, const Xf& b2w, v3& pa, v3& pb,
                                                    int& nbIter) {
  unsigned bits = 0;
  v3 p[4], q[4], y[4];
  double lam[4];
  gjk_for<0, 4, 1>([&](auto I) {
    constexpr unsigned i = decltype(I)::value;
    p[i] = q[i] = y[i] = mk(0., 0., 0.);
    lam[i] = 0.;
  });
  double mu = 0.;
  int numIterations = 0;
  const double relError = kEps;
  const double absError = 1.e-4 * relError;
  const v3 zero = mk(0., 0., 0.);
  v3 v = sub(xf_apply(a2w, support(a, zero)), xf_apply(b2w, support(b, zero)));
  v3 w = v;
  double dist = norm(v);
  while (bits < 15 && dist > 1.0e-15 && numIterations < 1000) {
    ++numIterations;
    const unsigned last = (unsigned)__builtin_ctz(~bits);  // the lowest free slot
    const v3 d = v;
    const v3 pl = support(a, vec_mat(neg(d), a2w.b));
    const v3 ql = support(b, vec_mat(d, b2w.b));
    w = sub(xf_apply(a2w, pl), xf_apply(b2w, ql));
    mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) break;
    if (gjk_degenerate(bits, y, w)) break;
    switch (last) {
      case 0: sv_put<0>(y, p, q, w, pl, ql); break;
      case 1: sv_put<1>(y, p, q, w, pl, ql); break;
      case 2: sv_put<2>(y, p, q, w, pl, ql); break;
      default: sv_put<3>(y, p, q, w, pl, ql); break;
    }
    bits |= 1u << last;
    v = sv_subalgorithm(y, bits, lam);
    dist = norm(v);
  }
  // computePoints (GJK_SV.cpp:348-365)
  pa = sv_point<15>(bits, p, lam);
  pb = sv_point<15>(bits, q, lam);
  nbIter = numIterations;
  return dist;
}

// The narrow phase selected by CollisionDetection Type (Base/src/Grains.cpp:
// 660-666): Johnson "GJK" (the default) or "GJK_SV".
template <bool SV>
__host__ __device__ inline double gjk_distance(const CShape& a, const CShape& b, const Xf& a2w,
                                               const Xf& b2w, v3& pa, v3& pb, int& nbIter) {
  if constexpr (SV)
    return closest_points_sv(a, b, a2w, b2w, pa, pb, nbIter);
  else
    return closest_points(a, b, a2w, b2w, pa, pb, nbIter);
}

// ------------------------------------------------------------ contact-error retry
// RigidBodyWithCrust::ClosestPoint_ErreurHandling (Geometry/src/
// RigidBodyWithCrust.cpp:534-588), reached when an overlap reaches the sum of
// the crusts (Particle.cpp:508-525, SimpleObstacle.cpp:235-252): GJK between
// the bodies shrunk by ten crusts (getTransformWithCrust(10, 0.5), :770-795),
// contact point midway between the witness points, overlap imposed to the
// crust sum along B -> A.  The sphere step keeps no sphere rotation, so the
// sphere's transform is its scaled identity (the rotation cancels for a
// sphere up to rounding).  Returns 1, or -1 when the shrunk bodies still
// touch (the second-level ContactError).

// Scale of one bounding-box axis, shrunk by ten crusts, capped at 0.5.
__host__ __device__ inline double retry_scale(double extent, double crust) {
  const double s = extent < kEps ? 1. : (extent - 10. * crust) / extent;
  return s < 0.5 ? 0.5 : s;
}

__host__ __device__ inline int retry_contact(v3 pA, v3 pB, double gjk_dist, double crA, double crB,
                                             v3& c, v3& ov, double& dist) {
  if (gjk_dist < kEps) return -1;
  v3 ba = sub(pA, pB);
  ba = dvd(ba, norm(ba));  // Vector3::normalize
  c = add(dvd(pA, 2.0), dvd(pB, 2.0));
  const double imposed = 1. * (crA + crB);
  ov = scl(ba, imposed);
  dist = -imposed;
  return 1;
}

__host__ __device__ inline Xf scaled_identity(v3 origin, double s) {
  const double I[9] = {1., 0., 0., 0., 1., 0., 0., 0., 1.};
  const double sc[3] = {s, s, s};
  return crust_xf(I, origin, sc);
}

// Two spheres.  closest_points starts from v = A - B (the supports of the
// zero direction are the centres) and takes the support points along -v and
// +v.  When the shrunk spheres are apart, the difference w of those points
// points along v, the second iteration's support points repeat the first
// ones to rounding and the optimality gap ends it with the one-point simplex
// of the first iteration: the witness points are those of the first
// iteration, here computed with the same operations (two iterations
// reported).  When they touch (w not along v), the second iteration's support
// points lie on the far sides, the segment between the two simplex points
// holds the origin and GJK ends at distance 0: the second-level error.
__host__ __device__ inline int sphere_sphere_retry(v3 A, double rA, double crA, v3 B, double rB,
                                                   double crB, v3& c, v3& ov, double& dist,
                                                   int& nbIter) {
  CShape sa{}, sb{};
  sa.type = sb.type = SHAPE_SPHERE;
  sa.ext[0] = rA;
  sb.ext[0] = rB;
  const Xf a2w = scaled_identity(A, retry_scale(rA, crA));
  const Xf b2w = scaled_identity(B, retry_scale(rB, crB));
  const v3 zero = mk(0., 0., 0.);
  const v3 v = sub(xf_apply(a2w, support(sa, zero)), xf_apply(b2w, support(sb, zero)));
  const v3 pA = xf_apply(a2w, support(sa, vec_mat(neg(v), a2w.b)));
  const v3 pB = xf_apply(b2w, support(sb, vec_mat(v, b2w.b)));
  nbIter = 2;
  const v3 w = sub(pA, pB);
  if (!(dot(w, v) > 0.)) return -1;
  return retry_contact(pA, pB, norm(w), crA, crB, c, ov, dist);
}

// A sphere against a box wall.  GJK between the shrunk sphere and the shrunk
// box converges to their exact closest points: the point of the shrunk box
// nearest to the centre, and the shrunk sphere's surface point towards it.
// Those points are taken here in closed form; the reference stops GJK at its
// tolerance (relative optimality gap 1e-10), which leaves its witness points
// within about 1e-7 m of these at an edge (1e-8 m on a face).  Reports no GJK
// iterations (nbIter = 0): no GJK runs.
__host__ __device__ inline int sphere_wall_retry(const Wall& wl, v3 A, double r, double cr, v3& c,
                                                 v3& ov, double& dist, int& nbIter) {
  nbIter = 0;
  const v3 l = xform(wl.inv_basis, wl.inv_origin, A);
  double q[3] = {l.x, l.y, l.z};
  bool inside = true;
  for (int d = 0; d < 3; ++d) {
    const double e = wl.ext[d] * retry_scale(wl.ext[d], wl.crust);
    if (q[d] > e) { q[d] = e; inside = false; }
    else if (q[d] < -e) { q[d] = -e; inside = false; }
  }
  if (inside) return -1;  // the centre lies in the shrunk box
  const v3 pB = xform(wl.basis, wl.origin, mk(q[0], q[1], q[2]));
  const v3 u = sub(pB, A);
  const double nu = norm(u);
  const double ra = r * retry_scale(r, cr);
  const v3 pA = add(A, scl(dvd(u, nu), ra));
  return retry_contact(pA, pB, nu - ra, cr, wl.crust, c, ov, dist);
}

// A body as seen by the narrow phase
struct CBody {
  const CShape* sh;
  const double* basis;  // row-major
  v3 origin;
  const double* scale;  // crust scaling per axis
  double crust, circ;
};

// RigidBodyWithCrust::ClosestPoint general case (RigidBodyWithCrust.cpp:274-348),
// with the Johnson (SV false) or signed-volume (SV true) GJK (:305-310).
// 1 = candidate (dist may be >= 0), 0 = no bounding-volume overlap,
// -1 = ContactError (shrunk bodies overlap).
template <bool SV>
__host__ __device__ inline int closest_point(const CBody& A, const CBody& B, v3& contact, v3& ov,
                                             double& dist, int& nbIter) {
  nbIter = 0;
  const v3 gab = sub(A.origin, B.origin);
  if (!(norm(gab) < A.circ + B.circ)) return 0;
  const Xf a2w = crust_xf(A.basis, A.origin, A.scale);
  const Xf b2w = crust_xf(B.basis, B.origin, B.scale);
  v3 pa, pb;
  int nit = 0;
  double distance = gjk_distance<SV>(*A.sh, *B.sh, a2w, b2w, pa, pb, nit);
  nbIter = nit;
  if (distance < kEps) return -1;
  pa = xf_apply(a2w, pa);
  pb = xf_apply(b2w, pb);
  const v3 ba = sub(pa, pb);
  contact = add(dvd(pa, 2.0), dvd(pb, 2.0));
  v3 o = dvd(ba, distance);
  o = round3(o);
  ov = scl(o, A.crust + B.crust - distance);
  distance -= A.crust + B.crust;
  dist = distance;
  return 1;
}

// ClosestPoint_ErreurHandling (RigidBodyWithCrust.cpp:534-588) for two
// convex bodies, after ClosestPoint found the crust-shrunk bodies overlapping:
// Johnson GJK between the bodies shrunk by ten crusts (getTransformWithCrust
// (10, 0.5), :770-795: per bounding-box axis (extent - 10 crust) / extent, at
// least 0.5); contact at the midpoint of the witness points, overlap vector
// the unit direction B -> A times the crust sum, overlap minus the crust sum.
// -1: the shrunk bodies still overlap (the second-level ContactError).  The
// reference calls the Johnson GJK here whichever narrow phase is selected
// (:545-546).
__host__ __device__ inline int closest_point_retry(const CBody& A, const CBody& B, v3& contact,
                                                   v3& ov, double& dist, int& nbIter) {
  double sa[3], sb[3];
  for (int d = 0; d < 3; ++d) {
    sa[d] = retry_scale(A.sh->bbox_ext[d], A.crust);
    sb[d] = retry_scale(B.sh->bbox_ext[d], B.crust);
  }
  const Xf a2w = crust_xf(A.basis, A.origin, sa);
  const Xf b2w = crust_xf(B.basis, B.origin, sb);
  v3 pa, pb;
  int nit = 0;
  const double distance = closest_points(*A.sh, *B.sh, a2w, b2w, pa, pb, nit);
  nbIter = nit;
  if (distance < kEps) return -1;
  return retry_contact(xf_apply(a2w, pa), xf_apply(b2w, pb), distance, A.crust, B.crust, contact,
                       ov, dist);
}

// ClosestPointRECTANGLE with a convex particle first (RigidBodyWithCrust.cpp:
// 1107-1226, branch :1172-1218), after ClosestPoint's circumscribed test
// (:282-296): the particle's full shape (not crust-shrunk) through its own
// transform, support of (-n) * basis (Matrix::tdot, Matrix.cpp:238-241),
// then as for a sphere: beyond the plane, projection strictly inside.
__host__ __device__ inline int convex_rect(const CBody& A, const Wall& wl, v3& contact, v3& ov,
                                           double& dist) {
  const v3 rp = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
  if (!(norm(sub(A.origin, rp)) < A.circ + wl.circ)) return 0;
  v3 n = mk(wl.basis[2], wl.basis[5], wl.basis[8]);
  n = scl(n, copysign(1., dot(n, sub(A.origin, rp))));
  const v3 d = neg(n);
  const double* b = A.basis;
  const v3 dl = mk(b[0] * d.x + b[3] * d.y + b[6] * d.z, b[1] * d.x + b[4] * d.y + b[7] * d.z,
                   b[2] * d.x + b[5] * d.y + b[8] * d.z);
  const v3 sp = support(*A.sh, dl);
  const double ao[3] = {A.origin.x, A.origin.y, A.origin.z};
  const v3 pa = xform(A.basis, ao, sp);
  if (!(dot(n, sub(pa, rp)) < 0.)) return 0;
  const v3 pb = add(scl(n, dot(sub(rp, pa), n)), pa);
  const v3 l = xform(wl.inv_basis, wl.inv_origin, pb);
  if (!(fabs(l.z) < kEps && fabs(l.y) < wl.ext[1] && fabs(l.x) < wl.ext[0])) return 0;
  contact = add(dvd(pa, 2.0), dvd(pb, 2.0));
  ov = sub(pb, pa);
  dist = -norm(ov);
  return 1;
}

// ParticleKinematics3D::computeAngularAccelerationBodyFixed
// (Component/src/ParticleKinematics3D.cpp:56-77), predictor mode
__host__ __device__ inline v3 ang_acc3(v3 tq, v3 om, double I0, double I1, double I2) {
  return mk(tq.x / (1. * I0) + om.y * om.z * (I1 - I2) / I0,
            tq.y / (1. * I1) + om.z * om.x * (I2 - I0) / I1,
            tq.z / (1. * I2) + om.x * om.y * (I0 - I1) / I2);
}

// SecondOrderLeapFrog SSPRK3 (TimeIntegration/src/SecondOrderLeapFrog.cpp:56-66)
__host__ __device__ inline v3 ssprk3_3d(v3 tq, v3 wbf, double I0, double I1, double I2,
                                        double dt) {
  v3 K1 = scl(ang_acc3(tq, wbf, I0, I1, I2), dt);
  v3 K2 = scl(ang_acc3(tq, add(wbf, K1), I0, I1, I2), dt);
  v3 K3 = scl(ang_acc3(tq, add(wbf, scl(add(K1, K2), 0.25)), I0, I1, I2), dt);
  return add(wbf, dvd(add(add(K1, K2), scl(K3, 4.)), 6.));
}

// Transform::composeLeftByRotation (Geometry/src/Transform.cpp:140-156)
__host__ __device__ inline void compose_left_rotation(double* mm, quat dq) {
  const double qx = dq.v.x, qy = dq.v.y, qz = dq.v.z, qw = dq.w;
  double nb[9];
  for (int i = 0; i < 3; ++i) {
    double px = qy * mm[6 + i] - qz * mm[3 + i] + qw * mm[0 + i];
    double py = qz * mm[0 + i] - qx * mm[6 + i] + qw * mm[3 + i];
    double pz = qx * mm[3 + i] - qy * mm[0 + i] + qw * mm[6 + i];
    double pw = -qx * mm[0 + i] - qy * mm[3 + i] - qz * mm[6 + i];
    nb[0 + i] = qy * pz - qz * py - pw * qx + qw * px;
    nb[3 + i] = qz * px - qx * pz - pw * qy + qw * py;
    nb[6 + i] = qx * py - qy * px - pw * qz + qw * pz;
  }
  for (int k = 0; k < 9; ++k) mm[k] = nb[k];
}

}  // namespace gdem
