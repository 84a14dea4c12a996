/*
 * gdem_oracle_convex.c -- CPU restatement of the Grains3D step for convex
 * particles (boxes, polyhedra) with the Johnson or the signed-volume GJK narrow
 * phase.
 *
 * TEST INFRASTRUCTURE ONLY (see gdem_oracle.h).  Paths are relative to
 * src/Grains3D/Grains/ of the reference.  Same structure as the sphere oracle:
 * particle list in insertion order, per-cell lists with head insertion, the
 * 13-cell half stencil, and the reference's operand order everywhere.
 *
 * Restatement choice: Polyhedron::support hill-climbs from a vertex kept in a
 * mutable member between calls (Polyhedron.cpp:265-300); here every call
 * starts from the polyhedron's first vertex with neighbours, so the result
 * does not depend on call history (it differs from the reference only when two
 * vertices tie within 1e-16 in the support direction).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gdem_oracle.h"
#include "orc_math.h"

#define MAXV 8
#define MAXSH 16

typedef struct {
  int type;          /* ORC_SHAPE_* */
  double ext[3];     /* box half extents / sphere radius in ext[0] */
  int nv;
  double v[MAXV][3];
  int cob_n[MAXV];
  int cob[MAXV][MAXV];
  int first_vertex;  /* Polyhedron.cpp:362-364 */
  double circ;       /* circumscribed radius */
  double bbox_ext[3];
  double vol;
  double inertia[6];
} shape_t;

typedef struct {
  int kind;
  double p[5];
  double beta, m2sqrt56;
} cmodel_t;

typedef struct {
  double basis[9], origin[3], ext[3];
  double bb_centre[3], bb_ext[3];
  double crust, req, circ, scale[3];
  int material;
  int rect;  /* a Rectangle (ext = LX/2, LY/2, 0) */
} cwall_t;

struct orc_csys {
  orc_config cfg;
  int nc[3];
  double edge[3];
  int ncell;
  int n;
  int nshape;
  shape_t shapes[MAXSH];
  double *x, *v, *w, *wbf, *q, *basis, *F, *T;
  double *mass, *inertia, *req, *crust, *scale;
  int32_t *shape, *mat;
  int nwalls;
  cwall_t walls[ORC_MAX_WALLS];
  cmodel_t models[ORC_MAX_MATERIALS][ORC_MAX_MATERIALS];
  int *head, *next, *prev, *cell;
  int stencil[13][3];
  int log_cap, log_n;
  int32_t *log_id0, *log_id1;
  double *log_pt, *log_ov, *log_d, *log_f;
  int64_t ngjk;
  int gjk_sv; /* 1: closest_points_GJK_SV, 0: Johnson closest_points */
  int *hist;  /* per-particle Polyhedron::m_curr_vertex, when histories are kept */
  int err_code;
  int32_t err_id0, err_id1;
  /* time of the current force evaluation; AppCollision contact features */
  double time;
  double ov_mean, ov_count, ov_max, ov_time_max, gjk_mean, gjk_count;
};

/* AppCollision::addToContactsFeatures (Base/src/AppCollision.cpp:160-190):
   running mean and maximum of the overlaps, running mean of the reported GJK
   iterations, over every closest-point result of the run. */
static void c_contact_features(orc_csys* s, double overlap, int nbIter) {
  if (overlap < 0.) {
    overlap = fabs(overlap);
    s->ov_mean = (s->ov_mean * s->ov_count + overlap) / (s->ov_count + 1.);
    s->ov_count += 1.;
    if (overlap > s->ov_max) {
      s->ov_max = overlap;
      s->ov_time_max = s->time;
    }
  }
  if (nbIter) {
    s->gjk_mean = (s->gjk_mean * s->gjk_count + (double)nbIter) / (s->gjk_count + 1.);
    s->gjk_count += 1.;
  }
}

static int cset_error(orc_csys* s, int code, int a, int b) {
  if (s->err_code == ORC_OK) {
    s->err_code = code;
    s->err_id0 = a;
    s->err_id1 = b;
  }
  return code;
}

/* ------------------------------------------------------------- supports */
/* Box::support (Geometry/src/Box.cpp:307-316) */
static v3 support_box(const shape_t* sh, v3 v) {
  double nrm = norm(v);
  if (nrm < ORC_EPSILON) return mk(0., 0., 0.);
  return mk(v.x < 0. ? -sh->ext[0] : sh->ext[0], v.y < 0. ? -sh->ext[1] : sh->ext[1],
            v.z < 0. ? -sh->ext[2] : sh->ext[2]);
}

/* Polyhedron::support (Geometry/src/Polyhedron.cpp:265-300), hill climbing on
   the vertex adjacency, moving only on a gain of at least 1e-16.  hist, when
   given, is the body's m_curr_vertex (Polyhedron.hh:134): the climb starts
   there and leaves it at the vertex returned, as the reference's mutable
   member does; without it every call starts from the first vertex with
   neighbours (the value after construction, Polyhedron.cpp:362-364). */
static v3 support_poly(const shape_t* sh, v3 v, int* hist) {
  double nrm = norm(v);
  if (!(nrm > ORC_EPSILON)) return mk(0., 0., 0.);
  int curr = hist ? *hist : sh->first_vertex, last = -1;
  double h = dot(ld(sh->v[curr]), v), d = 0.;
  for (;;) {
    int i = 0, n = sh->cob_n[curr];
    while (i != n &&
           (sh->cob[curr][i] == last || (d = dot(ld(sh->v[sh->cob[curr][i]]), v)) - h < 1.e-16))
      ++i;
    if (i == n) break;
    last = curr;
    curr = sh->cob[curr][i];
    h = d;
  }
  if (hist) *hist = curr;
  return ld(sh->v[curr]);
}

/* Sphere::support (Geometry/src/Sphere.cpp:99-109) */
static v3 support_sphere(const shape_t* sh, v3 v) {
  double s = norm(v);
  if (s > ORC_EPSILON) {
    double r = sh->ext[0] / s;
    return mk(v.x * r, v.y * r, v.z * r);
  }
  return mk(0., 0., 0.);
}

static v3 support(const shape_t* sh, v3 v, int* hist) {
  if (sh->type == ORC_SHAPE_BOX) return support_box(sh, v);
  if (sh->type == ORC_SHAPE_POLYHEDRON) return support_poly(sh, v, hist);
  return support_sphere(sh, v);
}

/* Transform: basis (row-major) + origin; operator() (Transform.cpp:312-317) */
typedef struct { double b[9]; double o[3]; } xf_t;

static v3 xf_apply(const xf_t* t, v3 p) {
  return mk(dot(mk(t->b[0], t->b[1], t->b[2]), p) + t->o[0],
            dot(mk(t->b[3], t->b[4], t->b[5]), p) + t->o[1],
            dot(mk(t->b[6], t->b[7], t->b[8]), p) + t->o[2]);
}

/* Vector3 * Matrix (Matrix.cpp:527-530 with Matrix::tdot, :238-241) */
static v3 vec_mat(v3 v, const double* m) {
  return mk(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
            m[2] * v.x + m[5] * v.y + m[8] * v.z);
}

/* RigidBodyWithCrust::getTransformWithCrust (RigidBodyWithCrust.cpp:749-760):
   the transform composed with the crust scaling (Matrix::multiplyByScalingMatrix,
   Matrix.cpp:347-358, scales column j by s_j) */
static void crust_xf(const double* basis, const double* origin, const double* sc, xf_t* t) {
  for (int i = 0; i < 3; ++i) {
    t->b[3 * i + 0] = basis[3 * i + 0] * sc[0];
    t->b[3 * i + 1] = basis[3 * i + 1] * sc[1];
    t->b[3 * i + 2] = basis[3 * i + 2] * sc[2];
    t->o[i] = origin[i];
  }
}

/* ------------------------------------------------------------- Johnson GJK
   Geometry/src/Convex.cpp:219-430 (sub-algorithm) and :536-627 (driver). */
typedef struct {
  v3 y[4];
  double dp[4][4];
  double det[16][4];
} gjk_t;

static void compute_det(unsigned bits, unsigned last, unsigned last_bit, unsigned all_bits,
                        gjk_t* g) {
  v3* y = g->y;
  double (*dp)[4] = g->dp;
  double (*det)[4] = g->det;
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if (bits & bit) dp[i][last] = dp[last][i] = dot(y[i], y[last]);
  dp[last][last] = dot(y[last], y[last]);
  det[last_bit][last] = 1.;
  for (unsigned j = 0, sj = 1; j < 4; ++j, sj <<= 1) {
    if (bits & sj) {
      unsigned s2 = sj | last_bit;
      det[s2][j] = dp[last][last] - dp[last][j];
      det[s2][last] = dp[j][j] - dp[j][last];
      for (unsigned k = 0, sk = 1; k < j; ++k, sk <<= 1) {
        if (bits & sk) {
          unsigned s3 = sk | s2;
          det[s3][k] = det[s2][j] * (dp[j][j] - dp[j][k]) + det[s2][last] * (dp[last][j] - dp[last][k]);
          det[s3][j] = det[sk | last_bit][k] * (dp[k][k] - dp[k][j]) +
                       det[sk | last_bit][last] * (dp[last][k] - dp[last][j]);
          det[s3][last] = det[sk | sj][k] * (dp[k][k] - dp[k][last]) +
                          det[sk | sj][j] * (dp[j][k] - dp[j][last]);
        }
      }
    }
  }
  if (all_bits == 15) {
    det[15][0] = det[14][1] * (dp[1][1] - dp[1][0]) + det[14][2] * (dp[2][1] - dp[2][0]) +
                 det[14][3] * (dp[3][1] - dp[3][0]);
    det[15][1] = det[13][0] * (dp[0][0] - dp[0][1]) + det[13][2] * (dp[2][0] - dp[2][1]) +
                 det[13][3] * (dp[3][0] - dp[3][1]);
    det[15][2] = det[11][0] * (dp[0][0] - dp[0][2]) + det[11][1] * (dp[1][0] - dp[1][2]) +
                 det[11][3] * (dp[3][0] - dp[3][2]);
    det[15][3] = det[7][0] * (dp[0][0] - dp[0][3]) + det[7][1] * (dp[1][0] - dp[1][3]) +
                 det[7][2] * (dp[2][0] - dp[2][3]);
  }
}

static int valid(unsigned s, unsigned all_bits, double (*det)[4]) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1) {
    if (all_bits & bit) {
      if (s & bit) {
        if (det[s][i] <= 1.0e-20) return 0; /* EPSILON3 */
      } else if (det[s | bit][i] > 0) {
        return 0;
      }
    }
  }
  return 1;
}

static v3 compute_vector(unsigned bits, const v3* y, double (*det)[4]) {
  double sum = 0.;
  v3 v = mk(0., 0., 0.);
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1) {
    if (bits & bit) {
      sum += det[bits][i];
      v = add(v, scl(y[i], det[bits][i]));
    }
  }
  return scl(v, 1. / sum);
}

static void compute_points(unsigned bits, double det[16][4], const v3* p, const v3* q, v3* p1,
                           v3* p2) {
  double sum = 0.;
  *p1 = mk(0., 0., 0.);
  *p2 = mk(0., 0., 0.);
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1) {
    if (bits & bit) {
      sum += det[bits][i];
      *p1 = add(*p1, scl(p[i], det[bits][i]));
      *p2 = add(*p2, scl(q[i], det[bits][i]));
    }
  }
  double s = 1. / sum;
  *p1 = scl(*p1, s);
  *p2 = scl(*p2, s);
}

static int proper(unsigned s, double (*det)[4]) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if ((s & bit) && det[s][i] <= 1.0e-20) return 0;
  return 1;
}

static int closest(unsigned* bits, unsigned last, unsigned last_bit, unsigned all_bits, gjk_t* g,
                   v3* v) {
  unsigned s;
  compute_det(*bits, last, last_bit, all_bits, g);
  for (s = *bits; s; --s) {
    if ((s & *bits) == s) {
      if (valid(s | last_bit, all_bits, g->det)) {
        *bits = s | last_bit;
        *v = compute_vector(*bits, g->y, g->det);
        return 1;
      }
    }
  }
  if (valid(last_bit, all_bits, g->det)) {
    *bits = last_bit;
    *v = g->y[last];
    return 1;
  }
  double min_dist2 = INFINITY;
  for (s = all_bits; s; --s) {
    if ((s & all_bits) == s) {
      if (proper(s, g->det)) {
        v3 u = compute_vector(s, g->y, g->det);
        double dist2 = dot(u, u);
        if (dist2 < min_dist2) {
          min_dist2 = dist2;
          *bits = s;
          *v = u;
        }
      }
    }
  }
  return 0;
}

static int degenerate(unsigned all_bits, const v3* y, v3 w) {
  for (unsigned i = 0, bit = 1; i < 4; ++i, bit <<= 1)
    if ((all_bits & bit) && y[i].x == w.x && y[i].y == w.y && y[i].z == w.z) return 1;
  return 0;
}

/* closest_points (Convex.cpp:536-627), tolerance GrainsExec::m_colDetTolerance
   = EPSILON (GrainsExec.cpp:51), no acceleration */
static double closest_points(const shape_t* a, const shape_t* b, const xf_t* a2w,
                             const xf_t* b2w, v3* pa, v3* pb, int* nbIter, int* ha, int* hb) {
  unsigned bits = 0, last = 0, last_bit = 0, all_bits = 0;
  v3 p[4], q[4];
  gjk_t g;
  memset(&g, 0, sizeof(g));
  memset(p, 0, sizeof(p));
  memset(q, 0, sizeof(q));
  double mu = 0.;
  int numIterations = 0;
  const double relError = ORC_EPSILON;
  const double absError = 1.e-4 * relError;
  v3 zero = mk(0., 0., 0.);
  v3 v = sub(xf_apply(a2w, support(a, zero, ha)), xf_apply(b2w, support(b, zero, hb)));
  v3 w = v;
  double dist = norm(v);
  while (bits < 15 && dist > 1.0e-15 && numIterations < 1000) {
    ++numIterations;
    last = 0;
    last_bit = 1;
    while (bits & last_bit) {
      ++last;
      last_bit <<= 1;
    }
    v3 d = v;
    p[last] = support(a, vec_mat(neg(d), a2w->b), ha);
    q[last] = support(b, vec_mat(d, b2w->b), hb);
    w = sub(xf_apply(a2w, p[last]), xf_apply(b2w, q[last]));
    mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) break;
    if (degenerate(all_bits, g.y, w)) break;
    g.y[last] = w;
    all_bits = bits | last_bit;
    if (!closest(&bits, last, last_bit, all_bits, &g, &v)) break;
    dist = norm(v);
  }
  compute_points(bits, g.det, p, q, pa, pb);
  *nbIter = numIterations;
  return dist;
}

/* ------------------------------------------------ signed-volume GJK (GJK_SV)
   Montanari, Petrinic and Barbieri's signed-volume sub-algorithm as the
   reference writes it (Geometry/src/GJK_SV.cpp): the simplex is reduced by
   comparing the signs of the sub-simplex volumes (areas, lengths) with the
   whole's, projecting onto the axis / plane of largest extent; the weights
   are those volume ratios.  Sums keep the reference's term order. */
static double vc(v3 a, unsigned i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* compareSigns (GJK_SV.cpp:9-13) */
static unsigned same_sign(double a, double b) { return (a > 0) == (b > 0); }

/* s1d (GJK_SV.cpp:18-89): segment y[i1] y[i2], i1 < i2 the two set bits */
static void sv_s1d(const v3* y, unsigned* bits, double* lambdas) {
  unsigned i1 = 4, i2 = 4;
  for (unsigned i = 0; i < 4; ++i) {
    if (*bits & (1u << i)) {
      if (i1 == 4) i1 = i;
      else { i2 = i; break; }
    }
  }
  v3 t = sub(y[i2], y[i1]);
  unsigned I = 0;
  double neg_tI = -t.x;
  if (fabs(t.y) > fabs(neg_tI)) { I = 1; neg_tI = -t.y; }
  if (fabs(t.z) > fabs(neg_tI)) { I = 2; neg_tI = -t.z; }
  double pI = dot(y[i2], t) / dot(t, t) * neg_tI + vc(y[i2], I);
  double C0 = -vc(y[i2], I) + pI, C1 = vc(y[i1], I) - pI;
  unsigned s0 = same_sign(neg_tI, C0), s1 = same_sign(neg_tI, C1);
  if (s0 + s1 == 2) {
    lambdas[i1] = C0 / neg_tI;
    lambdas[i2] = C1 / neg_tI;
  } else if (s0) {
    *bits &= ~(1u << i2);
    lambdas[i1] = 1.;
  } else {
    *bits &= ~(1u << i1);
    lambdas[i2] = 1.;
  }
}

/* the combination sum of lambdas[i] y[i] over the set bits, from zero */
static v3 sv_point(unsigned bits, const v3* y, const double* lambdas) {
  v3 r = mk(0., 0., 0.);
  for (unsigned i = 0; i < 4; ++i)
    if (bits & (1u << i)) r = add(r, scl(y[i], lambdas[i]));
  return r;
}

/* s2d (GJK_SV.cpp:94-235): triangle of the three set bits */
static void sv_s2d(const v3* y, unsigned* bits, double* lambdas) {
  unsigned k = 0, idx[3] = {0, 0, 0};
  for (unsigned i = 0; i < 4; ++i)
    if (*bits & (1u << i)) idx[k++] = i;
  const v3 a = y[idx[0]], b = y[idx[1]], c = y[idx[2]];
  v3 n = cross(sub(b, a), sub(c, a));
  v3 p0 = scl(n, dot(a, n) / dot(n, n));
  /* signed area of the projection with the largest magnitude */
  unsigned ix = 1, iy = 2;
  double mu_max = b.y * c.z + a.y * b.z + c.y * a.z - b.y * a.z - c.y * b.z - a.y * c.z;
  double mu = b.z * a.x + c.z * b.x + a.z * c.x - b.z * c.x - a.z * b.x - c.z * a.x;
  if (fabs(mu) > fabs(mu_max)) { mu_max = mu; ix = 0; }
  mu = b.x * c.y + a.x * b.y + c.x * a.y - b.x * a.y - c.x * b.y - a.x * c.y;
  if (fabs(mu) > fabs(mu_max)) { mu_max = mu; ix = 0; iy = 1; }
  /* areas with each vertex replaced by the projected origin: vertex k's
     area uses the other two, (u, w) = (b, c), (c, a), (a, b) */
  const v3 uu[3] = {b, c, a}, ww[3] = {c, a, b};
  double C[3];
  unsigned sg[3];
  const double px = vc(p0, ix), py = vc(p0, iy);
  for (int j = 0; j < 3; ++j) {
    const double ux = vc(uu[j], ix), uy = vc(uu[j], iy), wx = vc(ww[j], ix), wy = vc(ww[j], iy);
    C[j] = px * uy + py * wx + ux * wy - px * wy - py * ux - wx * uy;
    sg[j] = same_sign(mu_max, C[j]);
  }
  if (sg[0] + sg[1] + sg[2] == 3) {
    for (int j = 0; j < 3; ++j) lambdas[idx[j]] = C[j] / mu_max;
    return;
  }
  double d = 1.e9;
  unsigned new_bits = 0;
  for (int j = 0; j < 3; ++j) {
    if (sg[j]) continue;
    unsigned used = *bits & ~(1u << idx[j]);
    double nl[4] = {0., 0., 0., 0.};
    sv_s1d(y, &used, nl);
    v3 np = sv_point(used, y, nl);
    double d_star = dot(np, np);
    if (d_star < d) {
      new_bits = used;
      d = d_star;
      for (int i = 0; i < 4; ++i) lambdas[i] = nl[i];
    }
  }
  *bits = new_bits;
}

/* s3d (GJK_SV.cpp:240-325): the tetrahedron y[0..3] */
static void sv_s3d(const v3* y, unsigned* bits, double* lambdas) {
  const v3 a = y[0], b = y[1], c = y[2], e = y[3];
  double C[4];
  C[0] = e.x * c.y * b.z + c.x * b.y * e.z + b.x * e.y * c.z - b.x * c.y * e.z - c.x * e.y * b.z -
         e.x * b.y * c.z;
  C[1] = a.x * c.y * e.z + c.x * e.y * a.z + e.x * a.y * c.z - e.x * c.y * a.z - c.x * a.y * e.z -
         a.x * e.y * c.z;
  C[2] = e.x * b.y * a.z + b.x * a.y * e.z + a.x * e.y * b.z - a.x * b.y * e.z - b.x * e.y * a.z -
         e.x * a.y * b.z;
  C[3] = a.x * b.y * c.z + b.x * c.y * a.z + c.x * a.y * b.z - c.x * b.y * a.z - b.x * a.y * c.z -
         a.x * c.y * b.z;
  const double dM = C[0] + C[1] + C[2] + C[3];
  unsigned sg[4], nsame = 0;
  for (int j = 0; j < 4; ++j) nsame += sg[j] = same_sign(dM, C[j]);
  if (nsame == 4) {
    for (int i = 0; i < 4; ++i) lambdas[i] = C[i] / dM;
    return;
  }
  double d = 1.e9;
  unsigned new_bits = 0;
  for (unsigned j = 0; j < 4; ++j) {
    if (sg[j]) continue;
    unsigned used = *bits & ~(1u << j);
    double nl[4] = {0., 0., 0., 0.};
    sv_s2d(y, &used, nl);
    v3 np = sv_point(used, y, nl);
    double d_star = dot(np, np);
    if (d_star < d) {
      new_bits = used;
      d = d_star;
      for (int i = 0; i < 4; ++i) lambdas[i] = nl[i];
    }
  }
  *bits = new_bits;
}

/* sv_subalgorithm (GJK_SV.cpp:370-396) */
static v3 sv_subalgorithm(const v3* y, unsigned* bits, double* lambdas) {
  unsigned used = 0;
  for (unsigned i = 0; i < 4; ++i) used += (*bits >> i) & 1u;
  if (used == 1) {
    for (unsigned i = 0; i < 4; ++i)
      if (*bits & (1u << i)) lambdas[i] = 1.;
  } else if (used == 2) {
    sv_s1d(y, bits, lambdas);
  } else if (used == 3) {
    sv_s2d(y, bits, lambdas);
  } else {
    sv_s3d(y, bits, lambdas);
  }
  return sv_point(*bits, y, lambdas);
}

/* closest_points_GJK_SV (GJK_SV.cpp:428-524), tolerance EPSILON, no
   acceleration.  Differs from the Johnson driver in the sub-algorithm, in the
   degenerate test (against the current simplex only) and in the witness
   points (the weights sum to one). */
static double closest_points_sv(const shape_t* a, const shape_t* b, const xf_t* a2w,
                                const xf_t* b2w, v3* pa, v3* pb, int* nbIter, int* ha,
                                int* hb) {
  unsigned bits = 0, last = 0;
  v3 p[4], q[4], y[4];
  double lambdas[4] = {0., 0., 0., 0.};
  memset(p, 0, sizeof(p));
  memset(q, 0, sizeof(q));
  memset(y, 0, sizeof(y));
  double mu = 0.;
  int numIterations = 0;
  const double relError = ORC_EPSILON;
  const double absError = 1.e-4 * relError;
  v3 zero = mk(0., 0., 0.);
  v3 v = sub(xf_apply(a2w, support(a, zero, ha)), xf_apply(b2w, support(b, zero, hb)));
  v3 w = v;
  double dist = norm(v);
  while (bits < 15 && dist > 1.0e-15 && numIterations < 1000) {
    ++numIterations;
    for (last = 0; last < 4; ++last)
      if (!(bits & (1u << last))) break;
    v3 d = v;
    p[last] = support(a, vec_mat(neg(d), a2w->b), ha);
    q[last] = support(b, vec_mat(d, b2w->b), hb);
    w = sub(xf_apply(a2w, p[last]), xf_apply(b2w, q[last]));
    mu = dist - dot(v, w) / dist;
    if (mu < dist * relError || mu < absError) break;
    if (degenerate(bits, y, w)) break;
    y[last] = w;
    bits |= 1u << last;
    v = sv_subalgorithm(y, &bits, lambdas);
    dist = norm(v);
  }
  /* computePoints (GJK_SV.cpp:348-365) */
  *pa = sv_point(bits, p, lambdas);
  *pb = sv_point(bits, q, lambdas);
  *nbIter = numIterations;
  return dist;
}

/* ha, hb: the bodies' support histories (support_poly), or NULL */
static double gjk_distance(int sv, const shape_t* a, const shape_t* b, const xf_t* a2w,
                           const xf_t* b2w, v3* pa, v3* pb, int* nbIter, int* ha, int* hb) {
  return sv ? closest_points_sv(a, b, a2w, b2w, pa, pb, nbIter, ha, hb)
            : closest_points(a, b, a2w, b2w, pa, pb, nbIter, ha, hb);
}

/* ------------------------------------------------------------- shapes */
static void finish_shape(shape_t* sh) {
  /* bounding box of the shape at identity (Convex::bbox, Convex.cpp:29-41):
     extent along axis d = (support(e_d).d - support(-e_d).d) / 2 */
  for (int d = 0; d < 3; ++d) {
    v3 e = mk(d == 0 ? 1. : 0., d == 1 ? 1. : 0., d == 2 ? 1. : 0.);
    v3 sp = support(sh, e, NULL), sm = support(sh, neg(e), NULL);
    double mx = dot(e, sp), mn = dot(e, sm);
    sh->bbox_ext[d] = (mx - mn) / 2.;
  }
}

int orc_c_create(const orc_config* cfg, orc_csys** out) {
  orc_csys* s = (orc_csys*)calloc(1, sizeof(orc_csys));
  if (!s) return ORC_ERR_ARG;
  s->cfg = *cfg;
  for (int d = 0; d < 3; ++d) {
    int n = 1;
    double e = cfg->size[d];
    if (cfg->cell_size > ORC_EPSILON) {
      n = (int)((cfg->size[d] + ORC_EPSILON) / cfg->cell_size);
      if (!n) n = 1;
      e = cfg->size[d] / n;
    }
    if (cfg->periodic[d] && n < 3) { free(s); return ORC_ERR_ARG; }
    s->nc[d] = n;
    s->edge[d] = e;
  }
  s->ncell = s->nc[0] * s->nc[1] * s->nc[2];
  static const int cel[3][13] = {{-1, 0, +1, +1, -1, 0, +1, -1, 0, +1, -1, 0, +1},
                                 {0, 0, 0, 0, +1, +1, +1, +1, +1, +1, +1, +1, +1},
                                 {+1, +1, +1, 0, +1, +1, +1, 0, 0, 0, -1, -1, -1}};
  for (int j = 0; j < 13; ++j)
    for (int d = 0; d < 3; ++d) s->stencil[j][d] = cel[d][j];
  s->head = (int*)malloc(sizeof(int) * (size_t)s->ncell);
  for (int c = 0; c < s->ncell; ++c) s->head[c] = -1;
  *out = s;
  return ORC_OK;
}

static void cfree_particles(orc_csys* s) {
  free(s->x); free(s->v); free(s->w); free(s->wbf); free(s->q); free(s->basis); free(s->F);
  free(s->T); free(s->mass); free(s->inertia); free(s->req); free(s->crust); free(s->scale);
  free(s->shape); free(s->mat); free(s->next); free(s->prev); free(s->cell);
  free(s->hist);
  s->hist = NULL;
}

void orc_c_destroy(orc_csys* s) {
  if (!s) return;
  cfree_particles(s);
  free(s->head);
  free(s->log_id0); free(s->log_id1); free(s->log_pt); free(s->log_ov); free(s->log_d);
  free(s->log_f);
  free(s);
}

/* Box(LX, LY, LZ) (Box.cpp:27-33, half extents), Box::BuildInertia (:181-205),
   Box::computeCircumscribedRadius (:226-229), Box::getVolume (:273-279) */
int orc_c_add_box_shape(orc_csys* s, const double* half) {
  if (s->nshape >= MAXSH) return -1;
  shape_t* sh = &s->shapes[s->nshape];
  memset(sh, 0, sizeof(*sh));
  sh->type = ORC_SHAPE_BOX;
  memcpy(sh->ext, half, sizeof(double) * 3);
  const double ex = half[0], ey = half[1], ez = half[2];
  sh->vol = 8.0 * ex * ey * ez;
  sh->inertia[0] = 8.0 * ex * ey * ez * (ey * ey + ez * ez) / 3.0;
  sh->inertia[3] = 8.0 * ex * ey * ez * (ex * ex + ez * ez) / 3.0;
  sh->inertia[5] = 8.0 * ex * ey * ez * (ey * ey + ex * ex) / 3.0;
  sh->circ = norm(ld(half));
  finish_shape(sh);
  return s->nshape++;
}

/* Sphere(r) (Geometry/src/Sphere.cpp): radius in ext[0]; only for closest
   points (the particles of the convex step are boxes and polyhedra). */
int orc_c_add_sphere_shape(orc_csys* s, double r) {
  if (s->nshape >= MAXSH || !(r > 0.)) return -1;
  shape_t* sh = &s->shapes[s->nshape];
  memset(sh, 0, sizeof(*sh));
  sh->type = ORC_SHAPE_SPHERE;
  sh->ext[0] = r;
  sh->vol = 4.0 * ORC_PI * r * r * r / 3.0;
  sh->circ = r;
  finish_shape(sh);
  return s->nshape++;
}

/* GrainsExec::computeVolumeInertiaContrib (Base/src/GrainsExec.cpp:748-800),
   tetrahedron (0, A2, A3, A4) after F. Tonon, J. Math. Stat. 1 (2004) */
static void tet_contrib(v3 A2, v3 A3, v3 A4, double* vol, double* in) {
  double x1 = 0., x2 = A2.x, x3 = A3.x, x4 = A4.x, y1 = 0., y2 = A2.y, y3 = A3.y, y4 = A4.y,
         z1 = 0., z2 = A2.z, z3 = A3.z, z4 = A4.z;
  double det = fabs((x2 - x1) * (y3 - y1) * (z4 - z1) + (y2 - y1) * (z3 - z1) * (x4 - x1) +
                    (z2 - z1) * (x3 - x1) * (y4 - y1) - (z2 - z1) * (y3 - y1) * (x4 - x1) -
                    (x2 - x1) * (z3 - z1) * (y4 - y1) - (y2 - y1) * (x3 - x1) * (z4 - z1));
  *vol += det / 6.;
  in[0] += det * (y1 * y1 + y1 * y2 + y2 * y2 + y1 * y3 + y2 * y3 + y3 * y3 + y1 * y4 + y2 * y4 +
                  y3 * y4 + y4 * y4 + z1 * z1 + z1 * z2 + z2 * z2 + z1 * z3 + z2 * z3 + z3 * z3 +
                  z1 * z4 + z2 * z4 + z3 * z4 + z4 * z4) / 60.;
  in[1] -= det * (2. * x1 * z1 + x2 * z1 + x3 * z1 + x4 * z1 + x1 * z2 + 2. * x2 * z2 + x3 * z2 +
                  x4 * z2 + x1 * z3 + x2 * z3 + 2. * x3 * z3 + x4 * z3 + x1 * z4 + x2 * z4 +
                  x3 * z4 + 2. * x4 * z4) / 120.;
  in[2] -= det * (2. * x1 * y1 + x2 * y1 + x3 * y1 + x4 * y1 + x1 * y2 + 2. * x2 * y2 + x3 * y2 +
                  x4 * y2 + x1 * y3 + x2 * y3 + 2. * x3 * y3 + x4 * y3 + x1 * y4 + x2 * y4 +
                  x3 * y4 + 2. * x4 * y4) / 120.;
  in[3] += det * (x1 * x1 + x1 * x2 + x2 * x2 + x1 * x3 + x2 * x3 + x3 * x3 + x1 * x4 + x2 * x4 +
                  x3 * x4 + x4 * x4 + z1 * z1 + z1 * z2 + z2 * z2 + z1 * z3 + z2 * z3 + z3 * z3 +
                  z1 * z4 + z2 * z4 + z3 * z4 + z4 * z4) / 60.;
  in[4] -= det * (2. * y1 * z1 + y2 * z1 + y3 * z1 + y4 * z1 + y1 * z2 + 2. * y2 * z2 + y3 * z2 +
                  y4 * z2 + y1 * z3 + y2 * z3 + 2. * y3 * z3 + y4 * z3 + y1 * z4 + y2 * z4 +
                  y3 * z4 + 2. * y4 * z4) / 120.;
  in[5] += det * (x1 * x1 + x1 * x2 + x2 * x2 + x1 * x3 + x2 * x3 + x3 * x3 + x1 * x4 + x2 * x4 +
                  x3 * x4 + x4 * x4 + y1 * y1 + y1 * y2 + y2 * y2 + y1 * y3 + y2 * y3 + y3 * y3 +
                  y1 * y4 + y2 * y4 + y3 * y4 + y4 * y4) / 60.;
}

/* Polyhedron::BuildPolyhedron (Geometry/src/Polyhedron.cpp:324-365): vertex
   adjacency from the faces, volume and inertia from a fan of tetrahedra with
   the origin; Polyhedron::computeCircumscribedRadius (:201-210) */
int orc_c_add_polyhedron_shape(orc_csys* s, int nv, const double* verts, int nf,
                               const int32_t* face_size, const int32_t* face_idx) {
  if (s->nshape >= MAXSH || nv > MAXV || nv < 1) return -1;
  shape_t* sh = &s->shapes[s->nshape];
  memset(sh, 0, sizeof(*sh));
  sh->type = ORC_SHAPE_POLYHEDRON;
  sh->nv = nv;
  for (int i = 0; i < nv; ++i)
    for (int d = 0; d < 3; ++d) sh->v[i][d] = verts[3 * i + d];
  int off = 0;
  for (int f = 0; f < nf; ++f) {
    int m = face_size[f];
    const int32_t* fi = face_idx + off;
    for (int j = 0, k = m - 1; j < m; k = j++) {
      int a = fi[k];
      if (sh->cob_n[a] < MAXV) sh->cob[a][sh->cob_n[a]++] = fi[j];
    }
    if (m == 3) {
      tet_contrib(ld(sh->v[fi[0]]), ld(sh->v[fi[1]]), ld(sh->v[fi[2]]), &sh->vol, sh->inertia);
    } else {
      v3 G = mk(0., 0., 0.);
      for (int j = 0; j < m; ++j) G = add(G, ld(sh->v[fi[j]]));
      G = dvd(G, (double)m);
      for (int j = 0, k = m - 1; j < m; k = j++)
        tet_contrib(G, ld(sh->v[fi[k]]), ld(sh->v[fi[j]]), &sh->vol, sh->inertia);
    }
    off += m;
  }
  sh->first_vertex = 0;
  while (sh->first_vertex < nv && sh->cob_n[sh->first_vertex] == 0) ++sh->first_vertex;
  double ray = dot(ld(sh->v[0]), ld(sh->v[0])), d;
  for (int i = 1; i < nv; ++i)
    if ((d = dot(ld(sh->v[i]), ld(sh->v[i]))) > ray) ray = d;
  sh->circ = sqrt(ray);
  finish_shape(sh);
  return s->nshape++;
}

int orc_c_shape_info(const orc_csys* s, int shape, double* vol, double* inertia, double* circ,
                     double* bbox_ext) {
  if (shape < 0 || shape >= s->nshape) return ORC_ERR_ARG;
  const shape_t* sh = &s->shapes[shape];
  if (vol) *vol = sh->vol;
  if (inertia) memcpy(inertia, sh->inertia, sizeof(double) * 6);
  if (circ) *circ = sh->circ;
  if (bbox_ext) memcpy(bbox_ext, sh->bbox_ext, sizeof(double) * 3);
  return ORC_OK;
}

/* crust scaling (RigidBodyWithCrust.cpp:51-80): (extent - crust) / extent */
static void crust_scale(const double* ext, double crust, double* sc) {
  for (int d = 0; d < 3; ++d) sc[d] = ext[d] < ORC_EPSILON ? 1. : (ext[d] - crust) / ext[d];
}

/* Particle(RigidBodyWithCrust*, density, mat) (Component/src/Particle.cpp:179-232):
   m = rho V, I = rho I_shape, Component::getEquivalentSphereRadius
   (Component.cpp:205-208) */
int orc_c_set_particles(orc_csys* s, int n, const double* x, const double* v, const double* w,
                        const double* q, const double* basis, const int32_t* shape,
                        const double* density, const double* crust, const int32_t* material) {
  cfree_particles(s);
  s->n = n;
  size_t n3 = (size_t)n * 3;
  s->x = (double*)malloc(sizeof(double) * n3);
  s->v = (double*)malloc(sizeof(double) * n3);
  s->w = (double*)malloc(sizeof(double) * n3);
  s->wbf = (double*)malloc(sizeof(double) * n3);
  s->q = (double*)malloc(sizeof(double) * 4 * (size_t)n);
  s->basis = (double*)malloc(sizeof(double) * 9 * (size_t)n);
  s->F = (double*)calloc(n3, sizeof(double));
  s->T = (double*)calloc(n3, sizeof(double));
  s->mass = (double*)malloc(sizeof(double) * (size_t)n);
  s->inertia = (double*)malloc(sizeof(double) * n3);
  s->req = (double*)malloc(sizeof(double) * (size_t)n);
  s->crust = (double*)malloc(sizeof(double) * (size_t)n);
  s->scale = (double*)malloc(sizeof(double) * n3);
  s->shape = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  s->mat = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  s->next = (int*)malloc(sizeof(int) * (size_t)n);
  s->prev = (int*)malloc(sizeof(int) * (size_t)n);
  s->cell = (int*)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    int sh_id = shape[i];
    if (sh_id < 0 || sh_id >= s->nshape) return ORC_ERR_ARG;
    const shape_t* sh = &s->shapes[sh_id];
    double rho = density[i];
    s->shape[i] = sh_id;
    s->mat[i] = material ? material[i] : 0;
    s->mass[i] = rho * sh->vol;
    s->inertia[3 * i + 0] = sh->inertia[0] * rho;
    s->inertia[3 * i + 1] = sh->inertia[3] * rho;
    s->inertia[3 * i + 2] = sh->inertia[5] * rho;
    s->req[i] = pow((0.75 / ORC_PI) * sh->vol, 1. / 3.);
    s->crust[i] = crust[i];
    crust_scale(sh->bbox_ext, crust[i], s->scale + 3 * i);
    memcpy(s->x + 3 * i, x + 3 * i, sizeof(double) * 3);
    memcpy(s->v + 3 * i, v + 3 * i, sizeof(double) * 3);
    memcpy(s->w + 3 * i, w + 3 * i, sizeof(double) * 3);
    memcpy(s->q + 4 * i, q + 4 * i, sizeof(double) * 4);
    memcpy(s->basis + 9 * i, basis + 9 * i, sizeof(double) * 9);
    quat qq = qmk(mk(q[4 * i], q[4 * i + 1], q[4 * i + 2]), q[4 * i + 3]);
    st(s->wbf + 3 * i, qmulv(qconj(qq), qleftvec(qq, ld(w + 3 * i))));
    s->cell[i] = -1;
    s->next[i] = s->prev[i] = -1;
  }
  return ORC_OK;
}

int orc_c_add_wall(orc_csys* s, const double* basis, const double* centre,
                   const double* half, double crust, int material) {
  if (s->nwalls >= ORC_MAX_WALLS || material < 0 || material >= ORC_MAX_MATERIALS)
    return ORC_ERR_ARG;
  cwall_t* wl = &s->walls[s->nwalls++];
  memset(wl, 0, sizeof(*wl));
  memcpy(wl->basis, basis, sizeof(double) * 9);
  memcpy(wl->origin, centre, sizeof(double) * 3);
  memcpy(wl->ext, half, sizeof(double) * 3);
  wl->crust = crust;
  wl->material = material;
  double e0 = half[0], e1 = half[1], e2 = half[2], mn[3], mx[3];
  for (int i = 0; i < 3; ++i) {
    v3 row = mk(basis[3 * i], basis[3 * i + 1], basis[3 * i + 2]);
    v3 nrow = neg(row);
    v3 sm = norm(nrow) < ORC_EPSILON ? mk(0., 0., 0.)
                                      : mk(nrow.x < 0. ? -e0 : e0, nrow.y < 0. ? -e1 : e1,
                                           nrow.z < 0. ? -e2 : e2);
    v3 sp = norm(row) < ORC_EPSILON ? mk(0., 0., 0.)
                                     : mk(row.x < 0. ? -e0 : e0, row.y < 0. ? -e1 : e1,
                                          row.z < 0. ? -e2 : e2);
    mn[i] = centre[i] + dot(row, sm);
    mx[i] = centre[i] + dot(row, sp);
  }
  for (int i = 0; i < 3; ++i) {
    wl->bb_ext[i] = (mx[i] - mn[i]) / 2.;
    wl->bb_centre[i] = mn[i] + wl->bb_ext[i];
  }
  wl->req = pow((0.75 / ORC_PI) * (8.0 * e0 * e1 * e2), 1. / 3.);
  wl->circ = norm(ld(half));
  crust_scale(half, crust, wl->scale);
  return ORC_OK;
}

/* Rectangle obstacle (Rectangle.cpp:13-21): the box's bounding box with a
   zero third extent; equivalent radius of the area LX LY (getVolume,
   :182-185), circumscribed radius sqrt(LX^2 + LY^2) / 2 (:118-121) */
int orc_c_add_rectangle_wall(orc_csys* s, const double* basis, const double* centre, double lx,
                             double ly, double crust, int material) {
  if (!(lx > 0.) || !(ly > 0.)) return ORC_ERR_ARG;
  double half[3] = {lx / 2., ly / 2., 0.};
  int rc = orc_c_add_wall(s, basis, centre, half, crust, material);
  if (rc) return rc;
  cwall_t* wl = &s->walls[s->nwalls - 1];
  wl->rect = 1;
  double a = 2. * half[0], b = 2. * half[1];
  wl->req = pow((0.75 / ORC_PI) * (a * b), 1. / 3.);
  wl->circ = sqrt(a * a + b * b) / 2.;
  return ORC_OK;
}

int orc_c_set_contact_model(orc_csys* s, int a, int b, int kind, const double* p) {
  if (a < 0 || b < 0 || a >= ORC_MAX_MATERIALS || b >= ORC_MAX_MATERIALS) return ORC_ERR_ARG;
  cmodel_t m;
  memset(&m, 0, sizeof(m));
  m.kind = kind;
  memcpy(m.p, p, sizeof(double) * 5);
  m.beta = log(p[1]) / sqrt(ORC_PI * ORC_PI + log(p[1]) * log(p[1]));
  m.m2sqrt56 = -2. * sqrt(5. / 6.);
  s->models[a][b] = m;
  s->models[b][a] = m;
  return ORC_OK;
}

/* ------------------------------------------------------------- contact */
typedef struct {
  const shape_t* sh;
  const double* basis;
  double origin[3];
  const double* scale;
  double crust, circ;
  int* hist; /* support history (m_curr_vertex) of a polyhedron, or NULL */
} body_t;

/* RigidBodyWithCrust::ClosestPoint general case (RigidBodyWithCrust.cpp:274-348):
   circumscribed-sphere check, GJK between the crust-shrunk bodies, contact at
   the midpoint of the witness points, overlap vector (pA - pB)/d, rounded and
   scaled by (crustA + crustB - d).  Returns 1 contact candidate, 0 none,
   -1 ContactError (shrunk bodies overlap). */
static int closest_point(int sv, const body_t* A, const body_t* B, v3* contact, v3* ov,
                         double* dist, int64_t* ngjk, int* nbIter) {
  *nbIter = 0;
  v3 gab = sub(ld(A->origin), ld(B->origin));
  if (!(norm(gab) < A->circ + B->circ)) return 0;
  ++*ngjk;
  xf_t a2w, b2w;
  crust_xf(A->basis, A->origin, A->scale, &a2w);
  crust_xf(B->basis, B->origin, B->scale, &b2w);
  v3 pa, pb;
  int nit = 0;
  double distance = gjk_distance(sv, A->sh, B->sh, &a2w, &b2w, &pa, &pb, &nit, A->hist, B->hist);
  *nbIter = nit;
  if (distance < ORC_EPSILON) return -1;
  pa = xf_apply(&a2w, pa);
  pb = xf_apply(&b2w, pb);
  v3 ba = sub(pa, pb);
  *contact = add(dvd(pa, 2.0), dvd(pb, 2.0));
  v3 o = dvd(ba, distance);
  o = round3(o);
  *ov = scl(o, A->crust + B->crust - distance);
  distance -= A->crust + B->crust;
  *dist = distance;
  return 1;
}

/* Scale of one bounding-box axis shrunk by ten crusts, at least 0.5
   (RigidBodyWithCrust::getTransformWithCrust(10, 0.5),
   RigidBodyWithCrust.cpp:770-795). */
static double retry_scale(double extent, double crust) {
  double sc = extent < ORC_EPSILON ? 1. : (extent - 10. * crust) / extent;
  return sc < 0.5 ? 0.5 : sc;
}

/* RigidBodyWithCrust::ClosestPoint_ErreurHandling (RigidBodyWithCrust.cpp:
   534-588), called by Particle::InterAction (Particle.cpp:508-525) and
   SimpleObstacle::InterAction after ClosestPoint raised the ContactError: GJK
   between the bodies shrunk by ten crusts; contact at the midpoint of the
   witness points, overlap vector the normalised B -> A direction times the
   crust sum, overlap distance minus the crust sum.  -1: the shrunk bodies
   still overlap (the error is raised).  It calls the Johnson closest_points
   whichever narrow phase is selected (:545-546). */
static int closest_point_retry(const body_t* A, const body_t* B, v3* contact, v3* ov,
                               double* dist, int64_t* ngjk, int* nbIter) {
  double sa[3], sb[3];
  for (int d = 0; d < 3; ++d) {
    sa[d] = retry_scale(A->sh->bbox_ext[d], A->crust);
    sb[d] = retry_scale(B->sh->bbox_ext[d], B->crust);
  }
  ++*ngjk;
  xf_t a2w, b2w;
  crust_xf(A->basis, A->origin, sa, &a2w);
  crust_xf(B->basis, B->origin, sb, &b2w);
  v3 pa, pb;
  int nit = 0;
  double distance = closest_points(A->sh, B->sh, &a2w, &b2w, &pa, &pb, &nit, A->hist, B->hist);
  *nbIter = nit;
  if (distance < ORC_EPSILON) return -1;
  pa = xf_apply(&a2w, pa);
  pb = xf_apply(&b2w, pb);
  v3 ba = sub(pa, pb);
  ba = dvd(ba, norm(ba));
  *contact = add(dvd(pa, 2.0), dvd(pb, 2.0));
  double imposed = 1. * (A->crust + B->crust);
  *ov = scl(ba, imposed);
  *dist = -imposed;
  return 1;
}

/* Keep each polyhedron particle's support history (Polyhedron::m_curr_vertex,
   mutable, Polyhedron.cpp:265-300), every particle starting from its shape's
   first vertex with neighbours as after construction or a restart; the
   history then follows this oracle's order of closest-point calls, which is
   the reference's order of Particle::InterAction calls. */
int orc_c_set_support_history(orc_csys* s, int on) {
  free(s->hist);
  s->hist = NULL;
  if (!on) return ORC_OK;
  s->hist = (int*)malloc(sizeof(int) * (size_t)(s->n > 0 ? s->n : 1));
  if (!s->hist) return ORC_ERR_ARG;
  for (int i = 0; i < s->n; ++i) s->hist[i] = s->shapes[s->shape[i]].first_vertex;
  return ORC_OK;
}

/* One support call on a shape of this system: from vertex `start` (a
   polyhedron's history value; -1: the first vertex with neighbours), the
   support point of direction dir and, for a polyhedron, the vertex reached. */
int orc_c_support(const orc_csys* s, int shape, const double* dir, int start, double* point,
                  int* vertex) {
  if (shape < 0 || shape >= s->nshape) return ORC_ERR_ARG;
  const shape_t* sh = &s->shapes[shape];
  int h = start;
  if (sh->type == ORC_SHAPE_POLYHEDRON && (h < 0 || h >= sh->nv)) h = sh->first_vertex;
  st(point, support(sh, ld(dir), sh->type == ORC_SHAPE_POLYHEDRON ? &h : NULL));
  *vertex = sh->type == ORC_SHAPE_POLYHEDRON ? h : -1;
  return ORC_OK;
}

/* CollisionDetection Type="GJK" / "GJK_SV" (Base/src/Grains.cpp:660-666) */
int orc_c_set_narrow_phase(orc_csys* s, int sv) {
  if (sv != 0 && sv != 1) return ORC_ERR_ARG;
  s->gjk_sv = sv;
  return ORC_OK;
}

int orc_c_closest_points(const orc_csys* s, int shape_a, const double* basis_a,
                         const double* origin_a, const double* scale_a, int shape_b,
                         const double* basis_b, const double* origin_b, const double* scale_b,
                         double* pa, double* pb, int* nbIter, double* dist) {
  xf_t a2w, b2w;
  crust_xf(basis_a, origin_a, scale_a, &a2w);
  crust_xf(basis_b, origin_b, scale_b, &b2w);
  v3 A, B;
  *dist = gjk_distance(s->gjk_sv, &s->shapes[shape_a], &s->shapes[shape_b], &a2w, &b2w, &A, &B,
                       nbIter, NULL, NULL);
  st(pa, xf_apply(&a2w, A));
  st(pb, xf_apply(&b2w, B));
  return ORC_OK;
}

static inline v3 vel_at(v3 x, v3 v, v3 w, v3 pt) { return add(v, cross(w, sub(pt, x))); }

/* Hooke/Hertz performForcesCalculus, as in the sphere oracle */
static v3 cforce(const cmodel_t* m, v3 pt, v3 pen, double odist, v3 x0, v3 v0, v3 w0, double m0,
                 double req0, v3 x1, v3 v1, v3 w1, double m1, double req1) {
  v3 normal = dvd(pen, norm(pen));
  normal = round3(normal);
  v3 tmpV = sub(vel_at(x0, v0, w0, pt), vel_at(x1, v1, w1, pt));
  v3 v_n = scl(normal, dot(tmpV, normal));
  v3 v_t = sub(tmpV, v_n);
  v3 delFN, delFT;
  double normFN, muc;
  if (m->kind == ORC_MODEL_HOOKE) {
    double kn = m->p[0], etat_in = m->p[2];
    muc = m->p[3];
    delFN = scl(pen, kn);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double gamman = -2. * m->beta * sqrt(avmass * kn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double etat = etat_in == -1. ? -m->beta * sqrt(kn / avmass) : etat_in;
    double gammat = 2. * etat * avmass;
    delFT = scl(v_t, -gammat);
  } else {
    double Es = m->p[0], Gs = m->p[2];
    muc = m->p[3];
    double deltan = -odist;
    double Req = 1. / (1. / req0 + 1. / req1);
    double avmass = 1. / (1. / m0 + 1. / m1);
    double sq = sqrt(Req * deltan);
    double Sn = 2. * Es * sq;
    double St = 8. * Gs * sq;
    double kn = (4. / 3.) * Es * sq;
    delFN = scl(pen, kn);
    double gamman = m->m2sqrt56 * m->beta * sqrt(avmass * Sn);
    delFN = sub(delFN, scl(v_n, gamman));
    normFN = norm(delFN);
    double gammat = m->m2sqrt56 * m->beta * sqrt(avmass * St);
    delFT = scl(v_t, -gammat);
  }
  double fn = muc * normFN;
  double ft = norm(delFT);
  if (fn < ft) {
    double nvt = norm(v_t);
    v3 tangent = mk(0., 0., 0.);
    if (nvt > ORC_EPSILON) tangent = dvd(neg(v_t), nvt);
    delFT = scl(tangent, fn);
  }
  return add(delFN, delFT);
}

static void clog_contact(orc_csys* s, int a, int b, v3 pt, v3 ov, double d, v3 f) {
  if (s->log_n < s->log_cap) {
    int k = s->log_n;
    s->log_id0[k] = a;
    s->log_id1[k] = b;
    st(s->log_pt + 3 * k, pt);
    st(s->log_ov + 3 * k, ov);
    s->log_d[k] = d;
    st(s->log_f + 3 * k, f);
  }
  s->log_n++;
}

static body_t particle_body(const orc_csys* s, int p, v3 shift) {
  body_t b;
  b.sh = &s->shapes[s->shape[p]];
  b.basis = s->basis + 9 * p;
  st(b.origin, add(ld(s->x + 3 * p), shift));
  b.scale = s->scale + 3 * p;
  b.crust = s->crust[p];
  b.circ = b.sh->circ;
  b.hist = s->hist ? s->hist + p : NULL;
  return b;
}

/* ClosestPointRECTANGLE, particle first (RigidBodyWithCrust.cpp:1107-1226,
   branch :1172-1218), behind ClosestPoint's circumscribed test (:282-296):
   the particle's full shape through its transform (basis, centre) */
static int crect(const body_t* A, const cwall_t* wl, v3* point, v3* ov, double* dist) {
  v3 rp = ld(wl->origin), xa = ld(A->origin);
  if (!(norm(sub(xa, rp)) < A->circ + wl->circ)) return 0;
  v3 rn = mk(dot(mk(wl->basis[0], wl->basis[1], wl->basis[2]), mk(0., 0., 1.)),
             dot(mk(wl->basis[3], wl->basis[4], wl->basis[5]), mk(0., 0., 1.)),
             dot(mk(wl->basis[6], wl->basis[7], wl->basis[8]), mk(0., 0., 1.)));
  double sg = copysign(1., dot(rn, sub(xa, rp)));
  rn = mk(sg * rn.x, sg * rn.y, sg * rn.z);
  v3 d = neg(rn);
  const double* B = A->basis;
  v3 dl = mk(B[0] * d.x + B[3] * d.y + B[6] * d.z, B[1] * d.x + B[4] * d.y + B[7] * d.z,
             B[2] * d.x + B[5] * d.y + B[8] * d.z);
  v3 sp = support(A->sh, dl, A->hist);
  /* Transform::operator() (Transform.cpp:312-317) */
  v3 pa = mk(dot(mk(B[0], B[1], B[2]), sp) + A->origin[0], dot(mk(B[3], B[4], B[5]), sp) + A->origin[1],
             dot(mk(B[6], B[7], B[8]), sp) + A->origin[2]);
  if (!(dot(rn, sub(pa, rp)) < 0.)) return 0;
  double k = dot(sub(rp, pa), rn);
  v3 pb = add(mk(k * rn.x, k * rn.y, k * rn.z), pa);
  /* Rectangle::isIn (Rectangle.cpp:365-370) after the inverse transform
     (Transform::setToInverseTransform, Transform.cpp:71-77) */
  double io[3];
  for (int i = 0; i < 3; ++i) {
    v3 col = mk(wl->basis[i], wl->basis[3 + i], wl->basis[6 + i]);
    io[i] = dot(neg(col), rp);
  }
  v3 l = mk(dot(mk(wl->basis[0], wl->basis[3], wl->basis[6]), pb) + io[0],
            dot(mk(wl->basis[1], wl->basis[4], wl->basis[7]), pb) + io[1],
            dot(mk(wl->basis[2], wl->basis[5], wl->basis[8]), pb) + io[2]);
  if (!(fabs(l.z) < ORC_EPSILON && fabs(l.y) < wl->ext[1] && fabs(l.x) < wl->ext[0])) return 0;
  *point = add(mk(pa.x / 2.0, pa.y / 2.0, pa.z / 2.0), mk(pb.x / 2.0, pb.y / 2.0, pb.z / 2.0));
  *ov = sub(pb, pa);
  *dist = -norm(*ov);
  return 1;
}

/* Particle::InterAction (Particle.cpp:478-546) for two convex particles */
static int cinteract(orc_csys* s, int p0, int p1, v3 shift) {
  body_t A = particle_body(s, p0, mk(0., 0., 0.)), B = particle_body(s, p1, shift);
  v3 c, ov;
  double dist;
  int nit = 0;
  int rc = closest_point(s->gjk_sv, &A, &B, &c, &ov, &dist, &s->ngjk, &nit);
  if (rc < 0) rc = closest_point_retry(&A, &B, &c, &ov, &dist, &s->ngjk, &nit);
  if (rc < 0) return cset_error(s, ORC_ERR_CONTACT, p0, p1);
  if (rc == 0) return ORC_OK;
  c_contact_features(s, dist, nit);
  if (!(dist < 0.)) return ORC_OK;
  const cmodel_t* m = &s->models[s->mat[p0]][s->mat[p1]];
  if (m->kind == ORC_MODEL_NONE) return cset_error(s, ORC_ERR_ARG, p0, p1);
  v3 xa = ld(A.origin), xb = ld(B.origin);
  v3 f = cforce(m, c, ov, dist, xa, ld(s->v + 3 * p0), ld(s->w + 3 * p0), s->mass[p0], s->req[p0],
                xb, ld(s->v + 3 * p1), ld(s->w + 3 * p1), s->mass[p1], s->req[p1]);
  v3 f0 = scl(f, 1.), f1 = neg(f);
  st(s->F + 3 * p0, add(ld(s->F + 3 * p0), f0));
  st(s->T + 3 * p0, add(ld(s->T + 3 * p0), cross(sub(c, xa), f0)));
  st(s->F + 3 * p1, add(ld(s->F + 3 * p1), f1));
  st(s->T + 3 * p1, add(ld(s->T + 3 * p1), cross(sub(c, xb), f1)));
  clog_contact(s, p0, p1, c, ov, dist, f0);
  return ORC_OK;
}

/* One particle pair evaluated on the current state without accumulating:
   closest point of (p0, p1 moved by shift) and the contact force on p0.
   Returns 1 contact, 0 none, -1 ContactError. */
int orc_c_pair(orc_csys* s, int p0, int p1, const double* shift, double* point, double* ovec,
               double* overlap, double* force) {
  body_t A = particle_body(s, p0, mk(0., 0., 0.)), B = particle_body(s, p1, ld(shift));
  A.hist = B.hist = NULL; /* a diagnostic: no support history moves */
  v3 c, ov;
  double dist;
  int64_t ng = 0;
  int nit = 0;
  int rc = closest_point(s->gjk_sv, &A, &B, &c, &ov, &dist, &ng, &nit);
  if (rc < 0) rc = closest_point_retry(&A, &B, &c, &ov, &dist, &ng, &nit);
  if (rc <= 0) return rc;
  st(point, c);
  st(ovec, ov);
  *overlap = dist;
  if (!(dist < 0.)) return 0;
  const cmodel_t* m = &s->models[s->mat[p0]][s->mat[p1]];
  v3 f = cforce(m, c, ov, dist, ld(A.origin), ld(s->v + 3 * p0), ld(s->w + 3 * p0), s->mass[p0],
                s->req[p0], ld(B.origin), ld(s->v + 3 * p1), ld(s->w + 3 * p1), s->mass[p1],
                s->req[p1]);
  st(force, f);
  return 1;
}

/* SimpleObstacle::InterAction (SimpleObstacle.cpp:222-271): the particle is
   the first body of ClosestPoint, the box wall the second */
static int cinteract_wall(orc_csys* s, int p, int wi) {
  const cwall_t* wl = &s->walls[wi];
  const double* xp = s->x + 3 * p;
  const double rc_circ = s->shapes[s->shape[p]].circ;
  if (!(fabs(wl->bb_centre[0] - xp[0]) <= wl->bb_ext[0] + rc_circ &&
        fabs(wl->bb_centre[1] - xp[1]) <= wl->bb_ext[1] + rc_circ &&
        fabs(wl->bb_centre[2] - xp[2]) <= wl->bb_ext[2] + rc_circ))
    return ORC_OK;
  body_t A = particle_body(s, p, mk(0., 0., 0.));
  shape_t wsh;
  memset(&wsh, 0, sizeof(wsh));
  wsh.type = ORC_SHAPE_BOX;
  memcpy(wsh.ext, wl->ext, sizeof(wsh.ext));
  memcpy(wsh.bbox_ext, wl->ext, sizeof(wsh.bbox_ext));
  body_t B;
  B.sh = &wsh;
  B.basis = wl->basis;
  memcpy(B.origin, wl->origin, sizeof(B.origin));
  B.scale = wl->scale;
  B.crust = wl->crust;
  B.circ = wl->circ;
  B.hist = NULL;
  v3 c, ov;
  double dist;
  int nit = 0;
  int rc = wl->rect ? crect(&A, wl, &c, &ov, &dist)
                    : closest_point(s->gjk_sv, &A, &B, &c, &ov, &dist, &s->ngjk, &nit);
  if (rc < 0) rc = closest_point_retry(&A, &B, &c, &ov, &dist, &s->ngjk, &nit);
  if (rc < 0) return cset_error(s, ORC_ERR_CONTACT, p, -1 - wi);
  if (rc == 0) return ORC_OK;
  c_contact_features(s, dist, nit);
  if (!(dist < 0.)) return ORC_OK;
  const cmodel_t* m = &s->models[s->mat[p]][wl->material];
  if (m->kind == ORC_MODEL_NONE) return cset_error(s, ORC_ERR_ARG, p, -1 - wi);
  v3 zero = mk(0., 0., 0.);
  v3 xa = ld(xp);
  v3 f = cforce(m, c, ov, dist, xa, ld(s->v + 3 * p), ld(s->w + 3 * p), s->mass[p], s->req[p],
                ld(wl->origin), zero, zero, 1.e20, wl->req);
  st(s->F + 3 * p, add(ld(s->F + 3 * p), f));
  st(s->T + 3 * p, add(ld(s->T + 3 * p), cross(sub(c, xa), f)));
  clog_contact(s, p, -1 - wi, c, ov, dist, f);
  return ORC_OK;
}

/* One particle-wall contact evaluated on the current state (no InZone test). */
int orc_c_wall_pair(orc_csys* s, int p, int wi, double* point, double* ovec, double* overlap,
                    double* force) {
  const cwall_t* wl = &s->walls[wi];
  body_t A = particle_body(s, p, mk(0., 0., 0.));
  A.hist = NULL; /* a diagnostic: no support history moves */
  shape_t wsh;
  memset(&wsh, 0, sizeof(wsh));
  wsh.type = ORC_SHAPE_BOX;
  memcpy(wsh.ext, wl->ext, sizeof(wsh.ext));
  memcpy(wsh.bbox_ext, wl->ext, sizeof(wsh.bbox_ext));
  body_t B;
  B.sh = &wsh;
  B.basis = wl->basis;
  memcpy(B.origin, wl->origin, sizeof(B.origin));
  B.scale = wl->scale;
  B.crust = wl->crust;
  B.circ = wl->circ;
  B.hist = NULL;
  v3 c, ov;
  double dist;
  int64_t ng = 0;
  int nit = 0;
  int rc = wl->rect ? crect(&A, wl, &c, &ov, &dist)
                     : closest_point(s->gjk_sv, &A, &B, &c, &ov, &dist, &ng, &nit);
  if (rc < 0) rc = closest_point_retry(&A, &B, &c, &ov, &dist, &ng, &nit);
  if (rc <= 0) return rc;
  st(point, c);
  st(ovec, ov);
  *overlap = dist;
  if (!(dist < 0.)) return 0;
  const cmodel_t* m = &s->models[s->mat[p]][wl->material];
  v3 zero = mk(0., 0., 0.);
  v3 f = cforce(m, c, ov, dist, ld(s->x + 3 * p), ld(s->v + 3 * p), ld(s->w + 3 * p), s->mass[p],
                s->req[p], ld(wl->origin), zero, zero, 1.e20, wl->req);
  st(force, f);
  return 1;
}

static int cget_cell(const orc_csys* s, const double* p, int* ijk) {
  int id[3];
  for (int d = 0; d < 3; ++d) {
    id[d] = (int)floor((p[d] - s->cfg.origin[d]) / s->edge[d]);
    if (s->cfg.periodic[d]) {
      if (id[d] >= s->nc[d]) id[d] = s->nc[d] - 1;
      if (id[d] < 0) id[d] = 0;
    }
    if (id[d] < 0 || id[d] >= s->nc[d]) return -1;
  }
  if (ijk) { ijk[0] = id[0]; ijk[1] = id[1]; ijk[2] = id[2]; }
  return id[1] * s->nc[2] * s->nc[0] + id[2] * s->nc[0] + id[0];
}

static void c_push_front(orc_csys* s, int c, int p) {
  s->prev[p] = -1;
  s->next[p] = s->head[c];
  if (s->head[c] >= 0) s->prev[s->head[c]] = p;
  s->head[c] = p;
  s->cell[p] = c;
}

static void c_remove(orc_csys* s, int p) {
  int c = s->cell[p];
  if (s->prev[p] >= 0) s->next[s->prev[p]] = s->next[p];
  else s->head[c] = s->next[p];
  if (s->next[p] >= 0) s->prev[s->next[p]] = s->prev[p];
  s->next[p] = s->prev[p] = -1;
  s->cell[p] = -1;
}

static int c_link(orc_csys* s) {
  for (int c = 0; c < s->ncell; ++c) s->head[c] = -1;
  for (int p = 0; p < s->n; ++p) {
    int c = cget_cell(s, s->x + 3 * p, NULL);
    if (c < 0) return cset_error(s, ORC_ERR_SIMULATION, p, -1);
    c_push_front(s, c, p);
  }
  return ORC_OK;
}

static int c_link_update(orc_csys* s) {
  for (int p = 0; p < s->n; ++p) {
    int c = cget_cell(s, s->x + 3 * p, NULL);
    if (c < 0) return cset_error(s, ORC_ERR_SIMULATION, p, -1);
    if (c != s->cell[p]) {
      c_remove(s, p);
      c_push_front(s, c, p);
    }
  }
  return ORC_OK;
}

/* LinkedCell::ComputeForces (LinkedCell.cpp:1019-1103) */
static int c_compute_forces(orc_csys* s) {
  const double* g = s->cfg.gravity;
  s->log_n = 0;
  for (int p = 0; p < s->n; ++p) {
    double m = s->mass[p] * (1. - 0. / 1.);
    s->F[3 * p] = g[0] * m;
    s->F[3 * p + 1] = g[1] * m;
    s->F[3 * p + 2] = g[2] * m;
    s->T[3 * p] = s->T[3 * p + 1] = s->T[3 * p + 2] = 0.;
  }
  const int ni = s->nc[0];
  for (int p = 0; p < s->n; ++p) {
    int ijk[3];
    int c = cget_cell(s, s->x + 3 * p, ijk);
    if (c < 0) return cset_error(s, ORC_ERR_SIMULATION, p, -1);
    for (int q = s->next[p]; q >= 0; q = s->next[q]) {
      int rc = cinteract(s, p, q, mk(0., 0., 0.));
      if (rc) return rc;
    }
    for (int jj = 12; jj >= 0; --jj) {
      int t[3], ok = 1;
      double sh[3] = {0., 0., 0.};
      for (int d = 0; d < 3; ++d) {
        t[d] = ijk[d] + s->stencil[jj][d];
        int n = s->nc[d];
        if (t[d] < 0 || t[d] >= n) {
          if (!s->cfg.periodic[d]) { ok = 0; break; }
          if (t[d] < 0) { t[d] += n; sh[d] = -s->cfg.size[d]; }
          else { t[d] -= n; sh[d] = s->cfg.size[d]; }
        }
      }
      if (!ok) continue;
      int cn = t[1] * s->nc[2] * ni + t[2] * ni + t[0];
      for (int q = s->head[cn]; q >= 0; q = s->next[q]) {
        int rc = cinteract(s, p, q, mk(sh[0], sh[1], sh[2]));
        if (rc) return rc;
      }
    }
    for (int wi = 0; wi < s->nwalls; ++wi) {
      int rc = cinteract_wall(s, p, wi);
      if (rc) return rc;
    }
  }
  return ORC_OK;
}

/* ParticleKinematics3D::computeAngularAccelerationBodyFixed
   (Component/src/ParticleKinematics3D.cpp:56-77), predictor mode (Grains.cpp:18):
   torque / I plus the gyroscopic term */
static v3 ang_acc3(v3 tq, v3 om, const double* I) {
  return mk(tq.x / (1. * I[0]) + om.y * om.z * (I[1] - I[2]) / I[0],
            tq.y / (1. * I[1]) + om.z * om.x * (I[2] - I[0]) / I[1],
            tq.z / (1. * I[2]) + om.x * om.y * (I[0] - I[1]) / I[2]);
}

/* SecondOrderLeapFrog SSPRK3 (SecondOrderLeapFrog.cpp:56-66) */
static v3 ssprk3_3d(v3 tq, v3 wbf, const double* I, double dt) {
  v3 K1 = scl(ang_acc3(tq, wbf, I), dt);
  v3 K2 = scl(ang_acc3(tq, add(wbf, K1), I), dt);
  v3 K3 = scl(ang_acc3(tq, add(wbf, scl(add(K1, K2), 0.25)), I), dt);
  return add(wbf, dvd(add(add(K1, K2), scl(K3, 4.)), 6.));
}

/* ParticleKinematics::Move (ParticleKinematics.cpp:62-104) */
static double cmove_one(orc_csys* s, int p, double dtv, double dtd) {
  quat q = qmk(mk(s->q[4 * p], s->q[4 * p + 1], s->q[4 * p + 2]), s->q[4 * p + 3]);
  quat pc = qconj(q);
  v3 tq = qmulv(pc, qleftvec(q, ld(s->T + 3 * p)));
  double mcf = s->mass[p] * 1.;
  v3 v = add(ld(s->v + 3 * p), scl(dvd(ld(s->F + 3 * p), mcf), dtv));
  v3 dx = scl(v, dtd);
  v3 wbf = ssprk3_3d(tq, ld(s->wbf + 3 * p), s->inertia + 3 * p, dtv);
  v3 w = qmulv(q, qleftvec(pc, wbf));
  v3 avg = qmulv(q, qleftvec(pc, wbf));
  st(s->v + 3 * p, v);
  st(s->wbf + 3 * p, wbf);
  st(s->w + 3 * p, w);
  st(s->x + 3 * p, add(ld(s->x + 3 * p), dx));
  double nOmega = norm(avg);
  quat dq;
  if (nOmega > ORC_EPSILON) {
    double c = cos(nOmega * dtd / 2.);
    double sn = sin(nOmega * dtd / 2.);
    dq = qmk(scl(avg, sn * 1. / nOmega), c);
  } else {
    dq = qmk(mk(0., 0., 0.), 1.);
  }
  quat qn = qmul(dq, q);
  s->q[4 * p] = qn.v.x; s->q[4 * p + 1] = qn.v.y; s->q[4 * p + 2] = qn.v.z; s->q[4 * p + 3] = qn.w;
  compose_left_rotation(s->basis + 9 * p, dq);
  return norm(dx);
}

static int c_move(orc_csys* s, double dtv, double dtd) {
  for (int p = 0; p < s->n; ++p) {
    double depl = cmove_one(s, p, dtv, dtd);
    if (depl > s->crust[p]) return cset_error(s, ORC_ERR_MOTION, p, -1);
    for (int d = 0; d < 3; ++d) {
      if (!s->cfg.periodic[d]) continue;
      double lo = s->cfg.origin[d], L = s->cfg.size[d];
      double* xd = s->x + 3 * p + d;
      if (*xd >= lo + L) *xd -= L;
      else if (*xd < lo) *xd += L;
    }
  }
  return ORC_OK;
}

static int c_advance(orc_csys* s, double dtv) {
  for (int p = 0; p < s->n; ++p) {
    quat q = qmk(mk(s->q[4 * p], s->q[4 * p + 1], s->q[4 * p + 2]), s->q[4 * p + 3]);
    quat pc = qconj(q);
    v3 tq = qmulv(pc, qleftvec(q, ld(s->T + 3 * p)));
    double mcf = s->mass[p] * 1.;
    v3 v = add(ld(s->v + 3 * p), scl(dvd(ld(s->F + 3 * p), mcf), dtv));
    v3 wbf = ssprk3_3d(tq, ld(s->wbf + 3 * p), s->inertia + 3 * p, dtv);
    st(s->v + 3 * p, v);
    st(s->wbf + 3 * p, wbf);
    st(s->w + 3 * p, qmulv(q, qleftvec(pc, wbf)));
  }
  return ORC_OK;
}

int orc_c_prime(orc_csys* s) {
  s->err_code = ORC_OK;
  int rc = c_link(s);
  if (rc) return rc;
  return c_compute_forces(s);
}

int orc_c_step(orc_csys* s, int nsteps, double dt) {
  for (int k = 0; k < nsteps; ++k) {
    s->time += dt;
    int rc = c_move(s, 0.5 * dt, dt);
    if (rc) return rc;
    rc = c_link_update(s);
    if (rc) return rc;
    rc = c_compute_forces(s);
    if (rc) return rc;
    rc = c_advance(s, 0.5 * dt);
    if (rc) return rc;
  }
  return ORC_OK;
}

int orc_c_get_state(const orc_csys* s, double* x, double* v, double* w, double* q,
                    double* basis, double* F, double* T) {
  size_t n3 = sizeof(double) * 3 * (size_t)s->n;
  if (x) memcpy(x, s->x, n3);
  if (v) memcpy(v, s->v, n3);
  if (w) memcpy(w, s->w, n3);
  if (q) memcpy(q, s->q, sizeof(double) * 4 * (size_t)s->n);
  if (basis) memcpy(basis, s->basis, sizeof(double) * 9 * (size_t)s->n);
  if (F) memcpy(F, s->F, n3);
  if (T) memcpy(T, s->T, n3);
  return ORC_OK;
}

void orc_c_enable_contact_log(orc_csys* s, int cap) {
  free(s->log_id0); free(s->log_id1); free(s->log_pt); free(s->log_ov); free(s->log_d);
  free(s->log_f);
  int c = cap > 0 ? cap : 1;
  s->log_cap = cap;
  s->log_n = 0;
  s->log_id0 = (int32_t*)malloc(sizeof(int32_t) * (size_t)c);
  s->log_id1 = (int32_t*)malloc(sizeof(int32_t) * (size_t)c);
  s->log_pt = (double*)malloc(sizeof(double) * 3 * (size_t)c);
  s->log_ov = (double*)malloc(sizeof(double) * 3 * (size_t)c);
  s->log_d = (double*)malloc(sizeof(double) * (size_t)c);
  s->log_f = (double*)malloc(sizeof(double) * 3 * (size_t)c);
}

int orc_c_get_contacts(const orc_csys* s, int cap, int* n_out, int32_t* id0, int32_t* id1,
                       double* point, double* ovec, double* overlap, double* force) {
  *n_out = s->log_n;
  int n = s->log_n < s->log_cap ? s->log_n : s->log_cap;
  if (n > cap) n = cap;
  for (int k = 0; k < n; ++k) {
    if (id0) id0[k] = s->log_id0[k];
    if (id1) id1[k] = s->log_id1[k];
    if (point) memcpy(point + 3 * k, s->log_pt + 3 * k, sizeof(double) * 3);
    if (ovec) memcpy(ovec + 3 * k, s->log_ov + 3 * k, sizeof(double) * 3);
    if (overlap) overlap[k] = s->log_d[k];
    if (force) memcpy(force + 3 * k, s->log_f + 3 * k, sizeof(double) * 3);
  }
  return ORC_OK;
}

int orc_c_last_error(const orc_csys* s, int* code, int32_t* id0, int32_t* id1) {
  if (code) *code = s->err_code;
  if (id0) *id0 = s->err_id0;
  if (id1) *id1 = s->err_id1;
  return s->err_code;
}

/* AppCollision::getOverlapMax / getOverlapMean / getTimeOverlapMax /
   getNbIterGJKMean (Base/src/AppCollision.cpp:205-240): out[0..3], and the
   numbers of overlaps and of GJK reports counted, out[4..5]. */
void orc_c_get_contact_stats(const orc_csys* s, double* out) {
  out[0] = s->ov_max;
  out[1] = s->ov_mean;
  out[2] = s->ov_time_max;
  out[3] = s->gjk_mean;
  out[4] = s->ov_count;
  out[5] = s->gjk_count;
}

int64_t orc_c_gjk_calls(const orc_csys* s) { return s->ngjk; }
