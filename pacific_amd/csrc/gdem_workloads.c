/*
 * gdem_workloads.c -- seeded synthetic initial states for the DEM workloads.
 *
 * The reference draws insertion positions with libc random()
 * (Base/src/Grains.cpp:2047), which is not portable; these generators use
 * SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014) with Box-Muller normals, so every run and every
 * consumer (HIP path, CPU oracle, tests) sees bit-identical inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <omp.h>

#define GDW_PI 3.14159265358979323846

#define GDW_GAMMA 0x9E3779B97F4A7C15ULL

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static uint64_t splitmix64(uint64_t* s) { return mix64(*s += GDW_GAMMA); }

/* uniform in [0, 1) with 53 random bits */
static double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* The k-th draw (k >= 1) of u01 on a stream that starts at state s0, without
   stepping the stream: SplitMix64's state advances by a constant, so every
   draw is addressable and random sequential addition can test attempts in
   parallel and keep the first that succeeds (the result equals the serial
   addition's). */
static double u01_at(uint64_t s0, uint64_t k) {
  return (double)(mix64(s0 + k * GDW_GAMMA) >> 11) * (1.0 / 9007199254740992.0);
}

/* Attempts tested per parallel round of the additions: about twice the
   attempts the previous particle needed, between lo and hi. */
static long next_batch(long used, long lo, long hi) {
  long b = 2 * used;
  return b < lo ? lo : (b > hi ? hi : b);
}

static double normal01(uint64_t* s) {
  double u1 = u01(s), u2 = u01(s);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * GDW_PI * u2);
}

/* Jittered cubic lattice inside [lo, hi]: site (i,j,k) centre is
   lo + spacing/2 + spacing*(i,j,k), i fastest; the first n sites are used.
   Velocities are N(0, sigma_v^2) per component.  Returns the number of sites
   or -1 if fewer than n fit. */
int gdw_lattice(int n, double spacing, const double* lo, const double* hi, double jitter,
                double sigma_v, uint64_t seed, double* x, double* v) {
  int ns[3];
  for (int d = 0; d < 3; ++d) ns[d] = (int)floor((hi[d] - lo[d]) / spacing);
  long cap = (long)ns[0] * ns[1] * ns[2];
  if (cap < n) return -1;
  uint64_t st = seed * 0x2545F4914F6CDD1DULL + 1;
  for (int p = 0; p < n; ++p) {
    int i = p % ns[0], j = (p / ns[0]) % ns[1], k = p / (ns[0] * ns[1]);
    int id[3] = {i, j, k};
    for (int d = 0; d < 3; ++d)
      x[3 * p + d] = lo[d] + 0.5 * spacing + spacing * id[d] + jitter * (2.0 * u01(&st) - 1.0);
  }
  for (int p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) v[3 * p + d] = sigma_v * normal01(&st);
  return (int)cap;
}

/* Radii U[rmin, rmax]; returns the cube edge L that gives solid fraction phi. */
double gdw_radii(int n, double rmin, double rmax, double phi, uint64_t seed, double* r) {
  uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 7;
  double vs = 0.0;
  for (int p = 0; p < n; ++p) {
    r[p] = rmin + (rmax - rmin) * u01(&st);
    vs += 4.0 * GDW_PI * r[p] * r[p] * r[p] / 3.0;
  }
  return cbrt(vs / phi);
}

/* Random sequential addition of n spheres (radii given) in the triply periodic
   cube [0, L)^3 with minimum-image overlap rejection; velocities N(0, sigma_v^2).
   Attempt a (counted from 0 over the whole run) draws the centre from draws
   3a+1..3a+3 of the stream; attempts are tested in parallel rounds and the
   first success is kept, so the result does not depend on the thread count.
   Returns 0, or -1 if max_attempts is exhausted. */
typedef struct {
  const double* r;
  const double* x;
  const int* head;
  const int* next;
  double L, h;
  int nc;
  uint64_t s0;
} rsa_sph;

static int sphere_fits(const rsa_sph* R, int p, long a, double* c, long* cc) {
  int ic[3];
  for (int d = 0; d < 3; ++d) {
    c[d] = R->L * u01_at(R->s0, (uint64_t)(3 * a + d + 1));
    ic[d] = (int)(c[d] / R->h);
    if (ic[d] >= R->nc) ic[d] = R->nc - 1;
  }
  const int nc = R->nc;
  const double L = R->L;
  for (int dk = -1; dk <= 1; ++dk)
    for (int dj = -1; dj <= 1; ++dj)
      for (int di = -1; di <= 1; ++di) {
        int ia = (ic[0] + di + nc) % nc, ib = (ic[1] + dj + nc) % nc, ie = (ic[2] + dk + nc) % nc;
        for (int q = R->head[((long)ie * nc + ib) * nc + ia]; q >= 0; q = R->next[q]) {
          double d2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            double dd = R->x[3 * q + d] - c[d];
            if (dd > 0.5 * L) dd -= L;
            if (dd < -0.5 * L) dd += L;
            d2 += dd * dd;
          }
          double sr = R->r[p] + R->r[q];
          if (d2 <= sr * sr * (1.0 + 1e-9)) return 0;
        }
      }
  *cc = ((long)ic[2] * nc + ic[1]) * nc + ic[0];
  return 1;
}

int gdw_rsa_periodic(int n, const double* r, double L, double sigma_v, uint64_t seed,
                     long max_attempts, double* x, double* v) {
  double rmax = 0.0;
  for (int p = 0; p < n; ++p) rmax = r[p] > rmax ? r[p] : rmax;
  int nc = (int)floor(L / (2.0 * rmax));
  if (nc < 3) nc = 3;
  double h = L / nc;
  long ncell = (long)nc * nc * nc;
  int* head = (int*)malloc(sizeof(int) * ncell);
  int* next = (int*)malloc(sizeof(int) * (size_t)n);
  if (!head || !next) { free(head); free(next); return -1; }
  for (long c = 0; c < ncell; ++c) head[c] = -1;
  rsa_sph R = {r, x, head, next, L, h, nc, seed * 0xD1B54A32D192ED03ULL + 3};
  const long nth = omp_get_max_threads();
  long att = 0;     /* attempts used so far */
  long batch = 1;
  for (int p = 0; p < n; ++p) {
    const long first = att;
    long found = -1;
    double c[3];
    long cc = 0;
    while (found < 0) {
      if (att >= max_attempts) { free(head); free(next); return -1; }
      const long hi = att + batch < max_attempts ? att + batch : max_attempts;
      if (hi - att < 2 * nth) {
        for (long a = att; a < hi; ++a)
          if (sphere_fits(&R, p, a, c, &cc)) { found = a; break; }
      } else {
        long best = LONG_MAX;
#pragma omp parallel for schedule(static) reduction(min : best)
        for (long a = att; a < hi; ++a) {
          double ct[3];
          long ct_cell;
          if (a < best && sphere_fits(&R, p, a, ct, &ct_cell)) best = a;
        }
        if (best != LONG_MAX) {
          found = best;
          sphere_fits(&R, p, found, c, &cc);
        }
      }
      if (found < 0) {
        att = hi;
        batch = next_batch(2 * batch, 1, 256 * nth);
      }
    }
    att = found + 1;
    batch = next_batch(att - first, 1, 256 * nth);
    for (int d = 0; d < 3; ++d) x[3 * p + d] = c[d];
    next[p] = head[cc];
    head[cc] = p;
  }
  uint64_t st = R.s0 + (uint64_t)(3 * att) * GDW_GAMMA;
  for (int p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) v[3 * p + d] = sigma_v * normal01(&st);
  free(head);
  free(next);
  return 0;
}

/* Uniform random unit quaternions (x, y, z, w), Shoemake's method
   (Graphics Gems III, "Uniform random rotations", 1992). */
void gdw_random_quaternions(int n, uint64_t seed, double* q) {
  uint64_t st = seed * 0xA0761D6478BD642FULL + 11;
  for (int p = 0; p < n; ++p) {
    double u1 = u01(&st), u2 = u01(&st), u3 = u01(&st);
    double a = sqrt(1.0 - u1), b = sqrt(u1);
    q[4 * p + 0] = a * sin(2.0 * GDW_PI * u2);
    q[4 * p + 1] = a * cos(2.0 * GDW_PI * u2);
    q[4 * p + 2] = b * sin(2.0 * GDW_PI * u3);
    q[4 * p + 3] = b * cos(2.0 * GDW_PI * u3);
  }
}

/* Per-system draws for batched sweeps: uniform and log-uniform in [a, b]. */
void gdw_uniform(int n, double a, double b, int log_scale, uint64_t seed, double* out) {
  uint64_t st = seed * 0xE7037ED1A0B428DBULL + 5;
  for (int p = 0; p < n; ++p) {
    double u = u01(&st);
    out[p] = log_scale ? exp(log(a) + (log(b) - log(a)) * u) : a + (b - a) * u;
  }
}

/* ---------------------------------------------------------------- convex */
/* A convex polyhedron for the separating-axis test: vertices, face normals and
   edge directions in its local frame. */
typedef struct {
  int nv, nf, ne;
  double v[8][3], f[6][3], e[6][3];
  double circ;
  double insc;  /* radius of the largest ball about the origin inside the body */
} gdw_poly;

static void rot_apply(const double* R, const double* a, double* out) {
  for (int i = 0; i < 3; ++i) out[i] = R[3 * i] * a[0] + R[3 * i + 1] * a[1] + R[3 * i + 2] * a[2];
}

static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* Separated along some axis by more than margin (margin < 0 tolerates an
   overlap of |margin|). */
static int sat_separated(const gdw_poly* A, const double* RA, const double* cA, const gdw_poly* B,
                         const double* RB, const double* cB, double margin) {
  double va[8][3], vb[8][3], axes[6 + 6 + 36][3];
  int na = 0;
  for (int i = 0; i < A->nv; ++i) {
    rot_apply(RA, A->v[i], va[i]);
    for (int d = 0; d < 3; ++d) va[i][d] += cA[d];
  }
  for (int i = 0; i < B->nv; ++i) {
    rot_apply(RB, B->v[i], vb[i]);
    for (int d = 0; d < 3; ++d) vb[i][d] += cB[d];
  }
  for (int i = 0; i < A->nf; ++i) rot_apply(RA, A->f[i], axes[na++]);
  for (int i = 0; i < B->nf; ++i) rot_apply(RB, B->f[i], axes[na++]);
  double ea[6][3], eb[6][3];
  for (int i = 0; i < A->ne; ++i) rot_apply(RA, A->e[i], ea[i]);
  for (int i = 0; i < B->ne; ++i) rot_apply(RB, B->e[i], eb[i]);
  for (int i = 0; i < A->ne; ++i)
    for (int j = 0; j < B->ne; ++j) cross3(ea[i], eb[j], axes[na++]);
  for (int k = 0; k < na; ++k) {
    double* n = axes[k];
    double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l < 1e-12) continue;
    double mna = 1e300, mxa = -1e300, mnb = 1e300, mxb = -1e300;
    for (int i = 0; i < A->nv; ++i) {
      double p = (va[i][0] * n[0] + va[i][1] * n[1] + va[i][2] * n[2]) / l;
      mna = p < mna ? p : mna;
      mxa = p > mxa ? p : mxa;
    }
    for (int i = 0; i < B->nv; ++i) {
      double p = (vb[i][0] * n[0] + vb[i][1] * n[1] + vb[i][2] * n[2]) / l;
      mnb = p < mnb ? p : mnb;
      mxb = p > mxb ? p : mxb;
    }
    if (mxa + margin < mnb || mxb + margin < mna) return 1;
  }
  return 0;
}

/* Build a SAT description: kind 0 = box (half extents h), kind 1 = polyhedron
   with nv vertices and triangular faces (tri, 3 per face). */
static void poly_box(gdw_poly* P, const double* h) {
  memset(P, 0, sizeof(*P));
  P->nv = 8;
  for (int k = 0; k < 8; ++k) {
    P->v[k][0] = (k & 1) ? h[0] : -h[0];
    P->v[k][1] = (k & 2) ? h[1] : -h[1];
    P->v[k][2] = (k & 4) ? h[2] : -h[2];
  }
  P->nf = 3;
  P->ne = 3;
  for (int d = 0; d < 3; ++d) {
    P->f[d][d] = 1.0;
    P->e[d][d] = 1.0;
  }
  P->circ = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  P->insc = h[0] < h[1] ? (h[0] < h[2] ? h[0] : h[2]) : (h[1] < h[2] ? h[1] : h[2]);
}

static void poly_tri(gdw_poly* P, int nv, const double* verts, int nf, const int* tri) {
  memset(P, 0, sizeof(*P));
  P->nv = nv;
  double r2 = 0.0;
  for (int i = 0; i < nv; ++i) {
    for (int d = 0; d < 3; ++d) P->v[i][d] = verts[3 * i + d];
    double q = verts[3 * i] * verts[3 * i] + verts[3 * i + 1] * verts[3 * i + 1] +
               verts[3 * i + 2] * verts[3 * i + 2];
    r2 = q > r2 ? q : r2;
  }
  P->circ = sqrt(r2);
  P->nf = nf;
  for (int f = 0; f < nf; ++f) {
    const double* a = P->v[tri[3 * f]];
    const double* b = P->v[tri[3 * f + 1]];
    const double* c = P->v[tri[3 * f + 2]];
    double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    cross3(u, w, P->f[f]);
  }
  /* distance from the origin (inside) to the nearest face plane */
  P->insc = 1e300;
  for (int f = 0; f < nf; ++f) {
    const double* n = P->f[f];
    const double* a = P->v[tri[3 * f]];
    double d = fabs(n[0] * a[0] + n[1] * a[1] + n[2] * a[2]) / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    P->insc = d < P->insc ? d : P->insc;
  }
  /* unique edges of the triangles */
  int ne = 0, ea[12], eb[12];
  for (int f = 0; f < nf; ++f)
    for (int k = 0; k < 3; ++k) {
      int a = tri[3 * f + k], b = tri[3 * f + (k + 1) % 3];
      int lo = a < b ? a : b, hi = a < b ? b : a, dup = 0;
      for (int m = 0; m < ne; ++m) dup |= (ea[m] == lo && eb[m] == hi);
      if (!dup && ne < 6) {
        ea[ne] = lo;
        eb[ne] = hi;
        for (int d = 0; d < 3; ++d) P->e[ne][d] = P->v[hi][d] - P->v[lo][d];
        ++ne;
      }
    }
  P->ne = ne;
}

static void quat_to_basis(const double* q, double* b) {
  double x = q[0], y = q[1], z = q[2], w = q[3];
  b[0] = 1. - 2. * (y * y + z * z); b[1] = 2. * (x * y - z * w); b[2] = 2. * (x * z + y * w);
  b[3] = 2. * (x * y + z * w); b[4] = 1. - 2. * (x * x + z * z); b[5] = 2. * (y * z - x * w);
  b[6] = 2. * (x * z - y * w); b[7] = 2. * (y * z + x * w); b[8] = 1. - 2. * (x * x + y * y);
}

/* Random sequential addition of n convex particles in the box [lo, hi]
   (walls, no periodicity): type[i] = 0 box (half extents box_h), 1
   tetrahedron (tet_v, 4 triangular faces tet_f).  Orientation q (Shoemake,
   generated here), basis = R(q).  margin > 0 keeps a gap, margin < 0 allows an
   overlap of |margin| along the least-overlap axis.  Attempts are numbered
   over the whole run and drawn, tested in parallel and kept as in
   gdw_rsa_periodic.  Returns 0, or -1 - p when the attempts ran out while
   placing particle p. */
typedef struct {
  const gdw_poly* P;
  const int32_t* type;
  const double* x;
  const double* basis;
  const int* head;
  const int* next;
  const double* lo;
  const double* hi;
  const int* nc;
  double margin;
  uint64_t s0;
} rsa_cvx;

static int convex_fits(const rsa_cvx* R, int p, long a, double* c, long* cc) {
  const gdw_poly* A = &R->P[R->type[p] ? 1 : 0];
  const int* nc = R->nc;
  int ic[3];
  for (int d = 0; d < 3; ++d) {
    double lo = R->lo[d] + A->circ, hi = R->hi[d] - A->circ;
    c[d] = lo + (hi - lo) * u01_at(R->s0, (uint64_t)(3 * a + d + 1));
    ic[d] = (int)((c[d] - R->lo[d]) / ((R->hi[d] - R->lo[d]) / nc[d]));
    if (ic[d] >= nc[d]) ic[d] = nc[d] - 1;
  }
  const double margin = R->margin;
  for (int dk = -1; dk <= 1; ++dk)
    for (int dj = -1; dj <= 1; ++dj)
      for (int di = -1; di <= 1; ++di) {
        int ia = ic[0] + di, ib = ic[1] + dj, ie = ic[2] + dk;
        if (ia < 0 || ib < 0 || ie < 0 || ia >= nc[0] || ib >= nc[1] || ie >= nc[2]) continue;
        for (int q = R->head[((long)ie * nc[1] + ib) * nc[0] + ia]; q >= 0; q = R->next[q]) {
          const gdw_poly* B = &R->P[R->type[q] ? 1 : 0];
          double d2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            double dd = R->x[3 * q + d] - c[d];
            d2 += dd * dd;
          }
          double sr = A->circ + B->circ + (margin > 0 ? margin : 0.0);
          if (d2 > sr * sr) continue;
          /* the inscribed balls overlap by more than the tolerated overlap
             (or lie closer than the required gap): the separating-axis test
             would reject as well (its overlap along every axis is at least
             the balls' overlap); a relative 1e-9 keeps off its rounding */
          double si = (A->insc + B->insc + margin) * (1.0 - 1e-9);
          if (si > 0.0 && d2 < si * si) return 0;
          if (!sat_separated(A, R->basis + 9 * p, c, B, R->basis + 9 * q, R->x + 3 * q, margin))
            return 0;
        }
      }
  *cc = ((long)ic[2] * nc[1] + ic[1]) * nc[0] + ic[0];
  return 1;
}

int gdw_rsa_convex(int n, const int32_t* type, const double* box_h, const double* tet_v,
                   const int32_t* tet_f, const double* lo, const double* hi, double margin,
                   double sigma_v, uint64_t seed, long max_attempts, double* x, double* v,
                   double* q, double* basis) {
  gdw_poly P[2];
  poly_box(&P[0], box_h);
  int tri[12];
  for (int k = 0; k < 12; ++k) tri[k] = tet_f[k];
  poly_tri(&P[1], 4, tet_v, 4, tri);
  double rmax = P[0].circ > P[1].circ ? P[0].circ : P[1].circ;
  double h = 2.0 * rmax + (margin > 0 ? margin : 0.0);
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    nc[d] = (int)floor((hi[d] - lo[d]) / h);
    if (nc[d] < 1) nc[d] = 1;
  }
  long ncell = (long)nc[0] * nc[1] * nc[2];
  int* head = (int*)malloc(sizeof(int) * ncell);
  int* next = (int*)malloc(sizeof(int) * (size_t)n);
  if (!head || !next) { free(head); free(next); return -1; }
  for (long c = 0; c < ncell; ++c) head[c] = -1;
  gdw_random_quaternions(n, seed * 31 + 17, q);
  for (int i = 0; i < n; ++i) quat_to_basis(q + 4 * i, basis + 9 * i);
  rsa_cvx R = {P, type, x, basis, head, next, lo, hi, nc, margin,
               seed * 0xD1B54A32D192ED03ULL + 29};
  const long nth = omp_get_max_threads();
  long att = 0;
  long batch = 1;
  for (int p = 0; p < n; ++p) {
    const long first = att;
    long found = -1;
    double c[3];
    long cc = 0;
    while (found < 0) {
      if (att >= max_attempts) { free(head); free(next); return -1 - p; }
      const long top = att + batch < max_attempts ? att + batch : max_attempts;
      if (top - att < 2 * nth) {
        for (long a = att; a < top; ++a)
          if (convex_fits(&R, p, a, c, &cc)) { found = a; break; }
      } else {
        long best = LONG_MAX;
#pragma omp parallel for schedule(static) reduction(min : best)
        for (long a = att; a < top; ++a) {
          double ct[3];
          long ct_cell;
          if (a < best && convex_fits(&R, p, a, ct, &ct_cell)) best = a;
        }
        if (best != LONG_MAX) {
          found = best;
          convex_fits(&R, p, found, c, &cc);
        }
      }
      if (found < 0) {
        att = top;
        batch = next_batch(2 * batch, 1, 256 * nth);
      }
    }
    att = found + 1;
    batch = next_batch(att - first, 1, 256 * nth);
    for (int d = 0; d < 3; ++d) x[3 * p + d] = c[d];
    next[p] = head[cc];
    head[cc] = p;
  }
  uint64_t st = R.s0 + (uint64_t)(3 * att) * GDW_GAMMA;
  for (int p = 0; p < n; ++p)
    for (int d = 0; d < 3; ++d) v[3 * p + d] = sigma_v * normal01(&st);
  free(head);
  free(next);
  return 0;
}
