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

#define GDW_PI 3.14159265358979323846

static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* uniform in [0, 1) with 53 random bits */
static double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

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
   Returns 0, or -1 if max_attempts is exhausted. */
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
  uint64_t st = seed * 0xD1B54A32D192ED03ULL + 3;
  long attempts = 0;
  for (int p = 0; p < n; ++p) {
    for (;;) {
      if (++attempts > max_attempts) { free(head); free(next); return -1; }
      double c[3];
      int ic[3];
      for (int d = 0; d < 3; ++d) {
        c[d] = L * u01(&st);
        ic[d] = (int)(c[d] / h);
        if (ic[d] >= nc) ic[d] = nc - 1;
      }
      int ok = 1;
      for (int dk = -1; dk <= 1 && ok; ++dk)
        for (int dj = -1; dj <= 1 && ok; ++dj)
          for (int di = -1; di <= 1 && ok; ++di) {
            int a = (ic[0] + di + nc) % nc, b = (ic[1] + dj + nc) % nc, e = (ic[2] + dk + nc) % nc;
            for (int q = head[((long)e * nc + b) * nc + a]; q >= 0; q = next[q]) {
              double d2 = 0.0;
              for (int d = 0; d < 3; ++d) {
                double dd = x[3 * q + d] - c[d];
                if (dd > 0.5 * L) dd -= L;
                if (dd < -0.5 * L) dd += L;
                d2 += dd * dd;
              }
              double s = r[p] + r[q];
              if (d2 <= s * s * (1.0 + 1e-9)) { ok = 0; break; }
            }
          }
      if (!ok) continue;
      for (int d = 0; d < 3; ++d) x[3 * p + d] = c[d];
      long cc = ((long)ic[2] * nc + ic[1]) * nc + ic[0];
      next[p] = head[cc];
      head[cc] = p;
      break;
    }
  }
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
