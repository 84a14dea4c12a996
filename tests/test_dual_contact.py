"""The contact code of the differentiated step (pacific_amd/csrc/gdem_dual.h),
built for the host with hipcc (no GPU needed):

* with S = double it returns bit for bit what the forward's sphere_sphere,
  sphere_wall and contact_force (gdem_device.h) return, on random contacts
  (face, edge and corner regions of a rotated wall box included);
* with S = Dual the derivative of every output along every input direction
  matches a central difference of the double version (relative 1e-5, the
  truncation error of the difference in the sliding branch),
  including the Coulomb-sliding branch.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pacific_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

DRIVER = r"""
#include <cstdio>
#include <cmath>
#include <random>
#include "gdem_dual.h"
using namespace gdem;

static std::mt19937_64 rng(7);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static v3 R3(double s) { return mk(U(-s, s), U(-s, s), U(-s, s)); }

static Model hooke_model(double kn, double en, double mu) {
  Model m{};
  m.kind = 1;
  m.p[0] = kn; m.p[1] = en; m.p[2] = -1.; m.p[3] = mu; m.p[4] = 0.;
  m.beta = log(en) / sqrt(kPi * kPi + log(en) * log(en));
  m.m2sqrt56 = -2. * sqrt(5. / 6.);
  return m;
}

static Wall make_box_wall(double ang) {
  Wall wl{};
  const double c = cos(ang), s = sin(ang);
  const double b[9] = {c, -s, 0., s, c, 0., 0., 0., 1.};
  for (int i = 0; i < 9; ++i) wl.basis[i] = b[i];
  wl.origin[0] = 0.01; wl.origin[1] = -0.02; wl.origin[2] = 0.005;
  wl.ext[0] = 0.004; wl.ext[1] = 0.003; wl.ext[2] = 0.0005;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) wl.inv_basis[3 * i + j] = b[3 * j + i];
  for (int i = 0; i < 3; ++i) {
    v3 row = mk(wl.inv_basis[3 * i], wl.inv_basis[3 * i + 1], wl.inv_basis[3 * i + 2]);
    wl.inv_origin[i] = dot(neg(row), mk(wl.origin[0], wl.origin[1], wl.origin[2]));
  }
  wl.crust = 5e-5;
  wl.req = 2e-3;
  return wl;
}

static bool same(v3 a, V3<double> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

int main() {
  int forward_checked = 0, forward_bad = 0, deriv_checked = 0, deriv_bad = 0, sliding = 0;
  int wall_regions[3] = {0, 0, 0};
  const double r = 1e-3, cr = 5e-5, m0 = 1.05e-5, m1 = 0.9e-5;
  double worst = 0.;
  for (int trial = 0; trial < 4000; ++trial) {
    const double kn = exp(U(log(1e4), log(1e5))), en = U(0.5, 0.95), mu = U(0.05, 0.8);
    const Model md = hooke_model(kn, en, mu);
    HookeParams<double> hp{kn, md.beta, mu};
    const bool wall = trial % 2 == 1;
    v3 xA, vA = R3(0.3), wA = R3(200.), xB, vB = R3(0.3), wB = R3(200.);
    Wall wl = make_box_wall(U(0., 1.5));
    if (!wall) {
      xA = R3(0.01);
      v3 d = R3(1.);
      const double nd = norm(d);
      xB = add(xA, scl(d, (2. * r - U(1e-7, 9e-5)) / nd));
    } else {
      // a point near the box surface: face, edge or corner region
      const int region = trial / 2 % 3;
      v3 local = mk(U(-0.9, 0.9) * wl.ext[0], U(-0.9, 0.9) * wl.ext[1], wl.ext[2] + r - U(1e-7, 9e-5));
      if (region >= 1) local.x = wl.ext[0] + (r - U(1e-7, 9e-5)) / sqrt(2.);
      if (region >= 1) local.z = wl.ext[2] + (r - U(1e-7, 9e-5)) / sqrt(2.);
      if (region == 2) {
        local = mk(wl.ext[0], wl.ext[1], wl.ext[2]);
        local = add(local, scl(mk(1., 1., 1.), (r - U(1e-7, 9e-5)) / sqrt(3.)));
      }
      xA = xform(wl.basis, wl.origin, local);
      wall_regions[region]++;
    }
    // forward: templated double == forward functions, bit for bit
    v3 c, ov;
    double dist;
    V3<double> ct, ovt;
    double distt;
    int rc, rct;
    if (!wall) {
      rc = sphere_sphere(xA, r, cr, xB, r, cr, c, ov, dist);
      rct = ss_contact(lift3<double>(xA), r, cr, lift3<double>(xB), r, cr, ct, ovt, distt);
    } else {
      rc = sphere_wall(wl, xA, r, cr, c, ov, dist);
      rct = sw_contact(wl, lift3<double>(xA), r, cr, ct, ovt, distt);
    }
    if (rc != rct) { ++forward_bad; continue; }
    if (rc != 1 || !(dist < 0.)) continue;
    ++forward_checked;
    if (!same(c, ct) || !same(ov, ovt) || dist != distt) ++forward_bad;
    v3 f;
    V3<double> FA, TA, FB, TB;
    if (!wall) {
      f = contact_force(md, c, ov, dist, xA, vA, wA, m0, 1e-3, xB, vB, wB, m1, 1e-3);
      pair_torsors(lift3<double>(xA), lift3<double>(vA), lift3<double>(wA), r, cr, m0,
                   lift3<double>(xB), lift3<double>(vB), lift3<double>(wB), r, cr, m1, hp, FA, TA, FB, TB);
    } else {
      const v3 z = mk(0., 0., 0.);
      f = contact_force(md, c, ov, dist, xA, vA, wA, m0, 1e-3, mk(wl.origin[0], wl.origin[1], wl.origin[2]),
                        z, z, 1.e20, wl.req);
      wall_torsor(wl, lift3<double>(xA), lift3<double>(vA), lift3<double>(wA), r, cr, m0, hp, FA, TA);
    }
    if (!same(f, FA)) ++forward_bad;
    {
      // is this contact sliding? (Coulomb branch)
      v3 nrm = round3(dvd(ov, norm(ov)));
      v3 tv = sub(add(vA, cross(wA, sub(c, xA))), wall ? mk(0., 0., 0.) : add(vB, cross(wB, sub(c, xB))));
      v3 vt = sub(tv, scl(nrm, dot(tv, nrm)));
      (void)vt;
      const double ft = norm(sub(f, scl(nrm, dot(f, nrm))));
      const double fnn = fabs(dot(f, nrm));
      if (ft > 0.999 * mu * fnn) ++sliding;
    }
    // derivatives: dual vs central differences, every input direction
    const int ndir = wall ? 12 : 21;
    for (int s = 0; s < ndir; ++s) {
      auto eval = [&](double hstep, bool dual, double* out12) {
        v3 in[6] = {xA, vA, wA, xB, vB, wB};
        double th[3] = {kn, en, mu};
        const int nvec = wall ? 3 : 6;
        if (!dual) {
          if (s < 3 * nvec) {
            double* comp = &in[s / 3].x + (s % 3);
            *comp += hstep;
          } else {
            th[s - 3 * nvec] += hstep;
          }
          const double L = log(th[1]);
          HookeParams<double> p{th[0], L / sqrt(kPi * kPi + L * L), th[2]};
          V3<double> a, b, cc, d;
          bool ok;
          if (!wall)
            ok = pair_torsors(lift3<double>(in[0]), lift3<double>(in[1]), lift3<double>(in[2]), r, cr, m0,
                              lift3<double>(in[3]), lift3<double>(in[4]), lift3<double>(in[5]), r, cr, m1, p,
                              a, b, cc, d);
          else
            ok = wall_torsor(wl, lift3<double>(in[0]), lift3<double>(in[1]), lift3<double>(in[2]), r, cr,
                             m0, p, a, b);
          if (!ok) return false;
          double o[12] = {a.x, a.y, a.z, b.x, b.y, b.z, 0, 0, 0, 0, 0, 0};
          if (!wall) { o[6] = cc.x; o[7] = cc.y; o[8] = cc.z; o[9] = d.x; o[10] = d.y; o[11] = d.z; }
          for (int k = 0; k < 12; ++k) out12[k] = o[k];
          return true;
        }
        auto seed = [&](v3 a, int base) {
          return mk3(Dual{a.x, s == base ? 1. : 0.}, Dual{a.y, s == base + 1 ? 1. : 0.},
                     Dual{a.z, s == base + 2 ? 1. : 0.});
        };
        const double L = log(en);
        const double beta = L / sqrt(kPi * kPi + L * L);
        const double dbeta = kPi * kPi / (en * pow(kPi * kPi + L * L, 1.5));
        const int tb = wall ? 9 : 18;
        HookeParams<Dual> p{Dual{kn, s == tb ? 1. : 0.}, Dual{beta, s == tb + 1 ? dbeta : 0.},
                            Dual{mu, s == tb + 2 ? 1. : 0.}};
        V3<Dual> a, b, cc, d;
        bool ok;
        if (!wall)
          ok = pair_torsors(seed(xA, 0), seed(vA, 3), seed(wA, 6), r, cr, m0, seed(xB, 9), seed(vB, 12),
                            seed(wB, 15), r, cr, m1, p, a, b, cc, d);
        else
          ok = wall_torsor(wl, seed(xA, 0), seed(vA, 3), seed(wA, 6), r, cr, m0, p, a, b);
        if (!ok) return false;
        double o[12] = {a.x.d, a.y.d, a.z.d, b.x.d, b.y.d, b.z.d, 0, 0, 0, 0, 0, 0};
        if (!wall) { o[6] = cc.x.d; o[7] = cc.y.d; o[8] = cc.z.d; o[9] = d.x.d; o[10] = d.y.d; o[11] = d.z.d; }
        for (int k = 0; k < 12; ++k) out12[k] = o[k];
        return true;
      };
      const int nvec = wall ? 3 : 6;
      double scale;
      if (s < 3 * nvec) {
        const int g = (s / 3) % 3;
        scale = g == 0 ? 1e-9 : (g == 1 ? 1e-7 : 1e-4);
      } else {
        const double th[3] = {kn, en, mu};
        scale = 1e-7 * th[s - 3 * nvec];
      }
      double dd[12], fp[12], fm[12];
      if (!eval(0., true, dd) || !eval(scale, false, fp) || !eval(-scale, false, fm)) continue;
      // roundoff floor of the central difference: outputs carry ~1e-16
      // relative error, divided by the step
      double fmax = 0., err = 0., omax = 0.;
      for (int k = 0; k < 12; ++k) {
        const double fd = (fp[k] - fm[k]) / (2. * scale);
        fmax = fmax > fabs(fd) ? fmax : fabs(fd);
        omax = omax > fabs(fp[k]) ? omax : fabs(fp[k]);
        const double e = fabs(fd - dd[k]);
        err = err > e ? err : e;
      }
      const double noise = 1e-13 * omax / scale;
      if (fmax <= noise) continue;
      ++deriv_checked;
      const double rel = err / (fmax + noise);
      if (rel > worst) worst = rel;
      if (rel > 1e-5) ++deriv_bad;
    }
  }
  printf("%d %d %d %d %d %d %d %d %.3e\n", forward_checked, forward_bad, deriv_checked, deriv_bad,
         sliding, wall_regions[0], wall_regions[1], wall_regions[2], worst);
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_dual_contact_matches_forward_and_differences(tmp_path):
    src = tmp_path / "dual.hip"
    src.write_text(DRIVER)
    exe = tmp_path / "dual"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-ffp-contract=off",
                           "-I", CSRC, "-o", str(exe), str(src)])
    out = subprocess.check_output([str(exe)]).decode().split()
    fwd, fwd_bad, der, der_bad, sliding = (int(v) for v in out[:5])
    regions = [int(v) for v in out[5:8]]
    worst = float(out[8])
    assert fwd > 1000 and fwd_bad == 0
    assert der > 10000 and der_bad == 0, (der_bad, worst)
    assert sliding > 50
    assert min(regions) > 100
