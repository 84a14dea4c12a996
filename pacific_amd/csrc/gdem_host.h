// gdem_host.h -- host-side construction of the device records: grid, Morton
// row ranks, walls, contact models and sphere particle constants.  Shared by
// the single-system ABI (gdem_capi.hip) and the batched ABI (gdem_batch.hip),
// so a system gets the same bits through either.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_internal.h"

#pragma clang fp contract(off)

namespace gdem {

// LinkedCell::set (Base/src/LinkedCell.cpp:47-119) without ghost layers:
// n_d = int((L_d + 1e-10) / h), edge = L_d / n_d.  False on a bad request.
inline bool make_grid(const gdem_config& cfg, Grid& g) {
  for (int d = 0; d < 3; ++d) {
    int n = 1;
    double e = cfg.size[d];
    if (!(cfg.size[d] > 0.)) return false;
    if (cfg.cell_size > kEps) {
      n = (int)((cfg.size[d] + kEps) / cfg.cell_size);
      if (!n) n = 1;
      e = cfg.size[d] / n;
    }
    if (cfg.periodic[d] && n < 3) return false;
    g.nc[d] = n;
    g.edge[d] = e;
    g.origin[d] = cfg.origin[d];
    g.size[d] = cfg.size[d];
    g.periodic[d] = cfg.periodic[d] ? 1 : 0;
  }
  const int64_t ncell = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
  if (ncell > 8 * 1024 * 1024) return false;
  g.ncell = (int32_t)ncell;
  return true;
}

// Rank of each cell row (j, k) in a 2-D Morton (Z) order of (j, k); the cell
// key is row_rank[j * nk + k] * ni + i, so the cells of a row along x are
// consecutive in the sorted order and neighbouring rows stay close.
inline std::vector<int32_t> row_morton_rank(const int32_t nc[3]) {
  auto spread = [](uint64_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 31; ++b) r |= ((v >> b) & 1ull) << (2 * b);
    return r;
  };
  std::vector<std::pair<uint64_t, int32_t>> codes;
  codes.reserve((size_t)nc[1] * nc[2]);
  for (int j = 0; j < nc[1]; ++j)
    for (int k = 0; k < nc[2]; ++k)
      codes.emplace_back(spread((uint64_t)k) | (spread((uint64_t)j) << 1), (int32_t)(j * nc[2] + k));
  std::sort(codes.begin(), codes.end());
  std::vector<int32_t> rank(codes.size());
  for (size_t r = 0; r < codes.size(); ++r) rank[(size_t)codes[r].second] = (int32_t)r;
  return rank;
}

// Pre-filter margin of the single-precision distance test: coordinates lie
// within [lo, lo + L] (a particle outside a non-periodic grid is a
// SimulationError), a shifted image within L more.  Rounding each to float
// moves it by at most 2^-24 of its magnitude; 2^-19 of (max |coordinate| +
// max L) bounds the error of the float distance test by a wide factor.
inline float prefilter_margin(const Grid& g, float* sizef) {
  double mag = 0., lmax = 0.;
  for (int dd = 0; dd < 3; ++dd) {
    mag = std::max(mag, std::max(std::fabs(g.origin[dd]), std::fabs(g.origin[dd] + g.size[dd])));
    lmax = std::max(lmax, g.size[dd]);
    sizef[dd] = (float)g.size[dd];
  }
  return (float)(std::ldexp(1.0, -19) * (mag + lmax));
}

// Static box obstacle: inverse transform (Transform::setToInverseTransform,
// Geometry/src/Transform.cpp:71-77), bounding box (Convex::bbox,
// Geometry/src/Convex.cpp:29-41, with Box::support, Box.cpp:307-316,
// BBox::setValue, BBox.cpp:235-239) and equivalent-sphere radius of the box
// volume (Box::getVolume, Box.cpp:273-279).  A Rectangle (kind
// WALL_RECTANGLE, half = (LX/2, LY/2, 0)) has the same bounding box
// (Rectangle::support, Rectangle.cpp:139-152, is the box's with a zero
// third extent), the equivalent radius of its area LX LY taken as a volume
// (Rectangle::getVolume, :182-185) and circumscribed radius
// sqrt(LX^2 + LY^2) / 2 (:118-121).
inline Wall make_wall(const double* basis, const double* centre, const double* half, double crust,
                      int32_t material, int32_t kind = WALL_BOX) {
  Wall wl{};
  wl.kind = kind;
  memcpy(wl.basis, basis, sizeof(wl.basis));
  memcpy(wl.origin, centre, sizeof(wl.origin));
  memcpy(wl.ext, half, sizeof(wl.ext));
  wl.crust = crust;
  wl.material = material;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) wl.inv_basis[3 * i + j] = basis[3 * j + i];
  for (int i = 0; i < 3; ++i) {
    v3 row = mk(wl.inv_basis[3 * i], wl.inv_basis[3 * i + 1], wl.inv_basis[3 * i + 2]);
    wl.inv_origin[i] = dot(neg(row), mk(centre[0], centre[1], centre[2]));
  }
  const double e0 = half[0], e1 = half[1], e2 = half[2];
  double mn[3], mx[3];
  for (int i = 0; i < 3; ++i) {
    v3 row = mk(basis[3 * i], basis[3 * i + 1], basis[3 * i + 2]);
    v3 nrow = neg(row);
    v3 sm = norm(nrow) < kEps ? mk(0., 0., 0.)
                              : mk(nrow.x < 0. ? -e0 : e0, nrow.y < 0. ? -e1 : e1, nrow.z < 0. ? -e2 : e2);
    v3 sp = norm(row) < kEps ? mk(0., 0., 0.)
                             : mk(row.x < 0. ? -e0 : e0, row.y < 0. ? -e1 : e1, row.z < 0. ? -e2 : e2);
    mn[i] = centre[i] + dot(row, sm);
    mx[i] = centre[i] + dot(row, sp);
  }
  for (int i = 0; i < 3; ++i) {
    wl.bb_ext[i] = (mx[i] - mn[i]) / 2.;
    wl.bb_centre[i] = mn[i] + wl.bb_ext[i];
  }
  if (kind == WALL_RECTANGLE) {
    const double lx = 2. * e0, ly = 2. * e1;
    wl.req = pow((0.75 / kPi) * (lx * ly), 1. / 3.);
    wl.circ = sqrt(lx * lx + ly * ly) / 2.;
    return wl;
  }
  const double vol = 8.0 * e0 * e1 * e2;
  wl.req = pow((0.75 / kPi) * vol, 1. / 3.);
  return wl;
}

// Hooke / Hertz parameters the step supports: en in (0, 1], a rolling
// resistance kr >= 0.
inline bool model_params_ok(int32_t kind, const double* p) {
  if (kind == GDEM_MODEL_HOOKE || kind == GDEM_MODEL_HERTZ) {
    if (!(p[4] >= 0.) || !std::isfinite(p[4])) return false;
    return p[1] > 0. && p[1] <= 1.;
  }
  if (kind == GDEM_MODEL_HOOKE_MEMORY) return p[1] > 0. && p[1] <= 1. && p[4] >= 0. && p[5] >= 0.;
  if (kind == GDEM_MODEL_HERTZ_MEMORY) return p[1] > 0. && p[1] <= 1. && p[0] > 0. && p[4] >= 0.;
  return false;
}

// A Hooke or Hertz model with rolling resistance (kr != 0): the sphere
// paths that store one record per pair (packed step, differentiated
// systems) and the convex path do not carry its moment.
inline bool model_has_rolling(const Model& m) {
  return (m.kind == GDEM_MODEL_HOOKE || m.kind == GDEM_MODEL_HERTZ) && m.p[4] != 0.;
}

inline bool model_has_memory(int32_t kind) {
  return kind == GDEM_MODEL_HOOKE_MEMORY || kind == GDEM_MODEL_HERTZ_MEMORY;
}

// Model constants (HookeContactForceModel.cpp:19, HertzContactForceModel.cpp:20-21,
// HookeMemoryContactForceModel.cpp:28, HertzMemoryContactForceModel.cpp:29-30).
// p holds model_nparams(kind) values.
inline Model make_model(int32_t kind, const double* p) {
  Model m{};
  m.kind = kind;
  memcpy(m.p, p, sizeof(double) * model_nparams(kind));
  m.beta = log(p[1]) / sqrt(kPi * kPi + log(p[1]) * log(p[1]));
  m.m2sqrt56 = -2. * sqrt(5. / 6.);
  return m;
}

// Device record of one sphere: Sphere::getVolume (Geometry/src/Sphere.cpp:152-155),
// Sphere::BuildInertia (:61-70), Particle.cpp:158-164,
// Component::getEquivalentSphereRadius (Component/src/Component.cpp:205-208),
// ParticleKinematics::setAngularVelocity (ParticleKinematics.cpp:197-205).
inline void sphere_particle(const double* x, const double* v, const double* w, const double* q,
                            double rr, double rho, double crust, int32_t mat, int32_t id,
                            double4& pr, double4& vm, double4& wr, double4& wb, double4& qq,
                            Aux& ax) {
  const double vol = 4.0 * kPi * pow(rr, 3.) / 3.0;
  const double in0 = 8.0 * kPi * pow(rr, 5.) / 15.0;
  const double mass = rho * vol;
  const double inertia = in0 * rho;
  const double req = pow((0.75 / kPi) * vol, 1. / 3.);
  const quat qi = qmk(mk(q[0], q[1], q[2]), q[3]);
  const v3 wbf = qmulv(qconj(qi), qleftvec(qi, mk(w[0], w[1], w[2])));
  pr = make_double4(x[0], x[1], x[2], rr);
  vm = make_double4(v[0], v[1], v[2], mass);
  wr = make_double4(w[0], w[1], w[2], req);
  wb = make_double4(wbf.x, wbf.y, wbf.z, inertia);
  qq = make_double4(q[0], q[1], q[2], q[3]);
  ax.id = id;
  ax.mat = mat;
  ax.crust = crust;
}

}  // namespace gdem
