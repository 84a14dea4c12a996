// gdem_reload.cpp -- particle records of the reference's reload files in the
// 2014 format (text "_Particles.result", hybrid binary "_Particles.result.bin"),
// written and read through the public C ABI (gdem_get_state / gdem_set_state,
// gdem_get_memory / gdem_set_memory).  Reference paths are relative to
// src/Grains3D/Grains/.
//
// One record (Particle::write2014 / write2014_binary, Component/src/Particle.cpp:1168-1199):
//   int geometric type, int id, int tag, unsigned activity
//   transform (Transform::writeTransform2014[_binary], Geometry/src/Transform.cpp:416-440):
//     origin (3 doubles), basis row by row (Matrix::writeMatrix2014[_binary],
//     Geometry/src/Matrix.cpp:291-320), unsigned type
//   kinematics (ParticleKinematics::writeParticleKinematics2014[_binary],
//     Component/src/ParticleKinematics.cpp:331-355): translational velocity,
//     quaternion as w then x, y, z (Quaternion::writeQuaternion[_binary],
//     Geometry/src/Quaternion.cpp:524-540), angular velocity
//   the leap-frog's force and torque (SecondOrderLeapFrog::
//     writeParticleKinematics2014[_binary], TimeIntegration/src/SecondOrderLeapFrog.cpp:102-122)
//   contact memory (Component::writeContactMemory2014[_binary],
//     Component/src/Component.cpp:776-835): int count, then per entry
//     int id0, id1, id2, active, and kdelta, previous normal, spring torque.
// The binary file precedes each record with its byte count as a size_t
// (AllComponents::write, Component/src/AllComponents.cpp:1155-1169).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gdem.h"

namespace {

// Transform types (Geometry/include/Transform.hh:246-256)
constexpr unsigned kTranslation = 0x01, kRotation = 0x02;

struct Record {
  int32_t type = 0, id = 0, tag = 0;
  uint32_t active = 1;
  double origin[3] = {0., 0., 0.}, basis[9] = {1., 0., 0., 0., 1., 0., 0., 0., 1.};
  uint32_t ttype = 0;
  double v[3] = {0., 0., 0.}, q[4] = {1., 0., 0., 0.}, w[3] = {0., 0., 0.};  // q: w, x, y, z
  double F[3] = {0., 0., 0.}, T[3] = {0., 0., 0.};
  struct Mem {
    int32_t id0 = 1, id1 = 0, id2 = 1, active = 1;
    double kd[3], nrm[3], spr[3];
  };
  std::vector<Mem> mem;
};

// GrainsExec::doubleToString(ios::scientific, 16, x) (Base/src/GrainsExec.cpp:213-225):
// scientific with 16 digits after the point, or the default stream form of 0.
std::string num(double x) {
  char b[64];
  if (x != 0.)
    snprintf(b, sizeof b, "%.16e", x);
  else
    snprintf(b, sizeof b, "%s", std::signbit(x) ? "-0" : "0");
  return b;
}

std::string group3(const double* a) { return num(a[0]) + " " + num(a[1]) + " " + num(a[2]); }

void write_text(FILE* f, const Record& r) {
  std::string s = std::to_string(r.type) + " " + std::to_string(r.id) + " " +
                  std::to_string(r.tag) + " " + std::to_string(r.active ? 1 : 0) + " ";
  s += num(r.origin[0]) + " " + num(r.origin[1]) + " " + num(r.origin[2]) + " ";
  s += group3(r.basis) + " " + group3(r.basis + 3) + " " + group3(r.basis + 6);
  s += " " + std::to_string(r.ttype) + " ";
  s += group3(r.v) + " " + num(r.q[0]) + " " + group3(r.q + 1) + " " + group3(r.w);
  s += " " + group3(r.F) + " " + group3(r.T);
  s += " " + std::to_string((int)r.mem.size());
  for (const Record::Mem& m : r.mem) {
    s += " " + std::to_string(m.id0) + " " + std::to_string(m.id1) + " " + std::to_string(m.id2) +
         " " + std::to_string(m.active) + " " + group3(m.kd) + " " + group3(m.nrm) + " " +
         group3(m.spr);
  }
  s += "\n";
  fputs(s.c_str(), f);
}

template <class T>
void put(std::vector<char>& b, const T& v) {
  const char* p = reinterpret_cast<const char*>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}

void write_binary(FILE* f, const Record& r) {
  std::vector<char> b;
  put(b, r.type);
  put(b, r.id);
  put(b, r.tag);
  put(b, r.active);
  for (double x : r.origin) put(b, x);
  for (double x : r.basis) put(b, x);
  put(b, r.ttype);
  for (double x : r.v) put(b, x);
  for (double x : r.q) put(b, x);
  for (double x : r.w) put(b, x);
  for (double x : r.F) put(b, x);
  for (double x : r.T) put(b, x);
  put(b, (int32_t)r.mem.size());
  for (const Record::Mem& m : r.mem) {
    put(b, m.id0);
    put(b, m.id1);
    put(b, m.id2);
    put(b, m.active);
    for (double x : m.kd) put(b, x);
    for (double x : m.nrm) put(b, x);
    for (double x : m.spr) put(b, x);
  }
  const size_t nb = b.size();
  fwrite(&nb, sizeof nb, 1, f);
  fwrite(b.data(), 1, b.size(), f);
}

// Reads the next text record; false at the end of the file or on a format error.
bool read_text(FILE* f, Record& r, bool& bad) {
  bad = false;
  int active = 0;
  if (fscanf(f, "%d %d %d %d", &r.type, &r.id, &r.tag, &active) != 4) {
    bad = !feof(f);
    return false;
  }
  r.active = (uint32_t)active;
  auto rd = [&](double* a, int k) {
    for (int i = 0; i < k; ++i)
      if (fscanf(f, "%lf", &a[i]) != 1) return false;
    return true;
  };
  int nm = 0;
  if (!rd(r.origin, 3) || !rd(r.basis, 9) || fscanf(f, "%u", &r.ttype) != 1 || !rd(r.v, 3) ||
      !rd(r.q, 4) || !rd(r.w, 3) || !rd(r.F, 3) || !rd(r.T, 3) || fscanf(f, "%d", &nm) != 1 ||
      nm < 0) {
    bad = true;
    return false;
  }
  r.mem.assign((size_t)nm, Record::Mem{});
  for (Record::Mem& m : r.mem) {
    if (fscanf(f, "%d %d %d %d", &m.id0, &m.id1, &m.id2, &m.active) != 4 || !rd(m.kd, 3) ||
        !rd(m.nrm, 3) || !rd(m.spr, 3)) {
      bad = true;
      return false;
    }
  }
  return true;
}

template <class T>
bool get(const std::vector<char>& b, size_t& at, T& v) {
  if (at + sizeof(T) > b.size()) return false;
  memcpy(&v, b.data() + at, sizeof(T));
  at += sizeof(T);
  return true;
}

bool read_binary(FILE* f, Record& r, bool& bad) {
  bad = false;
  size_t nb = 0;
  if (fread(&nb, sizeof nb, 1, f) != 1) return false;
  if (nb > (1u << 30)) {
    bad = true;
    return false;
  }
  std::vector<char> b(nb);
  if (fread(b.data(), 1, nb, f) != nb) {
    bad = true;
    return false;
  }
  size_t at = 0;
  bool ok = get(b, at, r.type) && get(b, at, r.id) && get(b, at, r.tag) && get(b, at, r.active);
  for (double& x : r.origin) ok = ok && get(b, at, x);
  for (double& x : r.basis) ok = ok && get(b, at, x);
  ok = ok && get(b, at, r.ttype);
  for (double& x : r.v) ok = ok && get(b, at, x);
  for (double& x : r.q) ok = ok && get(b, at, x);
  for (double& x : r.w) ok = ok && get(b, at, x);
  for (double& x : r.F) ok = ok && get(b, at, x);
  for (double& x : r.T) ok = ok && get(b, at, x);
  int32_t nm = 0;
  ok = ok && get(b, at, nm) && nm >= 0;
  if (ok) r.mem.assign((size_t)nm, Record::Mem{});
  for (size_t k = 0; ok && k < r.mem.size(); ++k) {
    Record::Mem& m = r.mem[k];
    ok = get(b, at, m.id0) && get(b, at, m.id1) && get(b, at, m.id2) && get(b, at, m.active);
    for (double& x : m.kd) ok = ok && get(b, at, x);
    for (double& x : m.nrm) ok = ok && get(b, at, x);
    for (double& x : m.spr) ok = ok && get(b, at, x);
  }
  // the byte count must match what the record holds
  bad = !ok || at != nb;
  return !bad;
}

// Rotation matrix of a unit quaternion (w; x, y, z), row-major.
void rotation(const double* q, double* b) {
  const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
  b[0] = 1. - 2. * (qy * qy + qz * qz); b[1] = 2. * (qx * qy - qz * qw); b[2] = 2. * (qx * qz + qy * qw);
  b[3] = 2. * (qx * qy + qz * qw); b[4] = 1. - 2. * (qx * qx + qz * qz); b[5] = 2. * (qy * qz - qx * qw);
  b[6] = 2. * (qx * qz - qy * qw); b[7] = 2. * (qy * qz + qx * qw); b[8] = 1. - 2. * (qx * qx + qy * qy);
}

}  // namespace

extern "C" int gdem_write_particles_2014(gdem_ctx* ctx, const char* path, int32_t binary,
                                         int32_t id_offset, int32_t wall_id_offset,
                                         const int32_t* geom_type) {
  if (!ctx || !path) return GDEM_ERR_ARG;
  int64_t n = 0;
  int32_t nwalls = 0, convex = 0, memory = 0;
  int rc = gdem_info(ctx, &n, &nwalls, &convex, &memory);
  if (rc) return rc;
  const size_t nn = (size_t)n;
  std::vector<double> x(3 * nn), v(3 * nn), w(3 * nn), q(4 * nn), F(3 * nn), T(3 * nn), basis;
  rc = gdem_get_state(ctx, x.data(), v.data(), w.data(), q.data(), F.data(), T.data());
  if (rc) return rc;
  if (convex) {
    basis.resize(9 * nn);
    rc = gdem_get_basis(ctx, basis.data());
    if (rc) return rc;
  }
  std::vector<int32_t> owner, key;
  std::vector<double> kd, nrm, spr;
  if (memory) {
    int64_t cnt = 0;
    rc = gdem_get_memory(ctx, 0, &cnt, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    owner.resize((size_t)cnt);
    key.resize((size_t)cnt);
    kd.resize(3 * (size_t)cnt);
    nrm.resize(3 * (size_t)cnt);
    spr.resize(3 * (size_t)cnt);
    rc = gdem_get_memory(ctx, cnt, &cnt, owner.data(), key.data(), kd.data(), nrm.data(),
                         spr.data());
    if (rc) return rc;
  }
  FILE* f = fopen(path, binary ? "wb" : "w");
  if (!f) return GDEM_ERR_ARG;
  size_t e = 0;
  for (size_t i = 0; i < nn; ++i) {
    Record r;
    r.type = geom_type ? geom_type[i] : 0;
    r.id = (int32_t)i + id_offset;
    for (int d = 0; d < 3; ++d) {
      r.origin[d] = x[3 * i + d];
      r.v[d] = v[3 * i + d];
      r.w[d] = w[3 * i + d];
      r.F[d] = F[3 * i + d];
      r.T[d] = T[3 * i + d];
    }
    r.q[0] = q[4 * i + 3];  // w first
    r.q[1] = q[4 * i];
    r.q[2] = q[4 * i + 1];
    r.q[3] = q[4 * i + 2];
    if (convex)
      memcpy(r.basis, &basis[9 * i], sizeof r.basis);
    else
      rotation(r.q, r.basis);
    const bool moved = r.origin[0] != 0. || r.origin[1] != 0. || r.origin[2] != 0.;
    r.ttype = (moved ? kTranslation : 0u) | kRotation;
    // the memory entries of particle i, as the reference's map order: by key
    // tuple (1, partner id, 1)
    std::vector<Record::Mem> mine;
    for (; e < owner.size() && owner[e] == (int32_t)i; ++e) {
      Record::Mem m;
      m.id1 = key[e] >= 0 ? key[e] + id_offset : wall_id_offset + (-1 - key[e]);
      for (int d = 0; d < 3; ++d) {
        m.kd[d] = kd[3 * e + d];
        m.nrm[d] = nrm[3 * e + d];
        m.spr[d] = spr[3 * e + d];
      }
      mine.push_back(m);
    }
    std::sort(mine.begin(), mine.end(),
              [](const Record::Mem& a, const Record::Mem& b) { return a.id1 < b.id1; });
    r.mem = std::move(mine);
    if (binary)
      write_binary(f, r);
    else
      write_text(f, r);
  }
  const bool ok = !ferror(f);
  fclose(f);
  return ok ? GDEM_OK : GDEM_ERR_ARG;
}

extern "C" int gdem_read_particles_2014(gdem_ctx* ctx, const char* path, int32_t binary,
                                        int32_t id_offset, int32_t wall_id_offset,
                                        int32_t* geom_type_out) {
  if (!ctx || !path) return GDEM_ERR_ARG;
  int64_t n = 0;
  int32_t nwalls = 0, convex = 0, memory = 0;
  int rc = gdem_info(ctx, &n, &nwalls, &convex, &memory);
  if (rc) return rc;
  if (convex) return GDEM_ERR_ARG;
  const size_t nn = (size_t)n;
  FILE* f = fopen(path, binary ? "rb" : "r");
  if (!f) return GDEM_ERR_ARG;
  std::vector<Record> recs;
  for (;;) {
    Record r;
    bool bad = false;
    const bool got = binary ? read_binary(f, r, bad) : read_text(f, r, bad);
    if (bad) {
      fclose(f);
      return GDEM_ERR_ARG;
    }
    if (!got) break;
    recs.push_back(std::move(r));
  }
  fclose(f);
  if (recs.size() != nn) return GDEM_ERR_ARG;
  std::vector<double> x(3 * nn), v(3 * nn), w(3 * nn), q(4 * nn), F(3 * nn), T(3 * nn);
  std::vector<char> seen(nn, 0);
  std::vector<int32_t> owner, key;
  std::vector<double> kd, nrm, spr;
  for (const Record& r : recs) {
    const int64_t i = (int64_t)r.id - id_offset;
    if (i < 0 || i >= n || seen[(size_t)i]) return GDEM_ERR_ARG;
    seen[(size_t)i] = 1;
    if (geom_type_out) geom_type_out[i] = r.type;
    for (int d = 0; d < 3; ++d) {
      x[3 * i + d] = r.origin[d];
      v[3 * i + d] = r.v[d];
      w[3 * i + d] = r.w[d];
      F[3 * i + d] = r.F[d];
      T[3 * i + d] = r.T[d];
    }
    q[4 * i] = r.q[1];
    q[4 * i + 1] = r.q[2];
    q[4 * i + 2] = r.q[3];
    q[4 * i + 3] = r.q[0];
    for (const Record::Mem& m : r.mem) {
      if (!m.active) continue;  // inactive entries are erased before any use
      int32_t k;
      if (m.id1 - id_offset >= 0 && m.id1 - id_offset < n)
        k = m.id1 - id_offset;
      else if (m.id1 - wall_id_offset >= 0 && m.id1 - wall_id_offset < nwalls)
        k = -1 - (m.id1 - wall_id_offset);
      else
        return GDEM_ERR_ARG;
      owner.push_back((int32_t)i);
      key.push_back(k);
      kd.insert(kd.end(), m.kd, m.kd + 3);
      nrm.insert(nrm.end(), m.nrm, m.nrm + 3);
      spr.insert(spr.end(), m.spr, m.spr + 3);
    }
  }
  if (!owner.empty() && !memory) return GDEM_ERR_ARG;  // no memory model to hold it
  rc = gdem_set_state(ctx, x.data(), v.data(), w.data(), q.data(), F.data(), T.data());
  if (rc) return rc;
  if (memory)
    rc = gdem_set_memory(ctx, (int64_t)owner.size(), owner.data(), key.data(), kd.data(),
                         nrm.data(), spr.data());
  return rc;
}

// ---------------------------------------------------------------- main file
// Grains::saveReload (Base/src/Grains.cpp:2543-2581) and its parts; reading as
// Grains::readReload does for the parts a sphere step uses.
namespace {

// ostream << double with the stream's default format (precision 6), as the
// reference writes times, radii, crusts, extents and XML parameter values.
std::string dflt(double x) {
  std::ostringstream o;
  o << x;
  return o.str();
}

std::string base_name(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? path : path.substr(k + 1);
}

std::string dir_name(const std::string& path) {
  const size_t k = path.find_last_of('/');
  return k == std::string::npos ? std::string(".") : path.substr(0, k);
}

// Parameter names of each model kind, in gdem_set_contact_model's order
// (HookeContactForceModel.cpp:15-19, HertzContactForceModel.cpp:15-19,
// HookeMemoryContactForceModel.cpp:14-21, HertzMemoryContactForceModel.cpp:16-22).
const std::vector<std::string>& param_names(int32_t kind) {
  static const std::vector<std::string> hooke{"kn", "en", "etat", "muc", "kr"};
  static const std::vector<std::string> hertz{"Es", "en", "Gs", "muc", "kr"};
  static const std::vector<std::string> hooke_mem{"kn", "en", "etat", "muc", "kt", "mur", "etarpf",
                                                  "epst"};
  static const std::vector<std::string> hertz_mem{"Es", "en", "Gs", "muc", "mur", "etarpf", "epst"};
  static const std::vector<std::string> none;
  return kind == GDEM_MODEL_HOOKE ? hooke : kind == GDEM_MODEL_HERTZ ? hertz
       : kind == GDEM_MODEL_HOOKE_MEMORY ? hooke_mem : kind == GDEM_MODEL_HERTZ_MEMORY ? hertz_mem
       : none;
}

// ContactBuilderFactory::ContactType (Component/include/ContactBuilderFactory.hh:85-96)
int32_t xml_type(int32_t kind) {
  return kind == GDEM_MODEL_HOOKE ? 0 : kind == GDEM_MODEL_HOOKE_MEMORY ? 1
       : kind == GDEM_MODEL_HERTZ ? 2 : 3;
}
int32_t kind_of_xml(int32_t t) {
  return t == 0 ? GDEM_MODEL_HOOKE : t == 1 ? GDEM_MODEL_HOOKE_MEMORY : t == 2 ? GDEM_MODEL_HERTZ
       : t == 3 ? GDEM_MODEL_HERTZ_MEMORY : 0;
}

// Matrix::writeMatrix (Geometry/src/Matrix.cpp:276-283): three rows, the last
// without a line end.
std::string matrix_rows(const double* b) {
  return group3(b) + "\n" + group3(b + 3) + "\n" + group3(b + 6);
}

struct PType {
  double r, rho, crust;
  int32_t mat;
};

std::string mat_name(gdem_ctx* ctx, int32_t k) {
  char buf[256];
  if (gdem_get_material_name(ctx, k, buf, sizeof buf) == GDEM_OK && buf[0]) return buf;
  return "mat" + std::to_string(k);
}

std::string wall_name(gdem_ctx* ctx, int32_t k) {
  char buf[256];
  if (gdem_get_wall_name(ctx, k, buf, sizeof buf) == GDEM_OK && buf[0]) return buf;
  return "Obstacle" + std::to_string(k);
}

}  // namespace

extern "C" int gdem_write_reload(gdem_ctx* ctx, const char* root, const char* suffix, double time,
                                 int32_t binary, int32_t id_offset, int32_t wall_id_offset) {
  if (!ctx || !root || !suffix) return GDEM_ERR_ARG;
  int64_t n = 0;
  int32_t nwalls = 0, convex = 0, memory = 0;
  int rc = gdem_info(ctx, &n, &nwalls, &convex, &memory);
  if (rc) return rc;
  if (convex) return GDEM_ERR_ARG;  // sphere reference particles only
  const size_t nn = (size_t)n;
  std::vector<double> radius(nn), density(nn), crust(nn);
  std::vector<int32_t> material(nn);
  rc = gdem_get_particles(ctx, radius.data(), density.data(), crust.data(), material.data());
  if (rc) return rc;
  // reference particle types: distinct constants in order of first appearance
  std::vector<PType> types;
  std::vector<int32_t> geom(nn);
  for (size_t i = 0; i < nn; ++i) {
    size_t t = 0;
    for (; t < types.size(); ++t)
      if (types[t].r == radius[i] && types[t].rho == density[i] && types[t].crust == crust[i] &&
          types[t].mat == material[i])
        break;
    if (t == types.size()) types.push_back(PType{radius[i], density[i], crust[i], material[i]});
    geom[i] = (int32_t)t;
  }
  // materials in use (particles, walls, contact models)
  bool used[64] = {};
  for (const PType& t : types) used[t.mat] = true;
  std::vector<double> wb(9), wc(3), wh(3);
  for (int32_t k = 0; k < nwalls; ++k) {
    double wcr = 0.;
    int32_t wm = 0, wk = 0;
    rc = gdem_get_wall_shape(ctx, k, nullptr, nullptr, nullptr, &wcr, &wm, &wk);
    if (rc) return rc;
    if (wk != GDEM_WALL_BOX) return GDEM_ERR_ARG;  // box obstacles only
    if (wm >= 0 && wm < 31) used[wm] = true;
  }
  struct ModelEntry {
    int32_t key, kind;
    double p[8];
  };
  std::vector<ModelEntry> models;
  for (int32_t a = 0; a < 31; ++a)
    for (int32_t b = a; b < 31; ++b) {
      ModelEntry e{};
      if (gdem_get_contact_model(ctx, a, b, &e.kind, e.p) != GDEM_OK || e.kind == 0) continue;
      e.key = (1 << a) | (1 << b);
      used[a] = used[b] = true;
      models.push_back(e);
    }
  std::sort(models.begin(), models.end(),
            [](const ModelEntry& x, const ModelEntry& y) { return x.key < y.key; });
  const std::string rootS(root), reload = rootS + suffix;
  // ContactBuilderFactory::save: materials by name, models by key, each
  // model's parameters by name (std::map orders)
  {
    std::vector<std::pair<std::string, int32_t>> mats;
    for (int32_t k = 0; k < 31; ++k)
      if (used[k]) mats.push_back({mat_name(ctx, k), 1 << k});
    std::sort(mats.begin(), mats.end());
    std::string x = "<?xml version=\"1.0\" encoding=\"UTF-8\" standalone=\"no\" ?>\n<GRAINS>\n\n";
    x += "  <Materials>\n";
    for (const auto& m : mats)
      x += "    <Material value=\"" + std::to_string(m.second) + "\">" + m.first + "</Material>\n";
    x += "  </Materials>\n\n  <ContactForceModels>\n";
    for (const ModelEntry& e : models) {
      x += "    <ContactForceModel type=\"" + std::to_string(xml_type(e.kind)) + "\" value=\"" +
           std::to_string(e.key) + "\">\n      <Parameters>\n";
      const std::vector<std::string>& names = param_names(e.kind);
      std::vector<std::pair<std::string, double>> ps;
      for (size_t k = 0; k < names.size(); ++k) ps.push_back({names[k], e.p[k]});
      std::sort(ps.begin(), ps.end());
      for (const auto& pv : ps)
        x += "        <Parameter name=\"" + pv.first + "\">" + dflt(pv.second) + "</Parameter>\n";
      x += "      </Parameters>\n    </ContactForceModel>\n";
    }
    x += "  </ContactForceModels>\n\n</GRAINS>\n";
    FILE* f = fopen((rootS + "_MatContact.xml").c_str(), "w");
    if (!f) return GDEM_ERR_ARG;
    fputs(x.c_str(), f);
    fclose(f);
  }
  // the main file
  std::string s = "__Format2024__\n#Time " + dflt(time) + "\n";
  s += "<ContactForceModels> " + base_name(rootS + "_MatContact.xml") + " </ContactForceModels>\n";
  s += "\nNumberOfParticleTypes " + std::to_string(types.size()) + "\n";
  const double kPiH = 3.141592653589793;
  for (size_t t = 0; t < types.size(); ++t) {
    const PType& p = types[t];
    // the sphere's mass and principal inertia as gdem_set_particles forms them
    const double vol = 4.0 * kPiH * std::pow(p.r, 3.) / 3.0;
    const double mass = p.rho * vol;
    // Sphere::BuildInertia (Geometry/src/Sphere.cpp:61-70), then scaled by
    // the density (Particle.cpp:158-163)
    const double in0 = 8.0 * kPiH * std::pow(p.r, 5.) / 15.0;
    const double inertia = in0 * p.rho;
    const double inv = (1.0 / in0) / p.rho;
    s += "\n<Particle>\n*IDnumber " + std::to_string(t) + "\n*Material " + mat_name(ctx, p.mat) +
         "\n*Sphere " + dflt(p.r) + " *END\n*CrustThickness " + dflt(p.crust) + "\n";
    s += "*Type " + std::to_string(t) + "\n*Tag 0\n*Mass " + num(mass) + "\n*Density " +
         num(p.rho) + "\n";
    s += "*MomentOfInertiaTensor\n" + num(inertia) + " 0 0 " + num(inertia) + " 0 " + num(inertia) +
         "\n";
    s += "*InverseOfMOIT\n" + num(inv) + " 0 0 " + num(inv) + " 0 " + num(inv) + "\n";
    const double I3[9] = {1., 0., 0., 0., 1., 0., 0., 0., 1.}, z3[3] = {0., 0., 0.};
    s += "*Position 0\n" + group3(z3) + "\n" + matrix_rows(I3) + "\n";
    s += "*Activity 0\n*ParticleKinematics3D\n" + group3(z3) + "\n" + num(1.) + " " + group3(z3) +
         "\n" + group3(z3) + "\n</Particle>\n";
  }
  s += "\nRemainingNbParticlesToInsertPerType";
  for (size_t t = 0; t < types.size(); ++t) s += " 0";
  s += "\n\nRemainingKnownInsertionPositions 0\n";
  s += "\n<Obstacle>\n<Composite> __AllObstacles___ Standard\n";
  for (int32_t k = 0; k < nwalls; ++k) {
    double cr = 0.;
    int32_t wm = 0, wk = 0;
    rc = gdem_get_wall_shape(ctx, k, wb.data(), wc.data(), wh.data(), &cr, &wm, &wk);
    if (rc) return rc;
    const bool moved = wc[0] != 0. || wc[1] != 0. || wc[2] != 0.;
    const bool turned = !(wb[0] == 1. && wb[1] == 0. && wb[2] == 0. && wb[3] == 0. && wb[4] == 1. &&
                          wb[5] == 0. && wb[6] == 0. && wb[7] == 0. && wb[8] == 1.);
    const unsigned tt = (moved ? kTranslation : 0u) | (turned ? kRotation : 0u);
    s += "<Simple> " + wall_name(ctx, k) + " " + std::to_string(wall_id_offset + k) + "\n";
    s += "*Torsor " + dflt(wc[0]) + " " + dflt(wc[1]) + " " + dflt(wc[2]) + " 0 0 0 0 0 0\n";
    s += "*Material " + mat_name(ctx, wm) + "\n";
    s += "*Box " + dflt(wh[0]) + " " + dflt(wh[1]) + " " + dflt(wh[2]) + " *END\n*CrustThickness " +
         dflt(cr) + "\n";
    s += "*ToFluid 0\n*Position " + std::to_string(tt) + "\n" + group3(wc.data()) + "\n" +
         matrix_rows(wb.data()) + "\n 0\n</Simple>\n";
  }
  s += "</Composite>\n</Obstacle>\n\n";
  s += "NbFiles 1\n" + std::string(binary ? "Hybrid " : "Text ") + std::to_string(nn) + "\n";
  s += "#EndTime\n";
  {
    FILE* f = fopen((reload + ".result").c_str(), "w");
    if (!f) return GDEM_ERR_ARG;
    fputs(s.c_str(), f);
    fclose(f);
  }
  rc = gdem_write_particles_2014(ctx, (reload + "_Particles.result" + (binary ? ".bin" : "")).c_str(),
                                 binary, id_offset, wall_id_offset, geom.data());
  if (rc) return rc;
  FILE* f = fopen((rootS + "_RFTable.txt").c_str(), "a");
  if (!f) return GDEM_ERR_ARG;
  fputs((dflt(time) + " " + reload + "\n").c_str(), f);
  fclose(f);
  return GDEM_OK;
}

namespace {

// Tokens of a main file (whitespace-separated, as the reference's >> reads).
struct Tokens {
  std::ifstream in;
  std::string tok;
  bool next() { return bool(in >> tok); }
  bool is(const char* s) { return next() && tok == s; }
  bool dbl(double& x) {
    if (!next()) return false;
    char* end = nullptr;
    x = strtod(tok.c_str(), &end);
    return end && *end == '\0';
  }
  bool dbls(double* x, int k) {
    for (int i = 0; i < k; ++i)
      if (!dbl(x[i])) return false;
    return true;
  }
  bool integer(long long& v) {
    if (!next()) return false;
    char* end = nullptr;
    v = strtoll(tok.c_str(), &end, 10);
    return end && *end == '\0';
  }
};

struct RefType {
  std::string material;
  double r = 0., crust = 0., rho = 0.;
  long long type = -1;
};

struct Obst {
  std::string name, material;
  long long id = 0;
  double basis[9], origin[3], half[3], crust = 0.;
};

// One reference particle (Particle::write): the fields a sphere type has.
bool read_particle(Tokens& t, RefType& p) {
  while (t.next()) {
    const std::string k = t.tok;
    long long iv;
    double d[12];
    if (k == "</Particle>") return p.type >= 0 && p.r > 0. && p.rho > 0. && p.crust > 0.;
    if (k == "*IDnumber" || k == "*Tag" || k == "*Activity") {
      if (!t.integer(iv)) return false;
    } else if (k == "*Type") {
      if (!t.integer(p.type)) return false;
    } else if (k == "*Material") {
      if (!t.next()) return false;
      p.material = t.tok;
    } else if (k == "*Sphere") {
      if (!t.dbl(p.r) || !t.is("*END")) return false;
    } else if (k == "*CrustThickness") {
      if (!t.dbl(p.crust)) return false;
    } else if (k == "*Mass") {
      if (!t.dbl(d[0])) return false;
    } else if (k == "*Density") {
      if (!t.dbl(p.rho)) return false;
    } else if (k == "*MomentOfInertiaTensor" || k == "*InverseOfMOIT") {
      if (!t.dbls(d, 6)) return false;
    } else if (k == "*Position") {
      if (!t.integer(iv) || !t.dbls(d, 12)) return false;
    } else if (k == "*ParticleKinematics3D") {
      if (!t.dbls(d, 10)) return false;
    } else {
      return false;  // another shape or field: not a sphere type
    }
  }
  return false;
}

// A simple obstacle (SimpleObstacle::write) with a box shape.
bool read_simple(Tokens& t, std::vector<Obst>& out) {
  Obst o;
  if (!t.next()) return false;
  o.name = t.tok;
  if (!t.integer(o.id)) return false;
  double d[9];
  long long iv;
  bool shape = false;
  while (t.next()) {
    const std::string k = t.tok;
    if (k == "</Simple>") {
      if (!shape) return false;
      out.push_back(o);
      return true;
    }
    if (k == "*Torsor") {
      if (!t.dbls(d, 9)) return false;
    } else if (k == "*Material") {
      if (!t.next()) return false;
      o.material = t.tok;
    } else if (k == "*Box") {
      if (!t.dbls(o.half, 3) || !t.is("*END")) return false;
      shape = true;
    } else if (k == "*CrustThickness") {
      if (!t.dbl(o.crust)) return false;
    } else if (k == "*ToFluid") {
      if (!t.integer(iv)) return false;
    } else if (k == "*Position") {
      if (!t.integer(iv) || !t.dbls(o.origin, 3) || !t.dbls(o.basis, 9)) return false;
      // the contact memory follows the transform: a count, then entries of
      // three ids, a flag and three vectors (Component::writeContactMemory2014)
      long long cnt = 0;
      if (!t.integer(cnt) || cnt < 0) return false;
      for (long long e = 0; e < cnt; ++e)
        if (!t.dbls(d, 4) || !t.dbls(d, 9)) return false;
    } else {
      return false;
    }
  }
  return false;
}

bool read_composite(Tokens& t, std::vector<Obst>& out) {
  if (!t.next() || !t.next()) return false;  // name, type
  while (t.next()) {
    const std::string k = t.tok;
    if (k == "</Composite>") return true;
    if (k == "*Torsor") {
      double d[9];
      if (!t.dbls(d, 9)) return false;
    } else if (k == "<Simple>") {
      if (!read_simple(t, out)) return false;
    } else if (k == "<Composite>") {
      if (!read_composite(t, out)) return false;
    } else {
      return false;
    }
  }
  return false;
}

// Text between the first `open` at or after `at` and the next `close`.
bool between(const std::string& s, size_t& at, const std::string& open, const std::string& close,
             std::string& out) {
  const size_t a = s.find(open, at);
  if (a == std::string::npos) return false;
  const size_t b = s.find(close, a + open.size());
  if (b == std::string::npos) return false;
  out = s.substr(a + open.size(), b - a - open.size());
  at = b + close.size();
  return true;
}

// Value of attribute `name` in an element's opening tag.
bool attr(const std::string& tag, const std::string& name, std::string& v) {
  const size_t a = tag.find(name + "=\"");
  if (a == std::string::npos) return false;
  const size_t b = tag.find('"', a + name.size() + 2);
  if (b == std::string::npos) return false;
  v = tag.substr(a + name.size() + 2, b - a - name.size() - 2);
  return true;
}

}  // namespace

extern "C" int gdem_read_reload(const gdem_config* cfg, const char* main_path, int32_t id_offset,
                                int32_t wall_id_offset, gdem_ctx** out, double* time) {
  if (!cfg || !main_path || !out) return GDEM_ERR_ARG;
  *out = nullptr;
  const std::string mainS(main_path);
  Tokens t;
  t.in.open(mainS);
  if (!t.in || !t.is("__Format2024__") || !t.is("#Time")) return GDEM_ERR_ARG;
  double tm = 0.;
  if (!t.dbl(tm) || !t.is("<ContactForceModels>") || !t.next()) return GDEM_ERR_ARG;
  const std::string xml_file = dir_name(mainS) + "/" + base_name(t.tok);
  if (!t.is("</ContactForceModels>") || !t.is("NumberOfParticleTypes")) return GDEM_ERR_ARG;
  long long ntypes = 0;
  if (!t.integer(ntypes) || ntypes < 0) return GDEM_ERR_ARG;
  std::vector<RefType> types;
  for (long long k = 0; k < ntypes; ++k) {
    RefType p;
    if (!t.is("<Particle>") || !read_particle(t, p)) return GDEM_ERR_ARG;
    types.push_back(p);
  }
  if (!t.is("RemainingNbParticlesToInsertPerType")) return GDEM_ERR_ARG;
  long long iv = 0;
  for (long long k = 0; k < ntypes; ++k)
    if (!t.integer(iv)) return GDEM_ERR_ARG;
  if (!t.is("RemainingKnownInsertionPositions") || !t.integer(iv) || iv < 0) return GDEM_ERR_ARG;
  for (long long k = 0; k < iv; ++k) {
    double d[3];
    if (!t.dbls(d, 3)) return GDEM_ERR_ARG;
  }
  std::vector<Obst> obst;
  if (!t.is("<Obstacle>") || !t.is("<Composite>") || !read_composite(t, obst) ||
      !t.is("</Obstacle>"))
    return GDEM_ERR_ARG;
  long long nfiles = 0, npart = 0;
  if (!t.is("NbFiles") || !t.integer(nfiles) || nfiles != 1 || !t.next()) return GDEM_ERR_ARG;
  const std::string mode = t.tok;
  if ((mode != "Hybrid" && mode != "Text") || !t.integer(npart) || npart < 0 ||
      !t.is("#EndTime"))
    return GDEM_ERR_ARG;
  const bool binary = mode == "Hybrid";
  // materials and contact models (ContactBuilderFactory::reload,
  // Component/src/ContactBuilderFactory.cpp:206-262)
  std::string xml;
  {
    std::ifstream in(xml_file);
    if (!in) return GDEM_ERR_ARG;
    std::stringstream ss;
    ss << in.rdbuf();
    xml = ss.str();
  }
  std::vector<std::pair<std::string, int32_t>> mats;  // name, index
  {
    size_t at = 0;
    std::string body;
    if (!between(xml, at, "<Materials>", "</Materials>", body)) return GDEM_ERR_ARG;
    size_t m = 0;
    while ((m = body.find("<Material ", m)) != std::string::npos) {
      const size_t gt = body.find('>', m), end = body.find("</Material>", m);
      if (gt == std::string::npos || end == std::string::npos) return GDEM_ERR_ARG;
      std::string v;
      if (!attr(body.substr(m, gt - m), "value", v)) return GDEM_ERR_ARG;
      const long long value = strtoll(v.c_str(), nullptr, 10);
      int32_t idx = 0;
      while (idx < 31 && (1ll << idx) != value) ++idx;
      if (idx == 31) return GDEM_ERR_ARG;
      mats.push_back({body.substr(gt + 1, end - gt - 1), idx});
      m = end;
    }
  }
  auto mat_index = [&](const std::string& name) {
    for (const auto& m : mats)
      if (m.first == name) return m.second;
    return -1;
  };
  gdem_ctx* c = nullptr;
  int rc = gdem_create(cfg, &c);
  if (rc) return rc;
  auto fail = [&](int code) {
    gdem_destroy(c);
    return code;
  };
  for (const auto& m : mats) gdem_set_material_name(c, m.second, m.first.c_str());
  {
    size_t at = 0;
    std::string models;
    if (!between(xml, at, "<ContactForceModels>", "</ContactForceModels>", models))
      return fail(GDEM_ERR_ARG);
    size_t m = 0;
    while ((m = models.find("<ContactForceModel ", m)) != std::string::npos) {
      const size_t gt = models.find('>', m), end = models.find("</ContactForceModel>", m);
      if (gt == std::string::npos || end == std::string::npos) return fail(GDEM_ERR_ARG);
      const std::string tag = models.substr(m, gt - m), body = models.substr(gt + 1, end - gt - 1);
      std::string ts, vs;
      if (!attr(tag, "type", ts) || !attr(tag, "value", vs)) return fail(GDEM_ERR_ARG);
      const int32_t kind = kind_of_xml((int32_t)strtol(ts.c_str(), nullptr, 10));
      const long long key = strtoll(vs.c_str(), nullptr, 10);
      if (kind == 0 || key <= 0) return fail(GDEM_ERR_ARG);
      int32_t a = -1, b = -1;
      for (int32_t k = 0; k < 31; ++k)
        if (key & (1ll << k)) (a < 0 ? a : b) = k;
      if (b < 0) b = a;
      const std::vector<std::string>& names = param_names(kind);
      double p[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
      for (size_t k = 0; k < names.size(); ++k) {
        size_t pa = 0;
        std::string pv;
        const std::string open = "<Parameter name=\"" + names[k] + "\">";
        if (!between(body, pa, open, "</Parameter>", pv)) return fail(GDEM_ERR_ARG);
        p[k] = strtod(pv.c_str(), nullptr);
      }
      rc = gdem_set_contact_model(c, a, b, kind, p);
      if (rc) return fail(rc);
      m = end;
    }
  }
  // walls (obstacle k: id wall_id_offset + k in the contact memory)
  for (size_t k = 0; k < obst.size(); ++k) {
    const Obst& o = obst[k];
    const int32_t mi = mat_index(o.material);
    if (mi < 0) return fail(GDEM_ERR_ARG);
    rc = gdem_add_wall(c, o.basis, o.origin, o.half, o.crust, mi);
    if (rc) return fail(rc);
    gdem_set_wall_name(c, (int32_t)k, o.name.c_str());
  }
  // particle records: types give the constants, the records the state
  std::string pfile = mainS;
  if (pfile.size() > 7 && pfile.compare(pfile.size() - 7, 7, ".result") == 0)
    pfile.erase(pfile.size() - 7);
  pfile += std::string("_Particles.result") + (binary ? ".bin" : "");
  std::vector<Record> recs;
  {
    FILE* f = fopen(pfile.c_str(), binary ? "rb" : "r");
    if (!f) return fail(GDEM_ERR_ARG);
    for (;;) {
      Record r;
      bool bad = false;
      const bool got = binary ? read_binary(f, r, bad) : read_text(f, r, bad);
      if (bad) {
        fclose(f);
        return fail(GDEM_ERR_ARG);
      }
      if (!got) break;
      recs.push_back(std::move(r));
    }
    fclose(f);
  }
  const size_t n = recs.size();
  if ((long long)n != npart) return fail(GDEM_ERR_ARG);
  std::vector<double> x(3 * n), v(3 * n), w(3 * n), q(4 * n), radius(n), density(n), crust(n);
  std::vector<int32_t> material(n);
  std::vector<char> seen(n, 0);
  for (const Record& r : recs) {
    const long long i = (long long)r.id - id_offset;
    if (i < 0 || i >= (long long)n || seen[(size_t)i]) return fail(GDEM_ERR_ARG);
    seen[(size_t)i] = 1;
    const RefType* ty = nullptr;
    for (const RefType& p : types)
      if (p.type == r.type) ty = &p;
    if (!ty) return fail(GDEM_ERR_ARG);
    const int32_t mi = mat_index(ty->material);
    if (mi < 0) return fail(GDEM_ERR_ARG);
    radius[i] = ty->r;
    density[i] = ty->rho;
    crust[i] = ty->crust;
    material[i] = mi;
    for (int d = 0; d < 3; ++d) {
      x[3 * i + d] = r.origin[d];
      v[3 * i + d] = r.v[d];
      w[3 * i + d] = r.w[d];
    }
    q[4 * i] = r.q[1];
    q[4 * i + 1] = r.q[2];
    q[4 * i + 2] = r.q[3];
    q[4 * i + 3] = r.q[0];
  }
  rc = gdem_set_particles(c, (int64_t)n, x.data(), v.data(), w.data(), q.data(), radius.data(),
                          density.data(), crust.data(), material.data());
  if (rc) return fail(rc);
  // state, the leap-frog's F and T and the contact memory from the records
  rc = gdem_read_particles_2014(c, pfile.c_str(), binary ? 1 : 0, id_offset, wall_id_offset,
                                nullptr);
  if (rc) return fail(rc);
  if (time) *time = tm;
  *out = c;
  return GDEM_OK;
}
