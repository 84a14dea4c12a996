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
