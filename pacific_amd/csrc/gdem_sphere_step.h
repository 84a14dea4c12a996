// gdem_sphere_step.h -- per-particle bodies of the sphere detection and force
// kernels, shared by the single-system kernels (gdem_kernels.hip) and the
// batched persistent kernel (gdem_batch.hip), so both compute the same
// operations in the same order.
#pragma once
#include <hip/hip_runtime.h>

#include "gdem_internal.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

// ---------------------------------------------------------------- detect
struct Detector {
  const DetectArgs& a;
  PosSrc src;  // full-precision positions and records in storage order
  int t;
  float4 mef;
  bool have_me;
  v3 X;
  double r;
  Aux ax;
  int nout;

  __device__ __forceinline__ Detector(const DetectArgs& a_, const PosSrc& src_, int t_, float4 mef_)
      : a(a_), src(src_), t(t_), mef(mef_), have_me(false), X{}, r(0.), ax{}, nout(0) {}

  // Appends partner u (entry code) to this particle's out-list and the
  // mirror entry to u's in-list; id and id2 name the two in a capacity error.
  __device__ __forceinline__ void append(int u, int code, int32_t id, int32_t id2) {
    // each entry also records where its mirror sits in the partner's other
    // list (bits 7-11; an index at or beyond the capacity is never used)
    const int kin = atomicAdd(&a.in_count[u], 1);
    if (nout < kListCap) {
      a.out_list[(size_t)nout * a.lstride + t] = make_int2(u, code | (min(kin, 31) << 7));
    } else {
      atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)id);
    }
    if (kin < kListCap) {
      a.in_list[(size_t)kin * a.lstride + u] = make_int2(t, mirror_code(code) | (min(nout, 31) << 7));
    } else {
      atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)id2);
    }
    ++nout;
  }

  // One candidate u with periodic shift s (applied to u).  Cross-cell
  // candidates of the half stencil have this particle as the first body; a
  // same-cell pair is ordered by id.
  __device__ __forceinline__ void test(const float4 p2f, int u, int sx, int sy, int sz, bool same) {
    // single-precision pre-filter: rejects only pairs whose gap exceeds the
    // margin, which bounds the rounding of the float copy
    const float dxf = (p2f.x + sx * a.sizef[0]) - mef.x;
    const float dyf = (p2f.y + sy * a.sizef[1]) - mef.y;
    const float dzf = (p2f.z + sz * a.sizef[2]) - mef.z;
    const float rsf = mef.w + p2f.w + (float)a.skin + a.margin;
    if (dxf * dxf + dyf * dyf + dzf * dzf > rsf * rsf) return;
    if (a.skin > 0.) {
      // candidate: the pre-filter alone decides (every pair with gap below
      // the skin passes it, and the few within the margin beyond); the
      // contact test of every step is made on the candidates with the same
      // body order.  The in-list positions are taken afterwards (link()).
      const int32_t id = a.ids[t];
      const int code = pack_code(sx, sy, sz, same ? (id < a.ids[u]) : true);
      if (nout < kListCap) a.out_list[(size_t)nout * a.lstride + t] = make_int2(u, code);
      else atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)id);
      ++nout;
      return;
    }
    if (!have_me) {
      const int s = a.perm[t];
      const double4 me = src.pos(s);
      X = mk(me.x, me.y, me.z);
      r = me.w;
      ax = src.aux[s];
      have_me = true;
    }
    const int s2 = a.perm[u];
    const double4 p2 = src.pos(s2);
    const Aux ax2 = src.aux[s2];
    const double shx = sx * a.grid.size[0], shy = sy * a.grid.size[1], shz = sz * a.grid.size[2];
    const v3 Bn = mk(p2.x + shx, p2.y + shy, p2.z + shz);
    const bool first = same ? (ax.id < ax2.id) : true;
    v3 A, B, c, ov;
    double rA, rB, crA, crB, dist;
    if (first) {
      A = X; B = Bn; rA = r; rB = p2.w; crA = ax.crust; crB = ax2.crust;
    } else {
      A = mk(p2.x, p2.y, p2.z);
      B = mk(X.x + (-shx), X.y + (-shy), X.z + (-shz));
      rA = p2.w; rB = r; crA = ax2.crust; crB = ax.crust;
    }
    int rc = sphere_sphere(A, rA, crA, B, rB, crB, c, ov, dist);
    int it;
    if (rc < 0) rc = sphere_sphere_retry(A, rA, crA, B, rB, crB, c, ov, dist, it);
    if (rc < 0) {
      atomicMin(&a.err[kErrContact], first ? pack_ids(ax.id, ax2.id) : pack_ids(ax2.id, ax.id));
      return;
    }
    if (rc == 0 || !(dist < 0.)) return;
    append(u, pack_code(sx, sy, sz, first), ax.id, ax2.id);
  }

  // Skin lists: the in-list position of every out entry the sweep stored
  // (partner and code), from integer atomics issued four at a time, written
  // into the entry (bits 7-11) and with the mirror entry into the partner's
  // in-list.
  __device__ __forceinline__ void link() {
    const int n = min(nout, kListCap);
    for (int e0 = 0; e0 < n; e0 += 4) {
      int2 en[4];
      int kin[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (e0 + k < n) en[k] = a.out_list[(size_t)(e0 + k) * a.lstride + t];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (e0 + k < n) kin[k] = atomicAdd(&a.in_count[en[k].x], 1);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (e0 + k >= n) continue;
        const int e = e0 + k, u = en[k].x, code = en[k].y;
        a.out_list[(size_t)e * a.lstride + t] = make_int2(u, code | (min(kin[k], 31) << 7));
        if (kin[k] < kListCap)
          a.in_list[(size_t)kin[k] * a.lstride + u] = make_int2(t, mirror_code(code) | (e << 7));
        else
          atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)a.ids[u]);
      }
    }
  }

  __device__ __forceinline__ void range(int beg, int end, int sx, int sy, int sz, bool same) {
    int u = beg;
    for (; u + 1 < end; u += 2) {
      // issue both loads before either test
      const float4 p0 = a.pos[u], p1 = a.pos[u + 1];
      test(p0, u, sx, sy, sz, same);
      test(p1, u + 1, sx, sy, sz, same);
    }
    if (u < end) test(a.pos[u], u, sx, sy, sz, same);
  }
};

// Detection sweep of one particle (storage slot t of the sorted order), with
// positions from src.
__device__ __forceinline__ int detect_particle(const DetectArgs& a, const PosSrc& src, int t) {
  // the cell index comes from the full-precision position, as in the binning
  const double4 me = src.pos(a.perm[t]);
  int ijk[3];
  clamped_cell(a.grid, mk(me.x, me.y, me.z), ijk);
  Detector det(a, src, t, a.pos[t]);
  half_stencil_sweep(a.grid, a.row_rank, a.cell_start, t, ijk, det);
  if (a.skin > 0.) det.link();
  a.out_count[t] = det.nout;
  return det.nout;
}

__device__ __forceinline__ void detect_particle(const DetectArgs& a, int t) {
  detect_particle(a, PosSrc{a.in.pos_r, nullptr, a.in.aux}, t);
}

// ---------------------------------------------------------------- force
// Image shift (-1, 0, +1 periods) that brings a partner at offset delta from
// a particle nearest to it along axis d.  Pairs of a candidate list are closer
// than one cell edge, at most a third of the period, so the image is unique and
// equals the neighbour-cell shift a fresh sweep assigns.
__device__ __forceinline__ int image_shift(const Grid& g, int d, double delta) {
  if (!g.periodic[d]) return 0;
  const double h = 0.5 * g.size[d];
  return delta > h ? -1 : (delta < -h ? 1 : 0);
}

// Candidate at p2 clearly beyond contact with the sphere (X, r): the margin
// dwarfs the rounding of the exact test (sphere_sphere), which decides every
// other candidate.
__device__ __forceinline__ bool apart(const Grid& g, v3 X, double r, const double4& p2) {
  const int sx = image_shift(g, 0, p2.x - X.x);
  const int sy = image_shift(g, 1, p2.y - X.y);
  const int sz = image_shift(g, 2, p2.z - X.z);
  const double ex = (p2.x + sx * g.size[0]) - X.x;
  const double ey = (p2.y + sy * g.size[1]) - X.y;
  const double ez = (p2.z + sz * g.size[2]) - X.z;
  const double rs = r + p2.w;
  return ex * ex + ey * ey + ez * ez > rs * rs * (1. + 1e-10);
}

// Entries of a candidate list (cnt entries at l[e * a.lstride], partners in
// slot order) that are not clearly apart from the sphere (X, r), as a bit
// mask.  The partner positions are loaded four at a time with independent
// loads.
__device__ __forceinline__ uint32_t near_mask(const ForceArgs& a, v3 X, double r, const int2* l,
                                              int cnt) {
  uint32_t mask = 0;
  for (int e0 = 0; e0 < cnt; e0 += 4) {
    int p[4];
    double4 q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = e0 + k < cnt ? l[(size_t)(e0 + k) * a.lstride].x : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (p[k] >= 0) q[k] = a.in.pos_r[p[k]];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (p[k] >= 0 && !apart(a.grid, X, r, q[k])) mask |= 1u << (e0 + k);
  }
  return mask;
}

// Candidate screening of one particle between rebuilds (skin lists, storage
// slot t; the state stays in slot order): bit e of the low word marks out-list
// entry e, bit e of the high word in-list entry e, when the pair may touch.
// The force kernel walks the marked entries only: most candidates are apart,
// and testing them there would chain two dependent loads per candidate into a
// kernel that holds many registers per lane.
__device__ __forceinline__ void screen_particle(const ForceArgs& a, int t) {
  const double4 pr = a.in.pos_r[t];
  const v3 X = mk(pr.x, pr.y, pr.z);
  const int nout = min(a.out_count[t], kListCap);
  const int nin = min(a.in_count[t], kListCap);
  const uint32_t om = near_mask(a, X, pr.w, a.out_list + t, nout);
  const uint32_t im = near_mask(a, X, pr.w, a.in_list + t, nin);
  a.cmask[t] = (unsigned long long)om | ((unsigned long long)im << 32);
}

template <bool RECORD, bool MEM = false>
struct Reducer {
  const ForceArgs& a;
  int t;
  v3 X, V, W;
  double r, m, req;
  Aux ax;
  v3 F, T;
  int ncont, nrec;
  bool cand;      // list entries are skin candidates: test each for contact
  bool via_perm;  // the input copy is in the previous order (read through perm)
  int s_in;       // input slot of this particle (its memory is read there)
  int nhist;      // memory entries written for this evaluation
  StatAcc st;     // contact statistics (pairs counted by their first body)

  // The memory of the contact with key in this particle's input table
  // (Component::getContactMemory, Component.cpp:634-674, without creating).
  __device__ __forceinline__ bool find_memory(int key, ContactMemory& mem) const {
    const int nh = min(a.hin.count[s_in], kHistCap);
    for (int k = 0; k < nh; ++k) {
      const size_t at = (size_t)k * a.lstride + s_in;
      if (a.hin.key[at] == key) {
        const double4 kd = a.hin.kd[at], nr = a.hin.nrm[at], sp = a.hin.spr[at];
        mem.kd = mk(kd.x, kd.y, kd.z);
        mem.nrm = mk(nr.x, nr.y, nr.z);
        mem.spr = mk(sp.x, sp.y, sp.z);
        return true;
      }
    }
    return false;
  }

  // Records the contact's memory for the next evaluation in this particle's
  // output table: only contacts of this evaluation are kept, as the reference
  // erases the entries no contact touched (Component::updateContactMap,
  // Component.cpp:583-596).
  __device__ __forceinline__ void store_memory(int key, const ContactMemory& mem) {
    if (nhist < kHistCap) {
      const size_t at = (size_t)nhist * a.lstride + t;
      a.hout.key[at] = key;
      a.hout.kd[at] = make_double4(mem.kd.x, mem.kd.y, mem.kd.z, 0.);
      a.hout.nrm[at] = make_double4(mem.nrm.x, mem.nrm.y, mem.nrm.z, 0.);
      a.hout.spr[at] = make_double4(mem.spr.x, mem.spr.y, mem.spr.z, 0.);
    } else {
      atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)ax.id);
    }
    ++nhist;
  }

  // Memory-model force of a pair or wall contact: both partners evaluate it
  // from the first body's view (the second body holds the negated memory,
  // Component::addDeplContactInMap, Component.cpp:680-689), so both get the
  // same force, rolling torque and memory bit for bit.
  __device__ __forceinline__ v3 memory_force(const Model& md, int key, bool first, v3 c, v3 ov,
                                             double dist, v3 A, v3 VA, v3 WA, double mA,
                                             double reqA, v3 B, v3 VB, v3 WB, double mB,
                                             double reqB, v3& delM) {
    ContactMemory mem{mk(0., 0., 0.), mk(0., 0., 0.), mk(0., 0., 0.)};
    const bool found = find_memory(key, mem);
    if (found && !first) mem = ContactMemory{neg(mem.kd), neg(mem.nrm), neg(mem.spr)};
    const v3 f = memory_contact_force(md, c, ov, dist, A, VA, WA, mA, reqA, B, VB, WB, mB, reqB,
                                      a.dt_mem, found, mem, delM);
    if (first)
      store_memory(key, mem);
    else
      store_memory(key, ContactMemory{neg(mem.kd), neg(mem.nrm), neg(mem.spr)});
    return f;
  }

  // One contact of this particle's segment: partner slot p, code (shift to the
  // partner's image, first-body flag).  The pair is evaluated with A = first
  // body and B = second body moved to the first body's image, exactly as the
  // partner's thread does.
  __device__ __forceinline__ void contact(int p, int code) {
    const bool first = (code & 64) != 0;
    const int s2 = via_perm ? a.perm[p] : p;
    const double4 p2 = a.in.pos_r[s2];
    int sx, sy, sz;
    if (cand) {
      // either body may have crossed a periodic face since the list was
      // built, so the stored shift is stale: take the nearest image
      sx = image_shift(a.grid, 0, p2.x - X.x);
      sy = image_shift(a.grid, 1, p2.y - X.y);
      sz = image_shift(a.grid, 2, p2.z - X.z);
    } else {
      sx = (code & 3) - 1; sy = ((code >> 2) & 3) - 1; sz = ((code >> 4) & 3) - 1;
    }
    const double shx = sx * a.grid.size[0], shy = sy * a.grid.size[1], shz = sz * a.grid.size[2];
    if (cand && apart(a.grid, X, r, p2)) return;
    const double4 vm2 = a.in.vel_m[s2];
    const double4 wr2 = a.in.om_req[s2];
    const Aux ax2 = a.in.aux[s2];
    const v3 V2 = mk(vm2.x, vm2.y, vm2.z), W2 = mk(wr2.x, wr2.y, wr2.z);
    v3 A, B, c, ov;
    double rA, rB, crA, crB, dist;
    if (first) {
      A = X; B = mk(p2.x + shx, p2.y + shy, p2.z + shz);
      rA = r; rB = p2.w; crA = ax.crust; crB = ax2.crust;
    } else {
      A = mk(p2.x, p2.y, p2.z);
      B = mk(X.x + (-shx), X.y + (-shy), X.z + (-shz));
      rA = p2.w; rB = r; crA = ax2.crust; crB = ax.crust;
    }
    int rc = sphere_sphere(A, rA, crA, B, rB, crB, c, ov, dist);
    int nbIter = 1;  // ClosestPointSPHERESPHERE reports one iteration
    if (rc < 0) rc = sphere_sphere_retry(A, rA, crA, B, rB, crB, c, ov, dist, nbIter);
    if (rc < 0) {
      atomicMin(&a.err[kErrContact], first ? pack_ids(ax.id, ax2.id) : pack_ids(ax2.id, ax.id));
      return;
    }
    if (rc == 0) return;
    if (first && a.stat) st.add(dist, nbIter);
    if (!(dist < 0.)) return;
    const Model md = a.models[(first ? ax.mat : ax2.mat) * a.model_stride + (first ? ax2.mat : ax.mat)];
    v3 f;
    v3 delM = mk(0., 0., 0.);
    bool rolling = false;
    if (MEM && md.kind >= 3) {
      // memory key: the partner's id (Component.cpp:634, idmap (1, id, 1))
      if (first)
        f = memory_force(md, ax2.id, true, c, ov, dist, A, V, W, m, req, B, V2, W2, vm2.w, wr2.w,
                         delM);
      else
        f = memory_force(md, ax2.id, false, c, ov, dist, A, V2, W2, vm2.w, wr2.w, B, V, W, m,
                         req, delM);
      rolling = md.p[md.kind == 3 ? 5 : 4] != 0.;
    } else {
      if (first)
        f = contact_force(md, c, ov, dist, A, V, W, m, req, B, V2, W2, vm2.w, wr2.w, delM);
      else
        f = contact_force(md, c, ov, dist, A, V2, W2, vm2.w, wr2.w, B, V, W, m, req, delM);
      // addTorque(delM) / addTorque(-delM) (HookeContactForceModel.cpp:150-155)
      rolling = md.p[4] != 0.;
    }
    if (first) {
      F = add(F, f);
      T = add(T, cross(sub(c, A), f));
      if (rolling) T = add(T, delM);  // addTorque(delM) (HookeMemoryContactForceModel.cpp:236)
      ++ncont;
      if (RECORD) record(ax.id, ax2.id, c, ov, dist, f);
    } else {
      const v3 fn = neg(f);
      F = add(F, fn);
      T = add(T, cross(sub(c, B), fn));
      if (rolling) T = add(T, neg(delM));  // addTorque(-delM) (:241)
    }
  }

  __device__ __forceinline__ void record(int id0, int id1, v3 c, v3 ov, double dist, v3 f) {
    if (nrec < a.rec_cap) {
      ContactRec& cr = a.rec[(size_t)t * a.rec_cap + nrec];
      cr.id0 = id0; cr.id1 = id1;
      cr.point[0] = c.x; cr.point[1] = c.y; cr.point[2] = c.z;
      cr.ovec[0] = ov.x; cr.ovec[1] = ov.y; cr.ovec[2] = ov.z;
      cr.dist = dist;
      cr.force[0] = f.x; cr.force[1] = f.y; cr.force[2] = f.z;
    } else {
      atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)id0);
    }
    ++nrec;
  }
};

// The reaction of one particle-wall contact on the wall (the wall is the
// second body: Obstacle::addForce(point, -f) about its centre, then
// addTorque(-delM); Component/src/Obstacle.cpp:575-597, Torsor.cpp:184-189),
// stored as the particle's k-th wall contact for the load reduction.
__device__ __forceinline__ void wall_load(const ForceArgs& a, int t, int32_t id, int& k, int wi,
                                          v3 c, v3 wo, v3 f, v3 mwall) {
  if (k < kWallLoadCap) {
    const size_t at = (size_t)k * a.lstride + t;
    const v3 fw = neg(f);
    const v3 tw = add(cross(sub(c, wo), fw), mwall);
    a.wl_wall[at] = wi;
    a.wl_f[at] = make_double4(fw.x, fw.y, fw.z, 0.);
    a.wl_t[at] = make_double4(tw.x, tw.y, tw.z, 0.);
  } else {
    atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)id);
  }
  ++k;
}

// Force, torque and integrator tail of one particle (storage slot t of the
// sorted order).  Returns the number of contacts it is the first body of;
// drift = length of its displacement (0 without one).
template <int MODE, bool RECORD, bool MEM = false>
__device__ __forceinline__ int force_particle(const ForceArgs& a, int t, double& drift,
                                              StatAcc& stats) {
  // with skin lists the state moves through perm only after a rebuild;
  // otherwise it stays in its slot
  const bool skin = a.gate.w != nullptr;
  const bool rebuilt = skin && a.gate.rebuild();
  const int s = (!skin || rebuilt) ? a.perm[t] : t;
  const double4 pr = a.in.pos_r[s];
  const double4 vm = a.in.vel_m[s];
  const double4 wr = a.in.om_req[s];
  const double4 wb = a.in.wbf_I[s];
  const double4 qq = a.in.quat[s];
  drift = 0.;
  Reducer<RECORD, MEM> R{a, t};
  R.s_in = s;
  R.nhist = 0;
  R.ax = a.in.aux[s];
  R.X = mk(pr.x, pr.y, pr.z);
  R.V = mk(vm.x, vm.y, vm.z);
  R.W = mk(wr.x, wr.y, wr.z);
  R.r = pr.w;
  R.m = vm.w;
  R.req = wr.w;
  // InitializeForce: weight m (1 - 0/rho) g at the centre (Particle.cpp:414-419, 1409-1416)
  const double mw = R.m * (1. - 0. / 1.);
  R.F = mk(a.gravity[0] * mw, a.gravity[1] * mw, a.gravity[2] * mw);
  R.T = mk(0., 0., 0.);
  R.ncont = 0;
  R.nrec = 0;
  R.cand = skin;
  R.via_perm = !skin || rebuilt;
  // contact segment: detected contacts in sweep order, then the contacts
  // other particles detected, by increasing partner slot
  const int nout = min(a.out_count[t], kListCap);
  const int nin = min(a.in_count[t], kListCap);
  if (!skin) a.in_count[t] = 0;  // skin lists keep their entries until the next rebuild
  const int2* ol = a.out_list + t;
  const int2* il = a.in_list + t;
  // candidate lists: only the entries that may touch take part in the walk
  // (the others add nothing), in the same order
  uint32_t omask = (1u << nout) - 1u, imask = (1u << nin) - 1u;
  if (skin && !rebuilt) {
    const unsigned long long cm = a.cmask[t];
    omask = (uint32_t)cm;
    imask = (uint32_t)(cm >> 32);
  }
  for (uint32_t m = omask; m != 0u; m &= m - 1u) {
    const int2 en = ol[(size_t)__builtin_ctz(m) * a.lstride];
    R.contact(en.x, en.y);
  }
  long long last = -1;
  const int nlive = __builtin_popcount(imask);
  for (int e = 0; e < nlive; ++e) {
    long long key;
    const int f = next_incoming(il, nin, a.lstride, last, key, imask);
    if (f < 0) break;
    const int2 en = il[(size_t)f * a.lstride];
    R.contact(en.x, en.y);
    last = key;
  }
  // SimpleObstacle::InterAction (Component/src/SimpleObstacle.cpp:222-271)
  int nwl = 0;
  for (int wi = 0; wi < a.nwalls; ++wi) {
    const Wall& wl = a.walls[wi];
    if (!(fabs(wl.bb_centre[0] - R.X.x) <= wl.bb_ext[0] + R.r &&
          fabs(wl.bb_centre[1] - R.X.y) <= wl.bb_ext[1] + R.r &&
          fabs(wl.bb_centre[2] - R.X.z) <= wl.bb_ext[2] + R.r))
      continue;
    v3 c = mk(0., 0., 0.), ov = mk(0., 0., 0.);
    double dist;
    int nbIter = wl.kind == WALL_RECTANGLE ? 0 : 1;  // ClosestPointRECTANGLE reports none
    int rc = wl.kind == WALL_RECTANGLE ? sphere_rect(wl, R.X, R.r, c, ov, dist)
                                       : sphere_wall(wl, R.X, R.r, R.ax.crust, c, ov, dist);
    if (rc < 0) rc = sphere_wall_retry(wl, R.X, R.r, R.ax.crust, c, ov, dist, nbIter);
    if (rc < 0) {
      atomicMin(&a.err[kErrContact], pack_ids(R.ax.id, -1 - wi));
      continue;
    }
    if (rc == 0 || !(dist < 0.)) continue;
    if (a.stat) R.st.add(dist, nbIter);
    const Model md = a.models[R.ax.mat * a.model_stride + wl.material];
    // the wall's velocity at the contact point is vel + omega x (c - origin)
    // (Obstacle::getVelocityAtPoint, Component/src/Obstacle.cpp:128-133)
    const v3 wo = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
    const v3 wv = mk(wl.vel[0], wl.vel[1], wl.vel[2]);
    const v3 ww = mk(wl.omega[0], wl.omega[1], wl.omega[2]);
    v3 f;
    if (MEM && md.kind >= 3) {
      // the particle is the first body (SimpleObstacle.cpp:262-264); the
      // wall's own copy of the memory is never read back
      v3 delM;
      f = R.memory_force(md, -1 - wi, true, c, ov, dist, R.X, R.V, R.W, R.m, R.req, wo, wv, ww,
                         1.e20, wl.req, delM);
      R.F = add(R.F, f);
      R.T = add(R.T, cross(sub(c, R.X), f));
      const bool rolling = md.p[md.kind == 3 ? 5 : 4] != 0.;
      if (rolling) R.T = add(R.T, delM);
      if (a.wl_count)
        wall_load(a, t, R.ax.id, nwl, wi, c, wo, f, rolling ? neg(delM) : mk(0., 0., 0.));
    } else {
      v3 delM;
      f = contact_force(md, c, ov, dist, R.X, R.V, R.W, R.m, R.req, wo, wv, ww, 1.e20, wl.req,
                        delM);
      R.F = add(R.F, f);
      R.T = add(R.T, cross(sub(c, R.X), f));
      const bool rolling = md.p[4] != 0.;
      if (rolling) R.T = add(R.T, delM);
      if (a.wl_count)
        wall_load(a, t, R.ax.id, nwl, wi, c, wo, f, rolling ? neg(delM) : mk(0., 0., 0.));
    }
    ++R.ncont;
    if (RECORD) R.record(R.ax.id, -1 - wi, c, ov, dist, f);
  }
  stats = R.st;
  if (RECORD) a.rec_count[t] = R.nrec;
  if (a.wl_count) a.wl_count[t] = min(nwl, kWallLoadCap);
  if (MEM) a.hout.count[t] = min(R.nhist, kHistCap);
  const v3 F = R.F, T = R.T, X = R.X, V = R.V, W = R.W;
  const double m = R.m;

  // ---- tail
  v3 Xo = X, Vo = V, Wo = W;
  v3 WBF = mk(wb.x, wb.y, wb.z);
  quat q = qmk(mk(qq.x, qq.y, qq.z), qq.w);
  int ijk[3];
  cell_of(a.grid, X, ijk);
#pragma unroll
  for (int d = 0; d < 3; ++d) ijk[d] = min(max(ijk[d], 0), a.grid.nc[d] - 1);
  int key = row_base(a.grid, a.row_rank, ijk[1], ijk[2]) + ijk[0];
  if (MODE != MODE_FORCE) {
    const double I = wb.w;
    const quat pc = qconj(q);
    // torque in body frame (ParticleKinematics.cpp:115-117)
    const v3 tq = qmulv(pc, qleftvec(q, T));
    const double mcf = m * 1.;
    // closing half kick, SecondOrderLeapFrog::advanceVelocity (:74-94)
    Vo = add(V, scl(dvd(F, mcf), a.dt_vel));
    WBF = ssprk3_sphere(tq, WBF, I, a.dt_vel);
    Wo = qmulv(q, qleftvec(pc, WBF));
    if (MODE == MODE_MOVE) {
      // next step's kick and drift, SecondOrderLeapFrog::Move (:44-67) and
      // ParticleKinematics::Move (:62-104); q and T unchanged -> same tq
      Vo = add(Vo, scl(dvd(F, mcf), a.dt_vel));
      const v3 dx = scl(Vo, a.dt_disp);
      WBF = ssprk3_sphere(tq, WBF, I, a.dt_vel);
      Wo = qmulv(q, qleftvec(pc, WBF));
      const v3 avg = Wo;
      Xo = add(X, dx);
      const double nOmega = norm(avg);
      quat dq;
      if (nOmega > kEps) {
        const double cc = cos(nOmega * a.dt_disp / 2.);
        const double sn = sin(nOmega * a.dt_disp / 2.);
        dq = qmk(scl(avg, sn * 1. / nOmega), cc);
      } else {
        dq = qmk(mk(0., 0., 0.), 1.);
      }
      q = qmul(dq, q);
      const double depl = norm(dx);
      drift = depl;
      if (depl > R.ax.crust)
        atomicMin(&a.err[kErrMotion], (unsigned long long)(uint32_t)R.ax.id);
      double xo[3] = {Xo.x, Xo.y, Xo.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (!a.grid.periodic[d]) continue;
        const double lo = a.grid.origin[d], L = a.grid.size[d];
        if (xo[d] >= lo + L) xo[d] -= L;
        else if (xo[d] < lo) xo[d] += L;
      }
      Xo = mk(xo[0], xo[1], xo[2]);
      int nijk[3];
      if (cell_of(a.grid, Xo, nijk)) {
        key = row_base(a.grid, a.row_rank, nijk[1], nijk[2]) + nijk[0];
      } else {
        key = 0;
        atomicMin(&a.err[kErrSim], (unsigned long long)(uint32_t)R.ax.id);
      }
    }
  }
  a.out.pos_r[t] = make_double4(Xo.x, Xo.y, Xo.z, R.r);
  a.out.vel_m[t] = make_double4(Vo.x, Vo.y, Vo.z, m);
  a.out.om_req[t] = make_double4(Wo.x, Wo.y, Wo.z, R.req);
  a.out.wbf_I[t] = make_double4(WBF.x, WBF.y, WBF.z, wb.w);
  a.out.quat[t] = make_double4(q.v.x, q.v.y, q.v.z, q.w);
  a.out.aux[t] = R.ax;
  a.key_out[t] = key;
  if (a.F) {
    a.F[t] = make_double4(F.x, F.y, F.z, 0.);
    a.T[t] = make_double4(T.x, T.y, T.z, 0.);
  }
  return R.ncont;
}

template <int MODE, bool RECORD, bool MEM = false>
__device__ __forceinline__ int force_particle(const ForceArgs& a, int t, double& drift) {
  StatAcc st;
  return force_particle<MODE, RECORD, MEM>(a, t, drift, st);
}

// SecondOrderLeapFrog::Move (TimeIntegration/src/SecondOrderLeapFrog.cpp:44-67)
// via ParticleKinematics::Move (Component/src/ParticleKinematics.cpp:62-104) of
// the particle in storage slot i, in place; writes its new cell key.
// Returns the length of the displacement.
__device__ __forceinline__ double move_particle(const State& s, const double4* Fa, const double4* Ta,
                                                int i, const Grid& g, const int32_t* row_rank,
                                                double dtv, double dtd, int32_t* key,
                                                unsigned long long* err) {
  const double4 pr = s.pos_r[i], vm = s.vel_m[i], wr = s.om_req[i], wb = s.wbf_I[i],
                qq = s.quat[i];
  const double4 f4 = Fa[i], t4 = Ta[i];
  const Aux ax = s.aux[i];
  const v3 F = mk(f4.x, f4.y, f4.z), T = mk(t4.x, t4.y, t4.z);
  quat q = qmk(mk(qq.x, qq.y, qq.z), qq.w);
  const quat pc = qconj(q);
  const v3 tq = qmulv(pc, qleftvec(q, T));
  const double mcf = vm.w * 1.;
  const v3 V = add(mk(vm.x, vm.y, vm.z), scl(dvd(F, mcf), dtv));
  const v3 dx = scl(V, dtd);
  const v3 WBF = ssprk3_sphere(tq, mk(wb.x, wb.y, wb.z), wb.w, dtv);
  const v3 W = qmulv(q, qleftvec(pc, WBF));
  const v3 avg = W;
  v3 X = add(mk(pr.x, pr.y, pr.z), dx);
  const double nOmega = norm(avg);
  quat dq;
  if (nOmega > kEps) {
    const double cc = cos(nOmega * dtd / 2.);
    const double sn = sin(nOmega * dtd / 2.);
    dq = qmk(scl(avg, sn * 1. / nOmega), cc);
  } else {
    dq = qmk(mk(0., 0., 0.), 1.);
  }
  q = qmul(dq, q);
  const double depl = norm(dx);
  if (depl > ax.crust) atomicMin(&err[kErrMotion], (unsigned long long)(uint32_t)ax.id);
  double xo[3] = {X.x, X.y, X.z};
  for (int d = 0; d < 3; ++d) {
    if (!g.periodic[d]) continue;
    const double lo = g.origin[d], L = g.size[d];
    if (xo[d] >= lo + L) xo[d] -= L;
    else if (xo[d] < lo) xo[d] += L;
  }
  X = mk(xo[0], xo[1], xo[2]);
  int ijk[3];
  if (cell_of(g, X, ijk)) {
    key[i] = row_base(g, row_rank, ijk[1], ijk[2]) + ijk[0];
  } else {
    key[i] = 0;
    atomicMin(&err[kErrSim], (unsigned long long)(uint32_t)ax.id);
  }
  s.pos_r[i] = make_double4(X.x, X.y, X.z, pr.w);
  s.vel_m[i] = make_double4(V.x, V.y, V.z, vm.w);
  s.om_req[i] = make_double4(W.x, W.y, W.z, wr.w);
  s.wbf_I[i] = make_double4(WBF.x, WBF.y, WBF.z, wb.w);
  s.quat[i] = make_double4(q.v.x, q.v.y, q.v.z, q.w);
  return depl;
}


}  // namespace gdem
