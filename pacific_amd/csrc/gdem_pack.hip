// gdem_pack.hip -- the sphere step in the packed layout (skin candidate
// lists, memoryless contact models).  A step that keeps its lists runs:
//
//   k_bin_screen      : one thread per particle walks the candidates it
//                       detected (each pair sits in exactly one out-list) and
//                       screens them against the positions of the step; the
//                       candidates that may touch go to a queue.
//   k_scatter_contact : one lane per queued candidate evaluates the pair once
//                       with the reference's body order (RigidBodyWithCrust.cpp:
//                       596-640, HookeContactForceModel.cpp:44-114 /
//                       HertzContactForceModel.cpp:46-114) and stores its
//                       record (force on the first body and both torques) at
//                       the out-list entry; it flags the entry in the
//                       particle's out mask and the mirror entry in the
//                       partner's in mask (integer OR).
//   k_integ           : one thread per particle sums its flagged records in a
//                       fixed order -- detected contacts by entry, then
//                       received contacts by partner slot -- (Torsor.cpp:
//                       184-189, no floating-point atomics), adds the walls,
//                       and applies the leap-frog tail (SecondOrderLeapFrog.cpp:
//                       44-94 via ParticleKinematics.cpp:62-128), writing the
//                       other state copy.
//
// A step that rebuilds its lists runs the counting sort by cell in the same
// first kernels (count pass in k_bin_screen, scan, scatter pass in
// k_scatter_contact, per-cell sort), then k_sweep_contact sweeps the half
// stencil (LinkedCell.cpp:1019-1103), screens the new lists and evaluates the
// pairs, and k_integ.  The gate word of the step picks the branch on the
// device, so one captured graph serves both kinds of step.
//
// The per-pair arithmetic and the order of every sum are those of the fused
// kernel (gdem_sphere_step.h), so both give bit-identical results.  Contact
// statistics need fixed-order sums over the evaluations, which the queue does
// not give; with statistics on, gdem_step runs the canonical layout.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gdem_binning.h"
#include "gdem_internal.h"
#include "gdem_sphere_step.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

constexpr int kPB = 256;

static inline int pblocks(int n) { return (n + kPB - 1) / kPB; }

// ---------------------------------------------------------------- layout
// World angular velocity of slot s of copy `in` where it is not stored: the
// step that wrote the copy computed it as q (omega_body) q* with the
// orientation q from before its update (SecondOrderLeapFrog.cpp:74-94, then
// ParticleKinematics.cpp:62-104), which the other copy still holds at slot
// so.  The same operations in the same order: the value is bit-identical.
__device__ __forceinline__ v3 world_omega(const Pack& in, const Pack& other, int s, int so) {
  const double4 d = in.D[s];
  const quat q = qmk(mk(other.A[so].w, other.B[so].w, other.D[so].w), other.E[so]);
  return qmulv(q, qleftvec(qconj(q), mk(d.x, d.y, d.z)));
}

__global__ void __launch_bounds__(kPB) k_to_pack(State s, Pack pk, PackConst k0, PackConst k1,
                                                 int32_t* wsrc, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *wsrc = 0;  // the world angular velocity is stored
  if (i >= n) return;
  const double4 pr = s.pos_r[i], vm = s.vel_m[i], wr = s.om_req[i], wb = s.wbf_I[i],
                q = s.quat[i];
  pk.A[i] = make_double4(pr.x, pr.y, pr.z, q.x);
  pk.B[i] = make_double4(vm.x, vm.y, vm.z, q.y);
  pk.C[i] = make_double4(wr.x, wr.y, wr.z, 0.);
  pk.D[i] = make_double4(wb.x, wb.y, wb.z, q.z);
  pk.E[i] = q.w;
  const double2 rr = make_double2(pr.w, wr.w), mi = make_double2(vm.w, wb.w);
  const Aux ax = s.aux[i];
  k0.rr[i] = rr;
  k0.mi[i] = mi;
  k0.aux[i] = ax;
  k1.rr[i] = rr;
  k1.mi[i] = mi;
  k1.aux[i] = ax;
}

__global__ void __launch_bounds__(kPB) k_from_pack(Pack pk, Pack other, PackConst k, State s,
                                                   const int32_t* wsrc, const int32_t* perm,
                                                   const SkinWords* words, int p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 a = pk.A[i], b = pk.B[i], d = pk.D[i];
  const double2 rr = k.rr[i], mi = k.mi[i];
  v3 w;
  if (*wsrc != 0) {
    w = world_omega(pk, other, i, words->rebuilt[p] ? perm[i] : i);
  } else {
    const double4 c = pk.C[i];
    w = mk(c.x, c.y, c.z);
  }
  s.pos_r[i] = make_double4(a.x, a.y, a.z, rr.x);
  s.vel_m[i] = make_double4(b.x, b.y, b.z, mi.x);
  s.om_req[i] = make_double4(w.x, w.y, w.z, rr.y);
  s.wbf_I[i] = make_double4(d.x, d.y, d.z, mi.y);
  s.quat[i] = make_double4(a.w, b.w, d.w, pk.E[i]);
  s.aux[i] = k.aux[i];
}

// ---------------------------------------------------------------- pairs
// Screening of the candidates of particle t (slot order after the binning of
// a rebuild; the state is read through perm then): the out entries whose
// partner is not clearly apart (apart()), as a bit mask.  Partner positions
// are loaded four at a time with independent loads.
__device__ __forceinline__ uint32_t screen_particle(const PairArgs& a, int t, bool rebuild) {
  const DetectArgs& d = a.d;
  const PosSrc src{a.in.A, a.kin.rr, a.kin.aux};
  int nout = rebuild ? detect_particle(d, src, t) : d.out_count[t];
  nout = min(nout, kListCap);
  const int s = rebuild ? d.perm[t] : t;
  const double4 me = src.pos(s);
  const v3 X = mk(me.x, me.y, me.z);
  const double r = me.w;
  const int2* ol = d.out_list + t;
  uint32_t may = 0;
  for (int e0 = 0; e0 < nout; e0 += 4) {
    int pp[4];
    double4 q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) pp[k] = e0 + k < nout ? ol[(size_t)(e0 + k) * d.lstride].x : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (pp[k] >= 0) q[k] = src.pos(rebuild ? d.perm[pp[k]] : pp[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (pp[k] >= 0 && !apart(d.grid, X, r, q[k])) may |= 1u << (e0 + k);
  }
  return may;
}

// One queued candidate: out entry e of particle t, evaluated once with the
// reference's body order (RigidBodyWithCrust.cpp:596-640, the contact-error
// retry of :534-588, HookeContactForceModel.cpp:44-114 /
// HertzContactForceModel.cpp:46-114).  A contact stores its record at the
// entry and flags the entry in t's out mask and the mirror entry in the
// partner's in mask (integer OR).  Returns 1 for a contact.
__device__ __forceinline__ int contact_pair(const PairArgs& a, int t, int e, bool rebuild,
                                            bool wderived) {
  const DetectArgs& d = a.d;
  const PosSrc src{a.in.A, a.kin.rr, a.kin.aux};
  const int s = rebuild ? d.perm[t] : t;
  const double4 me = src.pos(s);
  const v3 X = mk(me.x, me.y, me.z);
  const double r = me.w;
  const int2 en = d.out_list[(size_t)e * d.lstride + t];
  const int u = en.x, code = en.y;
  const int su = rebuild ? d.perm[u] : u;
  const double4 p2 = src.pos(su);
  const double4 b = a.in.B[s];
  const double2 rr = a.kin.rr[s], mi = a.kin.mi[s];
  const Aux ax = a.kin.aux[s];
  const double4 b2 = a.in.B[su];
  const double2 rr2 = a.kin.rr[su], mi2 = a.kin.mi[su];
  const Aux ax2 = a.kin.aux[su];
  const v3 V = mk(b.x, b.y, b.z), V2 = mk(b2.x, b2.y, b2.z);
  v3 W, W2;
  if (wderived) {
    // the copy's world angular velocity is not stored (never on a rebuild
    // step: the first pair kernel of a rebuild stores it); after a rebuild
    // the other copy is in the previous order (perm)
    const bool moved = a.d.gate.w->rebuilt[a.d.gate.p] != 0;
    W = world_omega(a.in, a.other, s, moved ? d.perm[s] : s);
    W2 = world_omega(a.in, a.other, su, moved ? d.perm[su] : su);
  } else {
    const double4 c4 = a.in.C[s], c2 = a.in.C[su];
    W = mk(c4.x, c4.y, c4.z);
    W2 = mk(c2.x, c2.y, c2.z);
  }
  // either body may have crossed a periodic face since the list was built:
  // the partner is taken at its nearest image
  const int sx = image_shift(d.grid, 0, p2.x - X.x);
  const int sy = image_shift(d.grid, 1, p2.y - X.y);
  const int sz = image_shift(d.grid, 2, p2.z - X.z);
  const double shx = sx * d.grid.size[0], shy = sy * d.grid.size[1], shz = sz * d.grid.size[2];
  const bool first = (code & 64) != 0;
  v3 A, B, cp, ov;
  double rA, rB, crA, crB, dist;
  if (first) {
    A = X; B = mk(p2.x + shx, p2.y + shy, p2.z + shz);
    rA = r; rB = p2.w; crA = ax.crust; crB = ax2.crust;
  } else {
    A = mk(p2.x, p2.y, p2.z);
    B = mk(X.x + (-shx), X.y + (-shy), X.z + (-shz));
    rA = p2.w; rB = r; crA = ax2.crust; crB = ax.crust;
  }
  int rc = sphere_sphere(A, rA, crA, B, rB, crB, cp, ov, dist);
  int nbIter = 1;
  if (rc < 0) rc = sphere_sphere_retry(A, rA, crA, B, rB, crB, cp, ov, dist, nbIter);
  if (rc < 0) {
    atomicMin(&d.err[kErrContact], first ? pack_ids(ax.id, ax2.id) : pack_ids(ax2.id, ax.id));
    return 0;
  }
  if (rc == 0 || !(dist < 0.)) return 0;
  const Model md = a.models[(first ? ax.mat : ax2.mat) * a.model_stride + (first ? ax2.mat : ax.mat)];
  const double m = mi.x, req = rr.y;
  const v3 f = first ? contact_force(md, cp, ov, dist, A, V, W, m, req, B, V2, W2, mi2.x, rr2.y)
                     : contact_force(md, cp, ov, dist, A, V2, W2, mi2.x, rr2.y, B, V, W, m, req);
  const v3 ta = cross(sub(cp, A), f);
  const v3 tb = cross(sub(cp, B), neg(f));
  PairRec& R = a.rec[(size_t)e * d.lstride + t];
  R.f[0] = f.x; R.f[1] = f.y; R.f[2] = f.z;
  R.ta[0] = ta.x; R.ta[1] = ta.y; R.ta[2] = ta.z;
  R.tb[0] = tb.x; R.tb[1] = tb.y; R.tb[2] = tb.z;
  atomicOr(&a.omask[t], 1u << e);
  const int fi = (code >> 7) & 31;  // the mirror entry in the partner's in-list
  if (fi < kListCap) atomicOr(&a.imask[u], 1u << fi);
  return 1;
}

__device__ __forceinline__ void count_contacts(unsigned long long* counter, int nc,
                                               unsigned long long* sh) {
  unsigned long long c = (unsigned long long)nc;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kPB / 64; ++w) tot += sh[w];
    if (tot) counter[blockIdx.x] += tot;  // this workgroup's slot: no atomics
  }
}

// Queues the candidates of particle t that may touch (bit mask may) as
// t * 32 + e.  The queue has kQueueSegs segments (workgroup b appends to
// segment b mod kQueueSegs, one integer atomic per workgroup), so no counter
// sees more than a few hundred atomics per step.  Every thread of the
// workgroup calls.
__device__ __forceinline__ void queue_candidates(const PairArgs& a, int t, uint32_t may,
                                                 int* sh_wave, int* sh_base) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int cnt = __builtin_popcount(may);
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) sh_wave[wid] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kPB / 64; ++w) tot += sh_wave[w];
    *sh_base = tot ? atomicAdd(&a.qcount[blockIdx.x % kQueueSegs], tot) : 0;
  }
  __syncthreads();
  int pos = *sh_base + incl - cnt;
  for (int w = 0; w < wid; ++w) pos += sh_wave[w];
  int32_t* seg = a.queue + (size_t)(blockIdx.x % kQueueSegs) * a.qseg;
  for (uint32_t mm = may; mm != 0u; mm &= mm - 1u) {
    if (pos < a.qseg) seg[pos] = t * 32 + __builtin_ctz(mm);
    else atomicMin(&a.d.err[kErrCapacity], (unsigned long long)(uint32_t)t);
    ++pos;
  }
}

// The queued candidates, one per lane (grid-stride over the segments'
// lengths); every contact record lands at its fixed entry, so the order in
// which the queue was filled does not matter.
__device__ __forceinline__ void evaluate_queue(const PairArgs& a, bool rebuild, bool wderived) {
  __shared__ unsigned long long sh_cnt[kPB / 64];
  __shared__ int sh_pre[kQueueSegs + 1];
  static_assert(kQueueSegs == 64, "one lane of the first wave per queue segment");
  if (threadIdx.x < 64) {
    // segment lengths loaded by the 64 lanes at once, then a wave scan
    const int len = min(a.qcount[threadIdx.x], a.qseg);
    int incl = len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o, 64);
      if ((int)threadIdx.x >= o) incl += y;
    }
    sh_pre[threadIdx.x] = incl - len;
    if (threadIdx.x == 63) sh_pre[kQueueSegs] = incl;
  }
  __syncthreads();
  const int nq = sh_pre[kQueueSegs];
  int nc = 0;
  for (int q = blockIdx.x * kPB + threadIdx.x; q < nq; q += gridDim.x * kPB) {
    int lo = 0, hi = kQueueSegs;  // the segment k with pre[k] <= q < pre[k + 1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sh_pre[mid] <= q) lo = mid;
      else hi = mid;
    }
    const int w = a.queue[(size_t)lo * a.qseg + (q - sh_pre[lo])];
    nc += contact_pair(a, w >> 5, w & 31, rebuild, wderived);
  }
  count_contacts(a.contacts, nc, sh_cnt);
}

// The kernels of the pair stage.  Each step launches the same five kernels;
// a step that rebuilds the lists (gate word set) runs the counting sort by
// cell (count, scan, scatter, per-cell sort: gdem_binning.h) and then
// k_sweep_contact, one that does not screens its lists in k_bin_screen and
// evaluates the queue in k_scatter_contact, and the other kernels exit at
// once.  Sharing kernels between the two kinds of step keeps the launches
// that do nothing to two per step.

// Count pass of a rebuild, or the screening of the lists: the candidates
// that may touch go to the queue, and the out masks are cleared for the
// contact evaluation.  Block 0 does the skin-word duties of the evaluation.
__global__ void __launch_bounds__(kPB) k_bin_screen(Bins b, PosSrc src, SkinWords* words,
                                                    PairArgs a) {
  __shared__ int sh_wave[kPB / 64];
  __shared__ int sh_base;
  const bool rebuild = a.d.gate.rebuild();
  if (blockIdx.x == 0 && threadIdx.x == 0) skin_words_duty(a.d.gate, words, rebuild);
  if (rebuild) {
    if (*a.wsrc != 0) {
      // the sweep reads the copy through the new order: the world angular
      // velocity is stored first, from the orientations of the other copy at
      // the slots of the previous order (perm still holds it here)
      const bool moved = a.d.gate.w->rebuilt[a.d.gate.p] != 0;
      for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.d.n; i += gridDim.x * blockDim.x) {
        const v3 w = world_omega(a.in, a.other, i, moved ? a.d.perm[i] : i);
        a.in.C[i] = make_double4(w.x, w.y, w.z, 0.);
      }
    }
    count_pass(b.key, b.slot, b.cnt, a.d.n, true, a.d.in_count, src, a.d.grid, a.d.row_rank,
               a.d.err);
    return;
  }
  const int t = logical_index(a.d.nblk8);
  uint32_t may = 0;
  if (t < a.d.n) {
    may = screen_particle(a, t, false);
    a.omask[t] = 0u;
  }
  queue_candidates(a, t, may, sh_wave, &sh_base);
}

// Scatter pass of a rebuild, or the evaluation of the queued candidates.
__global__ void __launch_bounds__(kPB) k_scatter_contact(Bins b, PairArgs a) {
  if (a.d.gate.rebuild()) {
    // the first kernel stored the world angular velocity (no kernel before
    // this one reads the word after it)
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.wsrc = 0;
    scatter_pass(b.key, b.slot, b.local, b.bsum, b.perm, a.d.n);
    return;
  }
  evaluate_queue(a, false, *a.wsrc != 0);
}

// Rebuild steps only: the candidate sweep (LinkedCell.cpp:1019-1103) and the
// screening of the new lists; each workgroup compacts the candidates that may
// touch into a queue in LDS (kListCap entries per particle at most) and
// evaluates them, one per lane.
__global__ void __launch_bounds__(kPB) k_sweep_contact(PairArgs a) {
  __shared__ int sh_wave[kPB / 64];
  __shared__ int32_t sh_q[kPB * kListCap];
  __shared__ unsigned long long sh_cnt[kPB / 64];
  if (!a.d.gate.rebuild()) return;
  const int t = logical_index(a.d.nblk8);
  uint32_t may = 0;
  if (t < a.d.n) {
    may = screen_particle(a, t, true);
    a.omask[t] = 0u;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int cnt = __builtin_popcount(may);
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) sh_wave[wid] = incl;
  __syncthreads();
  int pos = incl - cnt, nq = 0;
#pragma unroll
  for (int w = 0; w < kPB / 64; ++w) {
    if (w < wid) pos += sh_wave[w];
    nq += sh_wave[w];
  }
  for (uint32_t mm = may; mm != 0u; mm &= mm - 1u) sh_q[pos++] = t * 32 + __builtin_ctz(mm);
  __syncthreads();
  int nc = 0;
  for (int q = threadIdx.x; q < nq; q += kPB) {
    const int w = sh_q[q];
    nc += contact_pair(a, w >> 5, w & 31, true, false);
  }
  count_contacts(a.contacts, nc, sh_cnt);
}

// ---------------------------------------------------------------- integrate
// Adds one stored pair record to the body it belongs to.
__device__ __forceinline__ void add_record(const PairRec& R, bool first, v3& F, v3& T) {
  const v3 f = mk(R.f[0], R.f[1], R.f[2]);
  if (first) {
    F = add(F, f);
    T = add(T, mk(R.ta[0], R.ta[1], R.ta[2]));
  } else {
    F = add(F, neg(f));
    T = add(T, mk(R.tb[0], R.tb[1], R.tb[2]));
  }
}

// Periodic wrap of coordinate x along axis d after the drift; false when x
// may have left the grid along a closed axis (Cell::GetCell then decides).
__device__ __forceinline__ bool wrap_axis(const IntegArgs& a, int d, double& x) {
  if (a.grid.periodic[d]) {
    const double lo = a.grid.origin[d], L = a.grid.size[d];
    if (x >= lo + L) x -= L;
    else if (x < lo) x += L;
    return true;
  }
  const double u = x - a.grid.origin[d];
  return u >= 0. && u <= a.gmax[d];
}

// Force, torque and leap-frog tail of particle t (its input at slot s).
// Returns its wall contacts; drift = length of its displacement.
template <int MODE, bool WALLS>
__device__ __forceinline__ int integ_particle(const IntegArgs& a, int t, int s, bool copy_const,
                                              double& drift) {
  const double4 A4 = a.in.A[s], B4 = a.in.B[s], D4 = a.in.D[s];
  const double qw = a.in.E[s];
  const double2 mi = a.kin.mi[s];
  const uint32_t om = a.omask[t];
  const uint32_t im = a.imask[t];
  const v3 X = mk(A4.x, A4.y, A4.z), V = mk(B4.x, B4.y, B4.z);
  // the world angular velocity is read only where it is used (wall
  // contacts, the force-only evaluation, which finds it stored)
  v3 W = mk(0., 0., 0.);
  if (WALLS || MODE == MODE_FORCE) {
    const double4 C4 = a.in.C[s];
    W = mk(C4.x, C4.y, C4.z);
  }
  v3 WBF = mk(D4.x, D4.y, D4.z);
  quat q = qmk(mk(A4.w, B4.w, D4.w), qw);
  const double m = mi.x, I = mi.y;
  // InitializeForce: weight m (1 - 0/rho) g at the centre (Particle.cpp:414-419, 1409-1416)
  const double mw = m * (1. - 0. / 1.);
  v3 F = mk(a.gravity[0] * mw, a.gravity[1] * mw, a.gravity[2] * mw);
  v3 T = mk(0., 0., 0.);
  // contact segment: detected contacts by entry, then received contacts in
  // the order (partner slot, code, position)
  for (uint32_t mm = om; mm != 0u; mm &= mm - 1u) {
    const size_t at = (size_t)__builtin_ctz(mm) * a.lstride + t;
    add_record(a.rec[at], (a.out_list[at].y & 64) != 0, F, T);
  }
  if (im != 0u) {
    a.imask[t] = 0u;  // ready for the next pair kernel
    const int nin = min(a.in_count[t], kListCap);
    const int2* il = a.in_list + t;
    long long last = -1;
    const int nlive = __builtin_popcount(im);
    for (int k = 0; k < nlive; ++k) {
      long long key;
      const int f = next_incoming(il, nin, a.lstride, last, key, im);
      if (f < 0) break;
      const int2 en = il[(size_t)f * a.lstride];
      const int e = (en.y >> 7) & 31;
      if (e < kListCap) add_record(a.rec[(size_t)e * a.lstride + en.x], (en.y & 64) != 0, F, T);
      last = key;
    }
  }
  int nwc = 0;
  if (WALLS) {
    // SimpleObstacle::InterAction (Component/src/SimpleObstacle.cpp:222-271)
    const double2 rr = a.kin.rr[s];
    const Aux ax = a.kin.aux[s];
    for (int wi = 0; wi < a.nwalls; ++wi) {
      const Wall& wl = a.walls[wi];
      if (!(fabs(wl.bb_centre[0] - X.x) <= wl.bb_ext[0] + rr.x &&
            fabs(wl.bb_centre[1] - X.y) <= wl.bb_ext[1] + rr.x &&
            fabs(wl.bb_centre[2] - X.z) <= wl.bb_ext[2] + rr.x))
        continue;
      v3 c = mk(0., 0., 0.), ov = mk(0., 0., 0.);
      double dist;
      int nbIter = 1;
      int rc = wl.kind == WALL_RECTANGLE ? sphere_rect(wl, X, rr.x, c, ov, dist)
                                         : sphere_wall(wl, X, rr.x, ax.crust, c, ov, dist);
      if (rc < 0) rc = sphere_wall_retry(wl, X, rr.x, ax.crust, c, ov, dist, nbIter);
      if (rc < 0) {
        atomicMin(&a.err[kErrContact], pack_ids(ax.id, -1 - wi));
        continue;
      }
      if (rc == 0 || !(dist < 0.)) continue;
      const Model md = a.models[ax.mat * a.model_stride + wl.material];
      // the wall's velocity at the contact point (Obstacle.cpp:128-133)
      const v3 wo = mk(wl.origin[0], wl.origin[1], wl.origin[2]);
      const v3 wv = mk(wl.vel[0], wl.vel[1], wl.vel[2]);
      const v3 ww = mk(wl.omega[0], wl.omega[1], wl.omega[2]);
      const v3 f = contact_force(md, c, ov, dist, X, V, W, m, rr.y, wo, wv, ww, 1.e20, wl.req);
      F = add(F, f);
      T = add(T, cross(sub(c, X), f));
      ++nwc;
    }
  }
  // ---- tail (as force_particle)
  v3 Xo = X, Vo = V, Wo = W;
  drift = 0.;
  if (MODE != MODE_FORCE) {
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
      // ParticleKinematics::Move (:62-104)
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
      // the record (id, crust) is read only when the drift may exceed a crust
      if (depl > a.crust_min) {
        const Aux ax = a.kin.aux[s];
        if (depl > ax.crust) atomicMin(&a.err[kErrMotion], (unsigned long long)(uint32_t)ax.id);
      }
      // every axis is wrapped (no short-circuit)
      const bool in0 = wrap_axis(a, 0, Xo.x);
      const bool in1 = wrap_axis(a, 1, Xo.y);
      const bool in2 = wrap_axis(a, 2, Xo.z);
      const bool inside = in0 && in1 && in2;
      if (!inside) {
        int ijk[3];
        if (!cell_of(a.grid, Xo, ijk))
          atomicMin(&a.err[kErrSim], (unsigned long long)(uint32_t)a.kin.aux[s].id);
      }
    }
  }
  // streaming stores: the new copy would otherwise sit dirty in the L2s
  // until the kernel's end (measured: 5 us per step less at 1M particles)
  typedef double nd4 __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store((nd4){Xo.x, Xo.y, Xo.z, q.v.x}, reinterpret_cast<nd4*>(a.out.A + t));
  __builtin_nontemporal_store((nd4){Vo.x, Vo.y, Vo.z, q.v.y}, reinterpret_cast<nd4*>(a.out.B + t));
  __builtin_nontemporal_store((nd4){WBF.x, WBF.y, WBF.z, q.v.z}, reinterpret_cast<nd4*>(a.out.D + t));
  __builtin_nontemporal_store(q.w, a.out.E + t);
  // a step of the step sequence without walls leaves the world angular
  // velocity to be derived (world_omega): 24 B per particle neither read
  // nor written here
  if (MODE != MODE_MOVE || WALLS)
    __builtin_nontemporal_store((nd4){Wo.x, Wo.y, Wo.z, 0.}, reinterpret_cast<nd4*>(a.out.C + t));
  if (copy_const) {
    // a rebuild reorders the constants; the evaluation after it fills the
    // other copy
    a.kout.rr[t] = a.kin.rr[s];
    a.kout.mi[t] = mi;
    a.kout.aux[t] = a.kin.aux[s];
  }
  if (a.F) {
    a.F[t] = make_double4(F.x, F.y, F.z, 0.);
    a.T[t] = make_double4(T.x, T.y, T.z, 0.);
  }
  return nwc;
}

template <int MODE, bool WALLS>
__global__ void __launch_bounds__(kPB) k_integ(IntegArgs a) {
  __shared__ unsigned long long sh_cnt[kPB / 64];
  const int t = logical_index(a.nblk8);
  const int p = a.gate.p;
  const bool rebuild = a.gate.rebuild();
  const bool copy_const = rebuild || a.words->rebuilt[p] != 0;
  double drift = 0.;
  int nwc = 0;
  if (t < a.n) {
    nwc = integ_particle<MODE, WALLS>(a, t, rebuild ? a.perm[t] : t, copy_const, drift);
    // path length since the rebuild; need word of the next evaluation (SkinWords)
    if (MODE == MODE_MOVE || rebuild) {
      const float len = skin_path(rebuild ? 0.f : a.path[t], rebuild, drift);
      a.path[t] = len;
      if (len > a.gate.limit) atomicOr(&a.words->need[1 - p], 1);
    }
  }
  if (t == 0) {
    a.words->rebuilt[1 - p] = rebuild ? 1 : 0;
    *a.wsrc = (MODE == MODE_MOVE && !WALLS) ? 1 : 0;
  }
  // the queue of the next screening starts empty
  if (blockIdx.x == 0 && threadIdx.x < kQueueSegs) a.qcount[threadIdx.x] = 0;
  if (WALLS) count_contacts(a.contacts, nwc, sh_cnt);
}

// Kick + drift of copy pk in place (SecondOrderLeapFrog::Move via
// ParticleKinematics::Move), from the stored F and T (the opening half of a
// step sequence).
__global__ void __launch_bounds__(kPB) k_move_pack(Pack pk, PackConst k, const double4* Fa,
                                                   const double4* Ta, int n, Grid g, double dtv,
                                                   double dtd, unsigned long long* err,
                                                   SkinGate gate, float* path, int32_t* wsrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double depl = 0.;
  if (i == 0) *wsrc = 0;  // stored below
  if (i < n) {
    const double4 A4 = pk.A[i], B4 = pk.B[i], D4 = pk.D[i];
    const double2 mi = k.mi[i];
    const double4 f4 = Fa[i], t4 = Ta[i];
    const v3 F = mk(f4.x, f4.y, f4.z), T = mk(t4.x, t4.y, t4.z);
    quat q = qmk(mk(A4.w, B4.w, D4.w), pk.E[i]);
    const quat pc = qconj(q);
    const v3 tq = qmulv(pc, qleftvec(q, T));
    const double mcf = mi.x * 1.;
    const v3 V = add(mk(B4.x, B4.y, B4.z), scl(dvd(F, mcf), dtv));
    const v3 dx = scl(V, dtd);
    const v3 WBF = ssprk3_sphere(tq, mk(D4.x, D4.y, D4.z), mi.y, dtv);
    const v3 W = qmulv(q, qleftvec(pc, WBF));
    const v3 avg = W;
    v3 X = add(mk(A4.x, A4.y, A4.z), dx);
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
    depl = norm(dx);
    const Aux ax = k.aux[i];
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
    if (!cell_of(g, X, ijk)) atomicMin(&err[kErrSim], (unsigned long long)(uint32_t)ax.id);
    pk.A[i] = make_double4(X.x, X.y, X.z, q.v.x);
    pk.B[i] = make_double4(V.x, V.y, V.z, q.v.y);
    pk.C[i] = make_double4(W.x, W.y, W.z, 0.);
    pk.D[i] = make_double4(WBF.x, WBF.y, WBF.z, q.v.z);
    pk.E[i] = q.w;
    // a move in place of copy p: the next evaluation reads need[p]
    const float len = skin_path(path[i], false, depl);
    path[i] = len;
    if (len > gate.limit) atomicOr(const_cast<int32_t*>(&gate.w->need[gate.p]), 1);
  }
}

// ---------------------------------------------------------------- launchers
void launch_to_pack(const State& s, const Pack& pk, const PackConst& k0, const PackConst& k1,
                    int32_t* wsrc, int32_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_pack, dim3(pblocks(std::max(n, 1))), dim3(kPB), 0, st, s, pk, k0, k1,
                     wsrc, n);
}

void launch_from_pack(const Pack& pk, const Pack& other, const PackConst& k, const State& s,
                      const int32_t* wsrc, const int32_t* perm, const SkinWords* words, int p,
                      int32_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_from_pack, dim3(pblocks(n)), dim3(kPB), 0, st, pk, other, k, s, wsrc, perm,
                     words, p, n);
}

void launch_pair(const Bins& b, const PosSrc& src, SkinWords* words, const PairArgs& a,
                 hipStream_t st, hipEvent_t start, hipEvent_t stop) {
  if (a.d.n <= 0) return;
  const dim3 grid(a.d.nblk8), block(kPB);
  // the queue holds a few candidates per hundred particles: an eighth of the
  // particles' grid covers it in one pass (grid-stride beyond); the scatter
  // pass strides over the particles
  const int cblocks = std::max(8, (a.d.nblk8 / 8 + 7) / 8 * 8);
  const dim3 cgrid(std::max(cblocks, std::min(pblocks(a.d.n), 1024)));
  const int32_t ncell = a.d.grid.ncell;
  if (start) hipExtLaunchKernelGGL(k_bin_screen, grid, block, 0, st, start, nullptr, 0, b, src, words, a);
  else hipLaunchKernelGGL(k_bin_screen, grid, block, 0, st, b, src, words, a);
  launch_scan(b, ncell, a.d.gate, st);
  hipLaunchKernelGGL(k_scatter_contact, cgrid, block, 0, st, b, a);
  launch_cellsort(b, src, ncell, a.d.n, a.d.gate, st);
  if (start) hipExtLaunchKernelGGL(k_sweep_contact, grid, block, 0, st, nullptr, stop, 0, a);
  else hipLaunchKernelGGL(k_sweep_contact, grid, block, 0, st, a);
}

template <int M, bool WALLS>
static void integ_launch(const IntegArgs& a, hipStream_t st, hipEvent_t start, hipEvent_t stop) {
  const dim3 grid(a.nblk8), block(kPB);
  if (start)
    hipExtLaunchKernelGGL((k_integ<M, WALLS>), grid, block, 0, st, start, stop, 0, a);
  else
    hipLaunchKernelGGL((k_integ<M, WALLS>), grid, block, 0, st, a);
}

template <bool WALLS>
static void integ_mode(int mode, const IntegArgs& a, hipStream_t st, hipEvent_t start,
                       hipEvent_t stop) {
  if (mode == MODE_FORCE) integ_launch<MODE_FORCE, WALLS>(a, st, start, stop);
  else if (mode == MODE_KICK) integ_launch<MODE_KICK, WALLS>(a, st, start, stop);
  else integ_launch<MODE_MOVE, WALLS>(a, st, start, stop);
}

void launch_integ(int mode, const IntegArgs& a, hipStream_t st, hipEvent_t start,
                  hipEvent_t stop) {
  if (a.n <= 0) return;
  if (a.nwalls > 0) integ_mode<true>(mode, a, st, start, stop);
  else integ_mode<false>(mode, a, st, start, stop);
}

void launch_move_pack(const Pack& pk, const PackConst& k, const double4* F, const double4* T,
                      int32_t n, const Grid& g, double dt_vel, double dt_disp,
                      unsigned long long* err, SkinGate gate, float* path, int32_t* wsrc,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_move_pack, dim3(pblocks(std::max(n, 1))), dim3(kPB), 0, st, pk, k, F, T, n,
                     g, dt_vel, dt_disp, err, gate, path, wsrc);
}

}  // namespace gdem
