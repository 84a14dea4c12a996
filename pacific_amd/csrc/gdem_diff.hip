// gdem_diff.hip -- differentiated DEM step for calibration: the loss of a
// trajectory (mean square position error at observation steps) and its
// adjoint gradient with respect to the Hooke parameters (kn, en, mu) of each
// system and to the initial velocities, for many independent sphere systems in
// one persistent launch.  C ABI: include/gdem.h, gdem_diff_*.
//
// Forward: the kick-drift-kick step of Grains::Simulation
// (Base/src/Grains.cpp:344-364) with the state kept in particle-id order.  For
// spheres the angular kick is w += (T / I) dt/2, which is what
// SecondOrderLeapFrog does through the body frame for an isotropic inertia.
// Checkpoints of (x, v, w, F, T) are stored every K steps.
//
// Adjoint: for each window of K steps (last first) the forward is recomputed
// from its checkpoint, storing the inputs (x', v_half, w_half) of every force
// evaluation; then the steps are reversed.  One step's adjoint is the
// transpose of the kicks and the drift (closed form) and of the force map:
// each contact's Jacobian is formed column by column by evaluating the same
// contact code (gdem_dual.h) with one dual direction seeded, and contracted
// with the force/torque adjoints of both bodies.  The contact set of a step is
// the one the forward found at that step (held fixed; the branch decisions
// are the forward's).  Each particle sums its own contributions in a fixed
// order and the per-system parameter gradient is a fixed-order block
// reduction: results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_dual.h"
#include "gdem_host.h"
#include "gdem_internal.h"
#include "gdem_sphere_step.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

struct DiffSys {
  Grid grid;
  float sizef[3];
  float margin;
  double gravity[3];
  int32_t n, off, cell_off, cs_off, row_off, wall_off, nwalls;
  int32_t substeps;  // steps of dt / substeps per step of dt
  double kn, en, mu, beta, dbeta;  // beta(en) and d beta / d en
};

enum { DIFF_LOSS = 0, DIFF_GRAD = 1, DIFF_RECORD = 2, DIFF_TANGENT = 3 };

struct DiffArgs {
  const DiffSys* sys;
  const int32_t* order;
  int32_t nsys;
  int32_t* next;
  int32_t mode, nsteps, K, nobs;
  const int32_t* obs;  // sorted observation steps (1..nsteps)
  double dt;
  int64_t ntot;
  const double4* prop;  // r, m, I, crust
  Aux* aux;
  const double4* x0;    // initial x, y, z, r
  const double4* v0;
  const double4* w0;
  double4 *x, *v, *w, *F, *T;                 // state, id order
  double4 *ck_x, *ck_v, *ck_w, *ck_F, *ck_T;  // [ncheckpoint][ntot]
  double4 *win_x, *win_v, *win_w;             // [K][ntot]
  const double4* target;                      // [nobs][ntot]
  double4* xobs;                              // [nobs][ntot] (record mode)
  int32_t *key, *slot, *perm, *cnt, *cell_start;
  float4* pos;
  const int32_t* row_rank;
  int2 *out_list, *in_list;
  int32_t *out_count, *in_count;
  const Wall* walls;
  double4 *ax, *av, *aw, *aF, *aT, *aFt, *aTt;  // adjoints, id order
  double4 *gX, *gV, *gW;                        // contact parts of one step's adjoint
  double* loss;    // per system
  double* gtheta;  // per system: d loss / d (kn, en, mu)
  const double* dtheta;  // tangent mode: direction in (kn, en, mu) per system
  double* dloss;         // tangent mode: directional derivative per system
  double4* gv0;    // d loss / d v0, id order
  unsigned long long* err;
};

__device__ __forceinline__ v3 ld3(double4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ double4 st3(v3 a, double w) { return make_double4(a.x, a.y, a.z, w); }


// Tangent (Jacobian-vector product) of one contact: inputs carry tangents
// tin, parameters (kn, beta, mu) tangents tp; returns the tangents of the
// force and torque on A (a_side) or on B.
struct PairTan {
  v3 xA, vA, wA, xB, vB, wB;
};
__device__ __forceinline__ bool pair_jvp(const PairIn& in, const PairTan& tin, double kn, double beta,
                                         double mu, v3 tp, bool a_side, v3& dF, v3& dT) {
  auto dv = [](v3 a, v3 t) { return mk3(Dual{a.x, t.x}, Dual{a.y, t.y}, Dual{a.z, t.z}); };
  HookeParams<Dual> p;
  p.kn = Dual{kn, tp.x};
  p.beta = Dual{beta, tp.y};
  p.mu = Dual{mu, tp.z};
  V3<Dual> FA, TA, FB, TB;
  if (!pair_torsors(dv(in.xA, tin.xA), dv(in.vA, tin.vA), dv(in.wA, tin.wA), in.rA, in.crA, in.mA,
                    dv(in.xB, tin.xB), dv(in.vB, tin.vB), dv(in.wB, tin.wB), in.rB, in.crB, in.mB, p,
                    FA, TA, FB, TB))
    return false;
  const V3<Dual> F = a_side ? FA : FB, T = a_side ? TA : TB;
  dF = mk(F.x.d, F.y.d, F.z.d);
  dT = mk(T.x.d, T.y.d, T.z.d);
  return true;
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* red) {
  // fixed-order tree over the workgroup
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = BS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + o];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

template <int BS>
__device__ __forceinline__ int block_scan_excl_d(int v, int* smem, int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) smem[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int w = (lane < BS / 64) ? smem[lane] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < BS / 64) smem[lane] = w;
  }
  __syncthreads();
  const int prefix = wid > 0 ? smem[wid - 1] : 0;
  total = smem[BS / 64 - 1];
  __syncthreads();
  return prefix + x - v;
}

// TAN: the forward-mode (DIFF_TANGENT) instantiation; the others carry no
// tangent code, which keeps their register allocation to what the gradient
// needs.
template <int BS, bool TAN>
struct DiffRunner {
  const DiffArgs& a;
  const DiffSys& sy;
  int s;
  int* smem;
  double* red;
  DetectArgs* sd;
  double lsum;      // this thread's share of the loss
  double dlsum;     // tangent mode: share of the directional derivative
  v3 gth;           // this thread's share of d loss / d (kn, en, mu)
  // a stiff system advances sub steps of dt / sub per step: nst steps of dt
  // in all, checkpoints every Ks of them, observation k after obs[k] * sub
  int sub = 1, nst = 0, Ks = 0;
  double dt = 0.;

  __device__ __forceinline__ unsigned long long* err() const { return a.err + (size_t)s * kErrWords; }

  __device__ __forceinline__ bool failed() {
    if (threadIdx.x == 0) {
      unsigned long long all = kNoErr;
      for (int k = 0; k < kErrWords; ++k)
        all &= __hip_atomic_load(&err()[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      smem[0] = all != kNoErr;
    }
    __syncthreads();
    const bool f = smem[0] != 0;
    __syncthreads();
    return f;
  }

  __device__ __forceinline__ double4* X() const { return a.x + sy.off; }
  __device__ __forceinline__ double4* V() const { return a.v + sy.off; }
  __device__ __forceinline__ double4* W() const { return a.w + sy.off; }
  __device__ __forceinline__ double4* Fo() const { return a.F + sy.off; }
  __device__ __forceinline__ double4* To() const { return a.T + sy.off; }
  __device__ __forceinline__ const double4* P() const { return a.prop + sy.off; }

  // Keys of the current positions, counting sort by cell (ties by id), sorted
  // float copy, half-stencil detection: contact lists by slot.
  __device__ __forceinline__ void bin_and_detect() {
    const int n = sy.n, nc = sy.grid.ncell;
    int32_t* key = a.key + sy.off;
    int32_t* slot = a.slot + sy.off;
    int32_t* perm = a.perm + sy.off;
    int32_t* cnt = a.cnt + sy.cell_off;
    int32_t* cs = a.cell_start + sy.cs_off;
    const int32_t* rr = a.row_rank + sy.row_off;
    for (int i = threadIdx.x; i < n; i += BS) {
      const double4 p = X()[i];
      int ijk[3];
      if (!cell_of(sy.grid, mk(p.x, p.y, p.z), ijk)) {
        atomicMin(&err()[kErrSim], (unsigned long long)(uint32_t)i);
        key[i] = 0;
      } else {
        key[i] = row_base(sy.grid, rr, ijk[1], ijk[2]) + ijk[0];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += BS) slot[i] = atomicAdd(&cnt[key[i]], 1);
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < nc; base += BS) {
      const int c = base + threadIdx.x;
      const int v = c < nc ? cnt[c] : 0;
      int total;
      const int ex = block_scan_excl_d<BS>(v, smem, total);
      if (c < nc) {
        cs[c] = carry + ex;
        cnt[c] = 0;
      }
      carry += total;
    }
    if (threadIdx.x == 0) cs[nc] = n;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += BS) perm[cs[key[i]] + slot[i]] = i;
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += BS) {
      const int beg = cs[c], end = cs[c + 1];
      for (int q = beg + 1; q < end; ++q) {
        const int v = perm[q];
        int b = q - 1;
        while (b >= beg && perm[b] > v) {
          perm[b + 1] = perm[b];
          --b;
        }
        perm[b + 1] = v;
      }
    }
    __syncthreads();
    float4* pos = a.pos + sy.off;
    for (int u = threadIdx.x; u < n; u += BS) {
      const double4 p = X()[perm[u]];
      pos[u] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
    }
    if (threadIdx.x == 0) {
      DetectArgs& d = *sd;
      d.pos = pos;
      d.sizef[0] = sy.sizef[0]; d.sizef[1] = sy.sizef[1]; d.sizef[2] = sy.sizef[2];
      d.margin = sy.margin;
      d.in.pos_r = X();
      d.in.aux = a.aux + sy.off;
      d.perm = perm;
      d.cell_start = cs;
      d.row_rank = rr;
      d.grid = sy.grid;
      d.n = n;
      d.lstride = n;
      d.out_list = a.out_list + (size_t)sy.off * kListCap;
      d.out_count = a.out_count + sy.off;
      d.in_list = a.in_list + (size_t)sy.off * kListCap;
      d.in_count = a.in_count + sy.off;
      d.err = err();
      d.nblk8 = 0;
      d.skin = 0.;
      d.gate = SkinGate{};
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += BS) detect_particle(*sd, t);
    __syncthreads();
  }

  // Body order and image of one list entry of particle p (slot t): returns
  // the partner id; meA tells whether p is the first body; shift is added to
  // the partner (if p is A) or to p (if p is B).
  __device__ __forceinline__ int decode(int u, int code, bool& meA, v3& shift) const {
    const int sx = (code & 3) - 1, sy_ = ((code >> 2) & 3) - 1, sz = ((code >> 4) & 3) - 1;
    meA = (code & 64) != 0;
    shift = mk(sx * sy.grid.size[0], sy_ * sy.grid.size[1], sz * sy.grid.size[2]);
    return a.perm[sy.off + u];
  }

  __device__ __forceinline__ PairIn pair_in(int p, int q, bool meA, v3 shift) const {
    // shift is applied to the partner q when p is A, to p when p is B
    const int A = meA ? p : q, B = meA ? q : p;
    const double4 pa = P()[A], pb = P()[B];
    PairIn in;
    const double4 xa = X()[A], xb = X()[B];
    in.xA = mk(xa.x, xa.y, xa.z);
    if (meA) in.xB = mk(xb.x + shift.x, xb.y + shift.y, xb.z + shift.z);
    else in.xB = mk(xb.x + (-shift.x), xb.y + (-shift.y), xb.z + (-shift.z));
    in.vA = ld3(V()[A]); in.wA = ld3(W()[A]);
    in.vB = ld3(V()[B]); in.wB = ld3(W()[B]);
    in.rA = pa.x; in.mA = pa.y; in.crA = pa.w;
    in.rB = pb.x; in.mB = pb.y; in.crB = pb.w;
    return in;
  }

  __device__ __forceinline__ HookeParams<double> params() const {
    HookeParams<double> h;
    h.kn = sy.kn;
    h.beta = sy.beta;
    h.mu = sy.mu;
    return h;
  }

  // Force and torque of every particle (weight + contacts in a fixed order + walls).
  __device__ __forceinline__ void forces() {
    const int n = sy.n;
    const HookeParams<double> hp = params();
    for (int t = threadIdx.x; t < n; t += BS) {
      const int p = a.perm[sy.off + t];
      const double4 pp = P()[p];
      const double mw = pp.y * (1. - 0. / 1.);
      v3 F = mk(sy.gravity[0] * mw, sy.gravity[1] * mw, sy.gravity[2] * mw);
      v3 T = mk(0., 0., 0.);
      auto one = [&](int u, int code) {
        bool meA;
        v3 shift;
        const int q = decode(u, code, meA, shift);
        const PairIn in = pair_in(p, q, meA, shift);
        V3<double> FA, TA, FB, TB;
        if (!pair_torsors(lift3<double>(in.xA), lift3<double>(in.vA), lift3<double>(in.wA), in.rA,
                          in.crA, in.mA, lift3<double>(in.xB), lift3<double>(in.vB),
                          lift3<double>(in.wB), in.rB, in.crB, in.mB, hp, FA, TA, FB, TB))
          return;
        if (meA) {
          F = add(F, val3(FA));
          T = add(T, val3(TA));
        } else {
          F = add(F, val3(FB));
          T = add(T, val3(TB));
        }
      };
      const int nout = min(a.out_count[sy.off + t], kListCap);
      const int2* ol = a.out_list + (size_t)sy.off * kListCap + t;
      for (int e = 0; e < nout; ++e) one(ol[(size_t)e * sy.n].x, ol[(size_t)e * sy.n].y);
      const int nin = min(a.in_count[sy.off + t], kListCap);
      a.in_count[sy.off + t] = 0;
      const int2* il = a.in_list + (size_t)sy.off * kListCap + t;
      long long last = -1;
      for (int e = 0; e < nin; ++e) {
        long long k;
        const int f = next_incoming(il, nin, sy.n, last, k);
        if (f < 0) break;
        one(il[(size_t)f * sy.n].x, il[(size_t)f * sy.n].y);
        last = k;
      }
      const v3 x = ld3(X()[p]), vv = ld3(V()[p]), ww = ld3(W()[p]);
      for (int wi = 0; wi < sy.nwalls; ++wi) {
        const Wall& wl = a.walls[sy.wall_off + wi];
        if (!(fabs(wl.bb_centre[0] - x.x) <= wl.bb_ext[0] + pp.x &&
              fabs(wl.bb_centre[1] - x.y) <= wl.bb_ext[1] + pp.x &&
              fabs(wl.bb_centre[2] - x.z) <= wl.bb_ext[2] + pp.x))
          continue;
        V3<double> FA, TA;
        V3<double> c, ov;
        double dist;
        const int rc = sw_contact(wl, lift3<double>(x), pp.x, pp.w, c, ov, dist);
        if (rc < 0) {
          atomicMin(&err()[kErrContact], ((unsigned long long)(uint32_t)p << 32) | (uint32_t)(-1 - wi));
          continue;
        }
        if (!wall_torsor(wl, lift3<double>(x), lift3<double>(vv), lift3<double>(ww), pp.x, pp.w,
                         pp.y, hp, FA, TA))
          continue;
        F = add(F, val3(FA));
        T = add(T, val3(TA));
      }
      Fo()[p] = st3(F, 0.);
      To()[p] = st3(T, 0.);
    }
    __syncthreads();
  }

  // First half kick and drift: v_h = v + (F/m) dt/2, w_h = w + (T/I) dt/2,
  // x' = x + v_h dt (ParticleKinematics.cpp:62-104, SecondOrderLeapFrog.cpp:44-67)
  __device__ __forceinline__ void kick_drift() {
    const double hdt = 0.5 * dt;
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 pp = P()[i];
      const v3 vh = add(ld3(V()[i]), scl(dvd(ld3(Fo()[i]), pp.y * 1.), hdt));
      const v3 wh = add(ld3(W()[i]), scl(dvd(ld3(To()[i]), pp.z), hdt));
      const v3 dx = scl(vh, dt);
      const double4 xo = X()[i];
      X()[i] = st3(add(mk(xo.x, xo.y, xo.z), dx), xo.w);
      V()[i] = st3(vh, 0.);
      W()[i] = st3(wh, 0.);
      if (norm(dx) > pp.w) atomicMin(&err()[kErrMotion], (unsigned long long)(uint32_t)i);
    }
    __syncthreads();
  }

  // Closing half kick (SecondOrderLeapFrog.cpp:74-94)
  __device__ __forceinline__ void kick() {
    const double hdt = 0.5 * dt;
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 pp = P()[i];
      V()[i] = st3(add(ld3(V()[i]), scl(dvd(ld3(Fo()[i]), pp.y * 1.), hdt)), 0.);
      W()[i] = st3(add(ld3(W()[i]), scl(dvd(ld3(To()[i]), pp.z), hdt)), 0.);
    }
    __syncthreads();
  }

  __device__ __forceinline__ int obs_index(int step) const {
    if (step % sub != 0) return -1;
    for (int k = 0; k < a.nobs; ++k)
      if (a.obs[k] * sub == step) return k;
    return -1;
  }

  __device__ __forceinline__ void save(double4* dst_x, double4* dst_v, double4* dst_w, double4* dst_F,
                       double4* dst_T) {
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      dst_x[sy.off + i] = X()[i];
      dst_v[sy.off + i] = V()[i];
      dst_w[sy.off + i] = W()[i];
      if (dst_F) dst_F[sy.off + i] = Fo()[i];
      if (dst_T) dst_T[sy.off + i] = To()[i];
    }
    __syncthreads();
  }

  __device__ __forceinline__ void restore(const double4* sx, const double4* sv, const double4* sw,
                          const double4* sF, const double4* sT) {
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      X()[i] = sx[sy.off + i];
      V()[i] = sv[sy.off + i];
      W()[i] = sw[sy.off + i];
      if (sF) Fo()[i] = sF[sy.off + i];
      if (sT) To()[i] = sT[sy.off + i];
    }
    __syncthreads();
  }

  // ---- forward-mode tangents (DIFF_TANGENT): tx, tv, tw, tF, tT live in
  // the adjoint arrays ax, av, aw, aF, aT
  __device__ __forceinline__ v3 dtheta_beta() const {
    // tangent of (kn, beta(en), mu) for the system's direction in (kn, en, mu)
    const double* d = a.dtheta + 3 * s;
    return mk(d[0], sy.dbeta * d[1], d[2]);
  }

  __device__ __forceinline__ void tangent_forces() {
    const int n = sy.n;
    const v3 tp = dtheta_beta();
    const double4 *tx = a.ax + sy.off, *tv = a.av + sy.off, *tw = a.aw + sy.off;
    for (int t = threadIdx.x; t < n; t += BS) {
      const int p = a.perm[sy.off + t];
      v3 dF = mk(0., 0., 0.), dT = mk(0., 0., 0.);
      auto one = [&](int u, int code) {
        bool meA;
        v3 shift;
        const int q = decode(u, code, meA, shift);
        const PairIn in = pair_in(p, q, meA, shift);
        const int A = meA ? p : q, B = meA ? q : p;
        PairTan tin;
        tin.xA = ld3(tx[A]); tin.vA = ld3(tv[A]); tin.wA = ld3(tw[A]);
        tin.xB = ld3(tx[B]); tin.vB = ld3(tv[B]); tin.wB = ld3(tw[B]);
        v3 f, tq;
        if (pair_jvp(in, tin, sy.kn, sy.beta, sy.mu, tp, meA, f, tq)) {
          dF = add(dF, f);
          dT = add(dT, tq);
        }
      };
      const int nout = min(a.out_count[sy.off + t], kListCap);
      const int2* ol = a.out_list + (size_t)sy.off * kListCap + t;
      for (int e = 0; e < nout; ++e) one(ol[(size_t)e * sy.n].x, ol[(size_t)e * sy.n].y);
      const int nin = min(a.in_count[sy.off + t], kListCap);
      const int2* il = a.in_list + (size_t)sy.off * kListCap + t;
      for (int e = 0; e < nin; ++e) one(il[(size_t)e * sy.n].x, il[(size_t)e * sy.n].y);
      const double4 pp = P()[p];
      const v3 x = ld3(X()[p]), vv = ld3(V()[p]), ww = ld3(W()[p]);
      const v3 dx = ld3(tx[p]), dvv = ld3(tv[p]), dww = ld3(tw[p]);
      auto dv = [](v3 a0, v3 t0) { return mk3(Dual{a0.x, t0.x}, Dual{a0.y, t0.y}, Dual{a0.z, t0.z}); };
      HookeParams<Dual> hp;
      hp.kn = Dual{sy.kn, tp.x};
      hp.beta = Dual{sy.beta, tp.y};
      hp.mu = Dual{sy.mu, tp.z};
      for (int wi = 0; wi < sy.nwalls; ++wi) {
        const Wall& wl = a.walls[sy.wall_off + wi];
        if (!(fabs(wl.bb_centre[0] - x.x) <= wl.bb_ext[0] + pp.x &&
              fabs(wl.bb_centre[1] - x.y) <= wl.bb_ext[1] + pp.x &&
              fabs(wl.bb_centre[2] - x.z) <= wl.bb_ext[2] + pp.x))
          continue;
        V3<Dual> FA, TA;
        if (!wall_torsor(wl, dv(x, dx), dv(vv, dvv), dv(ww, dww), pp.x, pp.w, pp.y, hp, FA, TA)) continue;
        dF = add(dF, mk(FA.x.d, FA.y.d, FA.z.d));
        dT = add(dT, mk(TA.x.d, TA.y.d, TA.z.d));
      }
      a.aF[sy.off + p] = st3(dF, 0.);
      a.aT[sy.off + p] = st3(dT, 0.);
    }
    __syncthreads();
  }

  // tangent of the first kick and the drift (uses the tangents of F, T at n)
  __device__ __forceinline__ void tangent_kick_drift() {
    const double hdt = 0.5 * dt;
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 pp = P()[i];
      const size_t g = sy.off + i;
      const v3 tvh = add(ld3(a.av[g]), scl(ld3(a.aF[g]), hdt / pp.y));
      const v3 twh = add(ld3(a.aw[g]), scl(ld3(a.aT[g]), hdt / pp.z));
      a.ax[g] = st3(add(ld3(a.ax[g]), scl(tvh, dt)), 0.);
      a.av[g] = st3(tvh, 0.);
      a.aw[g] = st3(twh, 0.);
    }
    __syncthreads();
  }

  __device__ __forceinline__ void tangent_kick() {
    const double hdt = 0.5 * dt;
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 pp = P()[i];
      const size_t g = sy.off + i;
      a.av[g] = st3(add(ld3(a.av[g]), scl(ld3(a.aF[g]), hdt / pp.y)), 0.);
      a.aw[g] = st3(add(ld3(a.aw[g]), scl(ld3(a.aT[g]), hdt / pp.z)), 0.);
    }
    __syncthreads();
  }

  // One forward step n -> n+1; optionally stores the force inputs in window slot j.
  __device__ __forceinline__ void step(int n1, int j) {
    constexpr bool tan = TAN;
    if (tan) tangent_kick_drift();
    kick_drift();
    bin_and_detect();
    if (j >= 0) {
      for (int i = threadIdx.x; i < sy.n; i += BS) {
        a.win_x[(size_t)j * a.ntot + sy.off + i] = X()[i];
        a.win_v[(size_t)j * a.ntot + sy.off + i] = V()[i];
        a.win_w[(size_t)j * a.ntot + sy.off + i] = W()[i];
      }
    }
    if (tan) tangent_forces();  // reads in_count before forces() clears it
    forces();
    if (tan) tangent_kick();
    kick();
    const int k = obs_index(n1);
    if (k >= 0 && j < 0) {
      for (int i = threadIdx.x; i < sy.n; i += BS) {
        const double4 x = X()[i];
        if (a.mode == DIFF_RECORD) {
          a.xobs[(size_t)k * a.ntot + sy.off + i] = x;
        } else {
          const double4 tg = a.target[(size_t)k * a.ntot + sy.off + i];
          const double dx = x.x - tg.x, dy = x.y - tg.y, dz = x.z - tg.z;
          lsum += dx * dx + dy * dy + dz * dz;
          if (tan) {
            const double4 t = a.ax[sy.off + i];
            dlsum += 2. * (dx * t.x + dy * t.y + dz * t.z);
          }
        }
      }
    }
  }

  // Adjoint of the force map at the current state arrays (x, v, w) for the
  // force/torque adjoints aFt/aTt: adds d/dx, d/dv, d/dw of every particle to
  // gx/gv/gw (per particle, written in place) and d/dtheta to this thread.
  __device__ __forceinline__ void force_vjp(double4* gx, double4* gv, double4* gw) {
    const int n = sy.n;
    const double4* aFt = a.aFt + sy.off;
    const double4* aTt = a.aTt + sy.off;
    for (int t = threadIdx.x; t < n; t += BS) {
      const int p = a.perm[sy.off + t];
      v3 dx = mk(0., 0., 0.), dv = mk(0., 0., 0.), dw = mk(0., 0., 0.);
      auto one = [&](int u, int code, bool own) {
        bool meA;
        v3 shift;
        const int q = decode(u, code, meA, shift);
        const PairIn in = pair_in(p, q, meA, shift);
        const int A = meA ? p : q, B = meA ? q : p;
        PairAdj adj;
        adj.FA = ld3(aFt[A]); adj.TA = ld3(aTt[A]);
        adj.FB = ld3(aFt[B]); adj.TB = ld3(aTt[B]);
        // one reverse pass gives every column: this particle's 9 inputs and,
        // if it detected the pair, the 3 parameters
        PairGrad pg;
        if (!pair_vjp(in, adj, sy.kn, sy.beta, sy.dbeta, sy.mu, pg)) return;
        dx = add(dx, meA ? pg.xA : pg.xB);
        dv = add(dv, meA ? pg.vA : pg.vB);
        dw = add(dw, meA ? pg.wA : pg.wB);
        if (own) gth = add(gth, pg.th);
      };
      const int nout = min(a.out_count[sy.off + t], kListCap);
      const int2* ol = a.out_list + (size_t)sy.off * kListCap + t;
      for (int e = 0; e < nout; ++e) one(ol[(size_t)e * sy.n].x, ol[(size_t)e * sy.n].y, true);
      const int nin = min(a.in_count[sy.off + t], kListCap);
      a.in_count[sy.off + t] = 0;
      const int2* il = a.in_list + (size_t)sy.off * kListCap + t;
      long long last = -1;
      for (int e = 0; e < nin; ++e) {
        long long k;
        const int f = next_incoming(il, nin, sy.n, last, k);
        if (f < 0) break;
        one(il[(size_t)f * sy.n].x, il[(size_t)f * sy.n].y, false);
        last = k;
      }
      const double4 pp = P()[p];
      const v3 x = ld3(X()[p]), vv = ld3(V()[p]), ww = ld3(W()[p]);
      const v3 aF = ld3(aFt[p]), aT = ld3(aTt[p]);
      for (int wi = 0; wi < sy.nwalls; ++wi) {
        const Wall& wl = a.walls[sy.wall_off + wi];
        if (!(fabs(wl.bb_centre[0] - x.x) <= wl.bb_ext[0] + pp.x &&
              fabs(wl.bb_centre[1] - x.y) <= wl.bb_ext[1] + pp.x &&
              fabs(wl.bb_centre[2] - x.z) <= wl.bb_ext[2] + pp.x))
          continue;
        PairGrad wg;
        if (!wall_vjp(wl, x, vv, ww, pp.x, pp.w, pp.y, aF, aT, sy.kn, sy.beta, sy.dbeta, sy.mu, wg))
          continue;
        dx = add(dx, wg.xA);
        dv = add(dv, wg.vA);
        dw = add(dw, wg.wA);
        gth = add(gth, wg.th);
      }
      gx[sy.off + p] = st3(dx, 0.);
      gv[sy.off + p] = st3(dv, 0.);
      gw[sy.off + p] = st3(dw, 0.);
    }
    __syncthreads();
  }

  // Adjoint of one step n -> n+1 whose force inputs are in the state arrays;
  // adjoint arrays hold the adjoint at boundary n+1 on entry, at n on exit.
  __device__ __forceinline__ void step_adjoint(int n1, double lscale) {
    const double hdt = 0.5 * dt;
    double4 *ax = a.ax + sy.off, *av = a.av + sy.off, *aw = a.aw + sy.off, *aF = a.aF + sy.off,
            *aT = a.aT + sy.off;
    const int k = obs_index(n1);
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 pp = P()[i];
      v3 gx = ld3(ax[i]);
      if (k >= 0) {
        const double4 x = X()[i];
        const double4 tg = a.target[(size_t)k * a.ntot + sy.off + i];
        gx = add(gx, scl(mk(x.x - tg.x, x.y - tg.y, x.z - tg.z), 2. * lscale));
      }
      ax[i] = st3(gx, 0.);
      // closing kick v' = v_h + (F'/m) dt/2, w' = w_h + (T'/I) dt/2
      a.aFt[sy.off + i] = st3(add(ld3(aF[i]), scl(ld3(av[i]), hdt / pp.y)), 0.);
      a.aTt[sy.off + i] = st3(add(ld3(aT[i]), scl(ld3(aw[i]), hdt / pp.z)), 0.);
    }
    __syncthreads();
    bin_and_detect();
    force_vjp(a.gX, a.gV, a.gW);
    // drift x' = x + v_h dt and first kick v_h = v + (F/m) dt/2, w_h = w + (T/I) dt/2
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 pp = P()[i];
      const v3 xt = add(ld3(ax[i]), ld3(a.gX[sy.off + i]));
      const v3 vt = add(add(ld3(av[i]), ld3(a.gV[sy.off + i])), scl(xt, dt));
      const v3 wt = add(ld3(aw[i]), ld3(a.gW[sy.off + i]));
      ax[i] = st3(xt, 0.);
      av[i] = st3(vt, 0.);
      aw[i] = st3(wt, 0.);
      aF[i] = st3(scl(vt, hdt / pp.y), 0.);
      aT[i] = st3(scl(wt, hdt / pp.z), 0.);
    }
    __syncthreads();
  }

  __device__ __forceinline__ void initial_state() {
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      X()[i] = a.x0[sy.off + i];
      V()[i] = a.v0[sy.off + i];
      W()[i] = a.w0[sy.off + i];
    }
    __syncthreads();
  }

  __device__ __forceinline__ void run() {
    lsum = 0.;
    dlsum = 0.;
    gth = mk(0., 0., 0.);
    sub = sy.substeps;
    dt = a.dt / (double)sub;
    nst = a.nsteps * sub;
    Ks = a.K * sub;
    const double lscale = 1. / ((double)a.nobs * 3. * (double)sy.n);
    initial_state();
    bin_and_detect();
    if constexpr (TAN) {
      // tangents of the initial state: positions fixed, velocities along the
      // given direction (in gv0), angular velocities fixed
      for (int i = threadIdx.x; i < sy.n; i += BS) {
        const double4 z = make_double4(0., 0., 0., 0.);
        a.ax[sy.off + i] = z;
        a.av[sy.off + i] = a.gv0[sy.off + i];
        a.aw[sy.off + i] = z;
      }
      __syncthreads();
      tangent_forces();
    }
    forces();
    if (a.mode == DIFF_GRAD) save(a.ck_x, a.ck_v, a.ck_w, a.ck_F, a.ck_T);
    bool ok = true;
    for (int n = 0; n < nst; ++n) {
      if (failed()) { ok = false; break; }
      step(n + 1, -1);
      if (a.mode == DIFF_GRAD && (n + 1) % Ks == 0 && n + 1 < nst) {
        const size_t c = (size_t)((n + 1) / Ks) * a.ntot;
        save(a.ck_x + c, a.ck_v + c, a.ck_w + c, a.ck_F + c, a.ck_T + c);
      }
    }
    if (ok && failed()) ok = false;
    const double L = block_sum<BS>(lsum, red) * lscale;
    if (threadIdx.x == 0) a.loss[s] = ok ? L : __builtin_nan("");
    if constexpr (TAN) {
      const double dL = block_sum<BS>(dlsum, red) * lscale;
      if (threadIdx.x == 0) a.dloss[s] = ok ? dL : __builtin_nan("");
    }
    if (a.mode != DIFF_GRAD || !ok) return;
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      const double4 z = make_double4(0., 0., 0., 0.);
      a.ax[sy.off + i] = z; a.av[sy.off + i] = z; a.aw[sy.off + i] = z;
      a.aF[sy.off + i] = z; a.aT[sy.off + i] = z;
    }
    __syncthreads();
    const int nwin = (nst + Ks - 1) / Ks;
    for (int wdx = nwin - 1; wdx >= 0; --wdx) {
      const int w0 = wdx * Ks;
      const int len = min(Ks, nst - w0);
      const size_t c = (size_t)wdx * a.ntot;
      restore(a.ck_x + c, a.ck_v + c, a.ck_w + c, a.ck_F + c, a.ck_T + c);
      for (int j = 0; j < len; ++j) step(w0 + j + 1, j);
      for (int j = len - 1; j >= 0; --j) {
        const size_t wj = (size_t)j * a.ntot;
        restore(a.win_x + wj, a.win_v + wj, a.win_w + wj, nullptr, nullptr);
        step_adjoint(w0 + j + 1, lscale);
      }
    }
    // a0 = forces of the initial state (the prime)
    initial_state();
    for (int i = threadIdx.x; i < sy.n; i += BS) {
      a.aFt[sy.off + i] = a.aF[sy.off + i];
      a.aTt[sy.off + i] = a.aT[sy.off + i];
    }
    __syncthreads();
    bin_and_detect();
    force_vjp(a.gX, a.gV, a.gW);
    for (int i = threadIdx.x; i < sy.n; i += BS)
      a.gv0[sy.off + i] = st3(add(ld3(a.av[sy.off + i]), ld3(a.gV[sy.off + i])), 0.);
    const double g0 = block_sum<BS>(gth.x, red);
    const double g1 = block_sum<BS>(gth.y, red);
    const double g2 = block_sum<BS>(gth.z, red);
    if (threadIdx.x == 0) {
      a.gtheta[3 * s] = g0;
      a.gtheta[3 * s + 1] = g1;
      a.gtheta[3 * s + 2] = g2;
    }
    __syncthreads();
  }
};

template <int BS, bool TAN>
__global__ void __launch_bounds__(BS) k_diff(DiffArgs a) {
  __shared__ int s_item;
  __shared__ int smem[BS / 64 > 1 ? BS / 64 : 2];
  __shared__ double red[BS];
  __shared__ DetectArgs sd;
  __shared__ DiffSys ssy;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(a.next, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= a.nsys) break;
    const int s = a.order[item];
    if (threadIdx.x == 0) ssy = a.sys[s];
    __syncthreads();
    DiffRunner<BS, TAN> r{a, ssy, s, smem, red, &sd, 0., 0., mk(0., 0., 0.)};
    r.run();
    __syncthreads();
  }
}

__global__ void k_diff_zero(int32_t* p) { *p = 0; }

}  // namespace gdem

using namespace gdem;

struct gdem_diff {
  int dev = 0;
  hipStream_t stream = nullptr;
  bool committed = false;
  int32_t block = 256, workgroups = 0;
  int32_t nsteps = 0, K = 50, nobs = 0;
  int64_t win_steps = 0;  // window slots: K times the largest substep count
  double dt = 0.;
  std::vector<int32_t> obs;
  std::vector<DiffSys> sys;
  std::vector<double4> prop, x0, v0, w0;
  std::vector<Aux> aux;
  std::vector<int32_t> rows;
  std::vector<Wall> walls;
  int64_t ntot = 0, ncell_tot = 0, ncs_tot = 0;
  // device
  DiffSys* d_sys = nullptr;
  int32_t *d_order = nullptr, *d_next = nullptr, *d_obs = nullptr;
  double4 *d_prop = nullptr, *d_x0 = nullptr, *d_v0 = nullptr, *d_w0 = nullptr;
  Aux* d_aux = nullptr;
  double4 *d_x = nullptr, *d_v = nullptr, *d_w = nullptr, *d_F = nullptr, *d_T = nullptr;
  double4 *d_ck = nullptr, *d_win = nullptr, *d_target = nullptr, *d_xobs = nullptr;
  int32_t *d_key = nullptr, *d_slot = nullptr, *d_perm = nullptr, *d_cnt = nullptr, *d_cs = nullptr,
          *d_rows = nullptr, *d_out_count = nullptr, *d_in_count = nullptr;
  float4* d_pos = nullptr;
  int2 *d_out_list = nullptr, *d_in_list = nullptr;
  Wall* d_walls = nullptr;
  double4* d_adj = nullptr;  // ax av aw aF aT aFt aTt gX gV gW gv0
  double *d_loss = nullptr, *d_gtheta = nullptr;
  unsigned long long* d_err = nullptr;
  int32_t err_sys = -1, err_code = 0, err_id0 = -1, err_id1 = -1;
};

#define DHIP(expr)                                                                 \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "gdem_diff: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return GDEM_ERR_HIP;                                                         \
    }                                                                              \
  } while (0)

extern "C" int gdem_diff_create(int32_t device, gdem_diff** out) {
  if (!out) return GDEM_ERR_ARG;
  *out = nullptr;
  gdem_diff* d = new gdem_diff();
  d->dev = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    return GDEM_ERR_HIP;
  }
  *out = d;
  return GDEM_OK;
}

extern "C" int gdem_diff_destroy(gdem_diff* d) {
  if (!d) return GDEM_OK;
  (void)hipSetDevice(d->dev);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  // d_v, d_w, d_F, d_T point into the d_x allocation
  void* ptrs[] = {d->d_sys, d->d_order, d->d_next, d->d_obs, d->d_prop, d->d_x0, d->d_v0, d->d_w0,
                  d->d_aux, d->d_x, d->d_ck, d->d_win, d->d_target,
                  d->d_xobs, d->d_key, d->d_slot, d->d_perm, d->d_cnt, d->d_cs, d->d_rows,
                  d->d_out_count, d->d_in_count, d->d_pos, d->d_out_list, d->d_in_list, d->d_walls,
                  d->d_adj, d->d_loss, d->d_gtheta, d->d_err};
  for (void* p : ptrs) (void)hipFree(p);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return GDEM_OK;
}

// beta(en) of HookeContactForceModel.cpp:19 and its derivative
// d beta / d en = pi^2 / (en (pi^2 + ln^2 en)^(3/2)).
static void beta_of(double en, double& beta, double& dbeta) {
  const double L = log(en);
  beta = L / sqrt(kPi * kPi + L * L);
  dbeta = kPi * kPi / (en * pow(kPi * kPi + L * L, 1.5));
}

extern "C" int gdem_diff_add_system(gdem_diff* d, const gdem_batch_system* in, int32_t* sys_id) {
  if (!d || !in || d->committed) return GDEM_ERR_ARG;
  const int64_t n = in->n;
  if (n <= 0 || d->ntot + n > 0x7fffffff) return GDEM_ERR_ARG;
  if (!in->x || !in->v || !in->w || !in->q || !in->radius || !in->density || !in->crust)
    return GDEM_ERR_ARG;
  if (in->kind != GDEM_MODEL_HOOKE || !model_params_ok(in->kind, in->pp)) return GDEM_ERR_ARG;
  if (in->pp[2] != -1.) return GDEM_ERR_ARG;  // tangential damping from en
  if (in->pp[4] != 0.) return GDEM_ERR_ARG;  // no rolling resistance in the adjoint
  for (int k = 0; k < 5; ++k)
    if (in->pw[k] != in->pp[k]) return GDEM_ERR_ARG;  // one parameter set per system
  if (in->nwalls < 0 || in->nwalls > kMaxWalls) return GDEM_ERR_ARG;
  if (in->nwalls > 0 && (!in->wall_basis || !in->wall_centre || !in->wall_half)) return GDEM_ERR_ARG;
  if (in->substeps < 0 || in->substeps > 4096) return GDEM_ERR_ARG;
  DiffSys sy{};
  sy.substeps = std::max<int32_t>(in->substeps, 1);
  if (!make_grid(in->cfg, sy.grid)) return GDEM_ERR_ARG;
  sy.margin = prefilter_margin(sy.grid, sy.sizef);
  for (int k = 0; k < 3; ++k) sy.gravity[k] = in->cfg.gravity[k];
  sy.n = (int32_t)n;
  sy.off = (int32_t)d->ntot;
  sy.cell_off = (int32_t)d->ncell_tot;
  sy.cs_off = (int32_t)d->ncs_tot;
  sy.row_off = (int32_t)d->rows.size();
  sy.wall_off = (int32_t)d->walls.size();
  sy.nwalls = in->nwalls;
  sy.kn = in->pp[0];
  sy.en = in->pp[1];
  sy.mu = in->pp[3];
  beta_of(sy.en, sy.beta, sy.dbeta);
  for (int64_t i = 0; i < n; ++i)
    if (!(in->radius[i] > 0.) || !(in->density[i] > 0.) || !(in->crust[i] > 0.)) return GDEM_ERR_ARG;
  const std::vector<int32_t> rank = row_morton_rank(sy.grid.nc);
  d->rows.insert(d->rows.end(), rank.begin(), rank.end());
  for (int k = 0; k < in->nwalls; ++k)
    d->walls.push_back(make_wall(in->wall_basis + 9 * k, in->wall_centre + 3 * k,
                                 in->wall_half + 3 * k, in->wall_crust, 1));
  for (int64_t i = 0; i < n; ++i) {
    double4 pr, vm, wr, wb, qq;
    Aux ax;
    sphere_particle(in->x + 3 * i, in->v + 3 * i, in->w + 3 * i, in->q + 4 * i, in->radius[i],
                    in->density[i], in->crust[i], 0, (int32_t)i, pr, vm, wr, wb, qq, ax);
    d->prop.push_back(make_double4(pr.w, vm.w, wb.w, in->crust[i]));
    d->x0.push_back(pr);
    d->v0.push_back(make_double4(vm.x, vm.y, vm.z, 0.));
    d->w0.push_back(make_double4(wr.x, wr.y, wr.z, 0.));
    d->aux.push_back(ax);
  }
  d->ntot += n;
  d->ncell_tot += sy.grid.ncell;
  d->ncs_tot += sy.grid.ncell + 1;
  if (sys_id) *sys_id = (int32_t)d->sys.size();
  d->sys.push_back(sy);
  return GDEM_OK;
}

template <class T>
static hipError_t dupload(T** p, const std::vector<T>& h) {
  hipError_t e = hipMalloc(p, sizeof(T) * std::max<size_t>(h.size(), 1));
  if (e != hipSuccess || h.empty()) return e;
  return hipMemcpy(*p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

extern "C" int gdem_diff_commit(gdem_diff* d, int32_t nsteps, double dt, int32_t checkpoint_every,
                                int32_t nobs, const int32_t* obs_steps) {
  if (!d || d->committed || d->sys.empty()) return GDEM_ERR_ARG;
  if (nsteps <= 0 || !(dt > 0.) || checkpoint_every <= 0 || nobs <= 0 || !obs_steps) return GDEM_ERR_ARG;
  for (int k = 0; k < nobs; ++k) {
    if (obs_steps[k] < 1 || obs_steps[k] > nsteps) return GDEM_ERR_ARG;
    if (k > 0 && obs_steps[k] <= obs_steps[k - 1]) return GDEM_ERR_ARG;
  }
  DHIP(hipSetDevice(d->dev));
  d->nsteps = nsteps;
  d->dt = dt;
  d->K = std::min(checkpoint_every, nsteps);
  d->nobs = nobs;
  d->obs.assign(obs_steps, obs_steps + nobs);
  const size_t N = (size_t)d->ntot;
  const size_t nck = (size_t)((nsteps + d->K - 1) / d->K);
  std::vector<int32_t> order(d->sys.size());
  std::iota(order.begin(), order.end(), 0);
  // largest work first: particles times substeps
  int32_t smax = 1;
  for (const DiffSys& sy : d->sys) smax = std::max(smax, sy.substeps);
  auto work = [&](int32_t k) { return (int64_t)d->sys[k].n * d->sys[k].substeps; };
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return work(x) > work(y); });
  DHIP(dupload(&d->d_sys, d->sys));
  DHIP(dupload(&d->d_order, order));
  DHIP(dupload(&d->d_obs, d->obs));
  DHIP(dupload(&d->d_prop, d->prop));
  DHIP(dupload(&d->d_x0, d->x0));
  DHIP(dupload(&d->d_v0, d->v0));
  DHIP(dupload(&d->d_w0, d->w0));
  DHIP(dupload(&d->d_aux, d->aux));
  DHIP(dupload(&d->d_rows, d->rows));
  DHIP(dupload(&d->d_walls, d->walls));
  DHIP(hipMalloc(&d->d_next, sizeof(int32_t)));
  DHIP(hipMalloc(&d->d_x, sizeof(double4) * N * 5));
  d->d_v = d->d_x + N;
  d->d_w = d->d_v + N;
  d->d_F = d->d_w + N;
  d->d_T = d->d_F + N;
  DHIP(hipMalloc(&d->d_ck, sizeof(double4) * N * nck * 5));
  // window of K steps of the stiffest system: K * smax force inputs
  DHIP(hipMalloc(&d->d_win, sizeof(double4) * N * (size_t)d->K * (size_t)smax * 3));
  d->win_steps = (int64_t)d->K * smax;
  DHIP(hipMalloc(&d->d_target, sizeof(double4) * N * (size_t)nobs));
  DHIP(hipMemset(d->d_target, 0, sizeof(double4) * N * (size_t)nobs));
  DHIP(hipMalloc(&d->d_xobs, sizeof(double4) * N * (size_t)nobs));
  DHIP(hipMalloc(&d->d_key, sizeof(int32_t) * N));
  DHIP(hipMalloc(&d->d_slot, sizeof(int32_t) * N));
  DHIP(hipMalloc(&d->d_perm, sizeof(int32_t) * N));
  DHIP(hipMalloc(&d->d_pos, sizeof(float4) * N));
  DHIP(hipMalloc(&d->d_cnt, sizeof(int32_t) * (size_t)std::max<int64_t>(d->ncell_tot, 1)));
  DHIP(hipMemset(d->d_cnt, 0, sizeof(int32_t) * (size_t)std::max<int64_t>(d->ncell_tot, 1)));
  DHIP(hipMalloc(&d->d_cs, sizeof(int32_t) * (size_t)std::max<int64_t>(d->ncs_tot, 1)));
  DHIP(hipMalloc(&d->d_out_list, sizeof(int2) * N * kListCap));
  DHIP(hipMalloc(&d->d_in_list, sizeof(int2) * N * kListCap));
  DHIP(hipMalloc(&d->d_out_count, sizeof(int32_t) * N));
  DHIP(hipMalloc(&d->d_in_count, sizeof(int32_t) * N));
  DHIP(hipMemset(d->d_in_count, 0, sizeof(int32_t) * N));
  DHIP(hipMalloc(&d->d_adj, sizeof(double4) * N * 11));
  DHIP(hipMemset(d->d_adj, 0, sizeof(double4) * N * 11));
  const size_t ns = d->sys.size();
  DHIP(hipMalloc(&d->d_loss, sizeof(double) * ns));
  DHIP(hipMalloc(&d->d_gtheta, sizeof(double) * ns * 3));
  DHIP(hipMalloc(&d->d_err, sizeof(unsigned long long) * kErrWords * ns));
  DHIP(hipMemset(d->d_err, 0xff, sizeof(unsigned long long) * kErrWords * ns));
  DHIP(hipDeviceSynchronize());
  d->committed = true;
  return GDEM_OK;
}

extern "C" int gdem_diff_set_launch(gdem_diff* d, int32_t block_size, int32_t workgroups) {
  if (!d) return GDEM_ERR_ARG;
  if (block_size != 64 && block_size != 128 && block_size != 256 && block_size != 512)
    return GDEM_ERR_ARG;
  if (workgroups < 0) return GDEM_ERR_ARG;
  d->block = block_size;
  d->workgroups = workgroups;
  return GDEM_OK;
}

extern "C" int gdem_diff_set_parameters(gdem_diff* d, const double* theta, const double* v0) {
  if (!d || !d->committed) return GDEM_ERR_ARG;
  DHIP(hipSetDevice(d->dev));
  if (theta) {
    for (size_t s = 0; s < d->sys.size(); ++s) {
      const double kn = theta[3 * s], en = theta[3 * s + 1], mu = theta[3 * s + 2];
      if (!(kn > 0.) || !(en > 0. && en <= 1.) || !(mu >= 0.)) return GDEM_ERR_ARG;
      DiffSys& sy = d->sys[s];
      sy.kn = kn;
      sy.en = en;
      sy.mu = mu;
      beta_of(en, sy.beta, sy.dbeta);
    }
    DHIP(hipMemcpy(d->d_sys, d->sys.data(), sizeof(DiffSys) * d->sys.size(), hipMemcpyHostToDevice));
  }
  if (v0) {
    std::vector<double4> h((size_t)d->ntot);
    for (size_t i = 0; i < h.size(); ++i) h[i] = make_double4(v0[3 * i], v0[3 * i + 1], v0[3 * i + 2], 0.);
    DHIP(hipMemcpy(d->d_v0, h.data(), sizeof(double4) * h.size(), hipMemcpyHostToDevice));
  }
  return GDEM_OK;
}

extern "C" int gdem_diff_set_targets(gdem_diff* d, const double* target) {
  if (!d || !d->committed || !target) return GDEM_ERR_ARG;
  DHIP(hipSetDevice(d->dev));
  const size_t m = (size_t)d->ntot * (size_t)d->nobs;
  std::vector<double4> h(m);
  for (size_t i = 0; i < m; ++i) h[i] = make_double4(target[3 * i], target[3 * i + 1], target[3 * i + 2], 0.);
  DHIP(hipMemcpy(d->d_target, h.data(), sizeof(double4) * m, hipMemcpyHostToDevice));
  return GDEM_OK;
}

static int diff_errors(gdem_diff* d) {
  const size_t ns = d->sys.size();
  std::vector<unsigned long long> h(ns * kErrWords);
  DHIP(hipMemcpy(h.data(), d->d_err, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
  for (size_t s = 0; s < ns; ++s) {
    const unsigned long long* e = &h[s * kErrWords];
    int code = GDEM_OK, id0 = -1, id1 = -1;
    if (e[kErrMotion] != kNoErr) { code = GDEM_ERR_MOTION; id0 = (int32_t)e[kErrMotion]; }
    else if (e[kErrSim] != kNoErr) { code = GDEM_ERR_SIMULATION; id0 = (int32_t)e[kErrSim]; }
    else if (e[kErrContact] != kNoErr) {
      code = GDEM_ERR_CONTACT;
      id0 = (int32_t)(e[kErrContact] >> 32);
      id1 = (int32_t)(uint32_t)(e[kErrContact] & 0xffffffffull);
    } else if (e[kErrCapacity] != kNoErr) { code = GDEM_ERR_CAPACITY; id0 = (int32_t)e[kErrCapacity]; }
    if (code != GDEM_OK) {
      d->err_sys = (int32_t)s;
      d->err_code = code;
      d->err_id0 = id0;
      d->err_id1 = id1;
      DHIP(hipMemset(d->d_err, 0xff, sizeof(unsigned long long) * h.size()));
      return code;
    }
  }
  return GDEM_OK;
}

static int diff_launch(gdem_diff* d, int mode, const double* dtheta = nullptr,
                       double* dloss = nullptr) {
  DHIP(hipSetDevice(d->dev));
  const size_t N = (size_t)d->ntot;
  const size_t nck = (size_t)((d->nsteps + d->K - 1) / d->K);
  DiffArgs a{};
  a.sys = d->d_sys;
  a.order = d->d_order;
  a.nsys = (int32_t)d->sys.size();
  a.next = d->d_next;
  a.mode = mode;
  a.nsteps = d->nsteps;
  a.K = d->K;
  a.nobs = d->nobs;
  a.obs = d->d_obs;
  a.dt = d->dt;
  a.ntot = d->ntot;
  a.prop = d->d_prop;
  a.aux = d->d_aux;
  a.x0 = d->d_x0;
  a.v0 = d->d_v0;
  a.w0 = d->d_w0;
  a.x = d->d_x; a.v = d->d_v; a.w = d->d_w; a.F = d->d_F; a.T = d->d_T;
  a.ck_x = d->d_ck;
  a.ck_v = a.ck_x + N * nck;
  a.ck_w = a.ck_v + N * nck;
  a.ck_F = a.ck_w + N * nck;
  a.ck_T = a.ck_F + N * nck;
  a.win_x = d->d_win;
  a.win_v = a.win_x + N * (size_t)d->win_steps;
  a.win_w = a.win_v + N * (size_t)d->win_steps;
  a.target = d->d_target;
  a.xobs = d->d_xobs;
  a.key = d->d_key; a.slot = d->d_slot; a.perm = d->d_perm; a.cnt = d->d_cnt; a.cell_start = d->d_cs;
  a.pos = d->d_pos;
  a.row_rank = d->d_rows;
  a.out_list = d->d_out_list; a.in_list = d->d_in_list;
  a.out_count = d->d_out_count; a.in_count = d->d_in_count;
  a.walls = d->d_walls;
  double4* g = d->d_adj;
  a.ax = g; a.av = g + N; a.aw = g + 2 * N; a.aF = g + 3 * N; a.aT = g + 4 * N;
  a.aFt = g + 5 * N; a.aTt = g + 6 * N; a.gX = g + 7 * N; a.gV = g + 8 * N; a.gW = g + 9 * N;
  a.gv0 = g + 10 * N;
  a.loss = d->d_loss;
  a.gtheta = d->d_gtheta;
  a.dtheta = dtheta;
  a.dloss = dloss;
  a.err = d->d_err;
  int wg = d->workgroups;
  if (wg <= 0) {
    hipDeviceProp_t prop;
    DHIP(hipGetDeviceProperties(&prop, d->dev));
    wg = prop.multiProcessorCount * (1024 / d->block) * 2;
  }
  wg = std::min<int>(wg, (int)d->sys.size());
  (void)hipGetLastError();  // an earlier failed call must not be reported as this launch's
  hipLaunchKernelGGL(k_diff_zero, dim3(1), dim3(1), 0, d->stream, d->d_next);
  switch (d->block) {
#define GDEM_LAUNCH_DIFF(B)                                                        \
  do {                                                                             \
    if (mode == DIFF_TANGENT)                                                      \
      hipLaunchKernelGGL((k_diff<B, true>), dim3(wg), dim3(B), 0, d->stream, a);  \
    else                                                                           \
      hipLaunchKernelGGL((k_diff<B, false>), dim3(wg), dim3(B), 0, d->stream, a); \
  } while (0)
    case 64: GDEM_LAUNCH_DIFF(64); break;
    case 128: GDEM_LAUNCH_DIFF(128); break;
    case 512: GDEM_LAUNCH_DIFF(512); break;
    default: GDEM_LAUNCH_DIFF(256); break;
#undef GDEM_LAUNCH_DIFF
  }
  DHIP(hipGetLastError());
  DHIP(hipStreamSynchronize(d->stream));
  return diff_errors(d);
}

extern "C" int gdem_diff_record(gdem_diff* d, double* xobs) {
  if (!d || !d->committed || !xobs) return GDEM_ERR_ARG;
  int rc = diff_launch(d, DIFF_RECORD);
  if (rc) return rc;
  const size_t m = (size_t)d->ntot * (size_t)d->nobs;
  std::vector<double4> h(m);
  DHIP(hipMemcpy(h.data(), d->d_xobs, sizeof(double4) * m, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < m; ++i) {
    xobs[3 * i] = h[i].x;
    xobs[3 * i + 1] = h[i].y;
    xobs[3 * i + 2] = h[i].z;
  }
  return GDEM_OK;
}

extern "C" int gdem_diff_loss(gdem_diff* d, double* loss) {
  if (!d || !d->committed || !loss) return GDEM_ERR_ARG;
  const int rc = diff_launch(d, DIFF_LOSS);
  DHIP(hipMemcpy(loss, d->d_loss, sizeof(double) * d->sys.size(), hipMemcpyDeviceToHost));
  return rc;
}

extern "C" int gdem_diff_gradient(gdem_diff* d, double* loss, double* grad_theta, double* grad_v0) {
  if (!d || !d->committed) return GDEM_ERR_ARG;
  const int rc = diff_launch(d, DIFF_GRAD);
  if (loss) DHIP(hipMemcpy(loss, d->d_loss, sizeof(double) * d->sys.size(), hipMemcpyDeviceToHost));
  if (grad_theta)
    DHIP(hipMemcpy(grad_theta, d->d_gtheta, sizeof(double) * 3 * d->sys.size(), hipMemcpyDeviceToHost));
  if (grad_v0) {
    const size_t N = (size_t)d->ntot;
    std::vector<double4> h(N);
    DHIP(hipMemcpy(h.data(), d->d_adj + 10 * N, sizeof(double4) * N, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < N; ++i) {
      grad_v0[3 * i] = h[i].x;
      grad_v0[3 * i + 1] = h[i].y;
      grad_v0[3 * i + 2] = h[i].z;
    }
  }
  return rc;
}

extern "C" int gdem_diff_directional(gdem_diff* d, const double* dtheta, const double* dv0,
                                     double* loss, double* dloss) {
  if (!d || !d->committed || !dtheta || !dv0 || !dloss) return GDEM_ERR_ARG;
  DHIP(hipSetDevice(d->dev));
  const size_t N = (size_t)d->ntot, ns = d->sys.size();
  std::vector<double4> h(N);
  for (size_t i = 0; i < N; ++i) h[i] = make_double4(dv0[3 * i], dv0[3 * i + 1], dv0[3 * i + 2], 0.);
  DHIP(hipMemcpy(d->d_adj + 10 * N, h.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
  double* d_dir = nullptr;
  DHIP(hipMalloc(&d_dir, sizeof(double) * ns * 4));
  DHIP(hipMemcpy(d_dir, dtheta, sizeof(double) * ns * 3, hipMemcpyHostToDevice));
  const int rc = diff_launch(d, DIFF_TANGENT, d_dir, d_dir + 3 * ns);
  if (loss) DHIP(hipMemcpy(loss, d->d_loss, sizeof(double) * ns, hipMemcpyDeviceToHost));
  DHIP(hipMemcpy(dloss, d_dir + 3 * ns, sizeof(double) * ns, hipMemcpyDeviceToHost));
  DHIP(hipFree(d_dir));
  return rc;
}

extern "C" int gdem_diff_last_error(gdem_diff* d, int32_t* sys, int32_t* code, int32_t* id0,
                                    int32_t* id1) {
  if (!d) return GDEM_ERR_ARG;
  if (sys) *sys = d->err_sys;
  if (code) *code = d->err_code;
  if (id0) *id0 = d->err_id0;
  if (id1) *id1 = d->err_id1;
  return d->err_code;
}

extern "C" int gdem_diff_info(gdem_diff* d, int32_t* nsys, int64_t* nparticles) {
  if (!d) return GDEM_ERR_ARG;
  if (nsys) *nsys = (int32_t)d->sys.size();
  if (nparticles) *nparticles = d->ntot;
  return GDEM_OK;
}
