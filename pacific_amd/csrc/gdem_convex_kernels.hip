// gdem_convex_kernels.hip -- the DEM step for convex particles (boxes,
// polyhedra) on gfx950.
//
//   detect : the sphere path's half-stencil sweep (13 cells as 5 slot ranges)
//            with the circumscribed-sphere test as single-precision
//            pre-filter; each surviving pair runs the Johnson GJK distance
//            between the crust-shrunk bodies (RigidBodyWithCrust::ClosestPoint,
//            Component/src/RigidBodyWithCrust.cpp:274-348) once, and the
//            contact (point, overlap vector, overlap) is stored in the
//            detecting particle's list with a reference in the partner's
//   force  : one thread per particle evaluates the force of each of its
//            stored contacts with the reference's body order (both partners
//            get bit-identical values), adds the wall contacts (GJK against
//            the wall box), and its tail applies the leap-frog kicks, the
//            drift, the gyroscopic angular update
//            (ParticleKinematics3D.cpp:56-77) and the rotation of the basis
//            (Transform::composeLeftByRotation, Transform.cpp:140-156)
//
// Paths are relative to src/Grains3D/Grains/ of the reference.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gdem_internal.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

constexpr int kCBlock = 256;

__device__ __forceinline__ int aux_material(const Aux& a) { return a.mat & 255; }
__device__ __forceinline__ int aux_shape(const Aux& a) { return a.mat >> 8; }

// Basis rows and crust scaling of one particle.
struct Orient {
  double b[9];
  double sc[3];
};

__device__ __forceinline__ Orient load_orient(const CState& c, int s) {
  const double4 r0 = c.rot0[s], r1 = c.rot1[s], r2 = c.rot2[s];
  Orient o;
  o.b[0] = r0.x; o.b[1] = r0.y; o.b[2] = r0.z; o.sc[0] = r0.w;
  o.b[3] = r1.x; o.b[4] = r1.y; o.b[5] = r1.z; o.sc[1] = r1.w;
  o.b[6] = r2.x; o.b[7] = r2.y; o.b[8] = r2.z; o.sc[2] = r2.w;
  return o;
}

// ---------------------------------------------------------------- detect
// Sweep: every candidate whose circumscribed sphere overlaps this particle's
// (the first test of RigidBodyWithCrust::ClosestPoint) is queued; the GJK runs
// in a second kernel with one lane per candidate, so lanes do not wait on
// each other's narrow phase.
struct CCandidates {
  const CDetectArgs& a;
  int t;
  float4 mef;
  bool have_me;
  v3 X;
  double circ;
  int32_t id;
  int ncand;

  __device__ __forceinline__ CCandidates(const CDetectArgs& a_, int t_, float4 mef_)
      : a(a_), t(t_), mef(mef_), have_me(false), X{}, circ(0.), id(0), ncand(0) {}

  __device__ __forceinline__ void test(const float4 p2f, int u, int sx, int sy, int sz, bool same) {
    const float dxf = (p2f.x + sx * a.sizef[0]) - mef.x;
    const float dyf = (p2f.y + sy * a.sizef[1]) - mef.y;
    const float dzf = (p2f.z + sz * a.sizef[2]) - mef.z;
    const float rsf = mef.w + p2f.w + a.margin;
    if (dxf * dxf + dyf * dyf + dzf * dzf > rsf * rsf) return;
    if (!have_me) {
      const int s = a.perm[t];
      const double4 p = a.in.pos_r[s];
      X = mk(p.x, p.y, p.z);
      circ = p.w;
      id = a.in.aux[s].id;
      have_me = true;
    }
    const int s2 = a.perm[u];
    const double4 p2 = a.in.pos_r[s2];
    const int32_t id2 = a.in.aux[s2].id;
    const double shx = sx * a.grid.size[0], shy = sy * a.grid.size[1], shz = sz * a.grid.size[2];
    const bool first = same ? (id < id2) : true;
    // circumscribed spheres, first body A, second body B moved to A's image
    v3 gab;
    double cA, cB;
    if (first) {
      gab = sub(X, mk(p2.x + shx, p2.y + shy, p2.z + shz));
      cA = circ; cB = p2.w;
    } else {
      gab = sub(mk(p2.x, p2.y, p2.z), mk(X.x + (-shx), X.y + (-shy), X.z + (-shz)));
      cA = p2.w; cB = circ;
    }
    if (!(norm(gab) < cA + cB)) return;
    if (ncand >= kCandCap) {
      atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)id);
      return;
    }
    a.cand[(size_t)t * kCandCap + ncand] = make_int2(u, pack_code(sx, sy, sz, first));
    ++ncand;
  }

  __device__ __forceinline__ void range(int beg, int end, int sx, int sy, int sz, bool same) {
    for (int u = beg; u < end; ++u) test(a.pos[u], u, sx, sy, sz, same);
  }
};

__global__ void __launch_bounds__(kCBlock) k_cdetect(CDetectArgs a) {
  const int t = logical_index(a.nblk8);
  int nc = 0;
  if (t < a.n) {
    const double4 me = a.in.pos_r[a.perm[t]];
    int ijk[3];
    clamped_cell(a.grid, mk(me.x, me.y, me.z), ijk);
    CCandidates det(a, t, a.pos[t]);
    half_stencil_sweep(a.grid, a.row_rank, a.cell_start, t, ijk, det);
    nc = det.ncand;
    // SimpleObstacle::InterAction's InZone test (Component/src/SimpleObstacle.cpp:222-271)
    for (int wi = 0; wi < a.nwalls; ++wi) {
      const Wall& wl = a.walls[wi];
      if (!(fabs(wl.bb_centre[0] - me.x) <= wl.bb_ext[0] + me.w &&
            fabs(wl.bb_centre[1] - me.y) <= wl.bb_ext[1] + me.w &&
            fabs(wl.bb_centre[2] - me.z) <= wl.bb_ext[2] + me.w))
        continue;
      if (nc >= kCandCap) {
        atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)a.in.aux[a.perm[t]].id);
        break;
      }
      a.cand[(size_t)t * kCandCap + nc] = make_int2(wi, kWallCode);
      ++nc;
    }
    a.cand_count[t] = nc;
    a.mask[t] = 0u;
  }
  // queue the candidates: one atomic per wave
  const int lane = threadIdx.x & 63;
  int incl = nc;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  const int total = __shfl(incl, 63, 64);
  int base = 0;
  if (lane == 63 && total > 0) base = atomicAdd(a.nwork, total);
  base = __shfl(base, 63, 64);
  const int off = base + incl - nc;
  for (int j = 0; j < nc; ++j) a.work[off + j] = t * kCandCap + j;
}

// Narrow phase of one candidate: RigidBodyWithCrust::ClosestPoint
// (Component/src/RigidBodyWithCrust.cpp:274-348) between the crust-shrunk
// bodies, first body A at its position, second body B moved to A's image.
// SV: the signed-volume GJK (CollisionDetection Type="GJK_SV") instead of
// the Johnson one.
template <bool SV>
__global__ void __launch_bounds__(kCBlock) k_cgjk(CGjkArgs a) {
  const int nw = *a.nwork;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nw; k += gridDim.x * blockDim.x) {
    const int item = a.work[k];
    const int t = item / kCandCap, j = item - t * kCandCap;
    const int2 cd = a.cand[item];
    const int u = cd.x, code = cd.y;
    if (code == kWallCode) {
      // the particle is the first body, the wall box the second
      const int s1 = a.perm[t];
      const double4 p1 = a.in.pos_r[s1];
      const Aux ax1 = a.in.aux[s1];
      const Orient o1 = load_orient(a.cin, s1);
      const Wall& wl = a.walls[u];
      const double4 wsc = a.wall_sc[u];
      const double wscale[3] = {wsc.x, wsc.y, wsc.z};
      const CBody A{&a.shapes[aux_shape(ax1)], o1.b, mk(p1.x, p1.y, p1.z), o1.sc, ax1.crust, p1.w};
      const CBody B{&a.wall_shapes[u], wl.basis, mk(wl.origin[0], wl.origin[1], wl.origin[2]),
                    wscale, wl.crust, wsc.w};
      v3 c = mk(0., 0., 0.), ov = mk(0., 0., 0.);
      double dist = 0.;
      int nit = 0;  // ClosestPointRECTANGLE reports none
      int rc = wl.kind == WALL_RECTANGLE ? convex_rect(A, wl, c, ov, dist)
                                         : closest_point<SV>(A, B, c, ov, dist, nit);
      if (rc < 0) rc = closest_point_retry(A, B, c, ov, dist, nit);
      if (rc < 0) {
        atomicMin(&a.err[kErrContact], pack_ids(ax1.id, -1 - u));
        continue;
      }
      if (a.stats) {
        CRec& r = a.rec[item];
        r.dist = dist;
        r.nit = nit;
        r.evaluated = rc != 0;
      }
      if (rc == 0 || !(dist < 0.)) continue;
      CRec& r = a.rec[item];
      r.c[0] = c.x; r.c[1] = c.y; r.c[2] = c.z;
      r.ov[0] = ov.x; r.ov[1] = ov.y; r.ov[2] = ov.z;
      r.dist = dist;
      atomicOr(&a.mask[t], 1u << j);
      continue;
    }
    const int sx = (code & 3) - 1, sy = ((code >> 2) & 3) - 1, sz = ((code >> 4) & 3) - 1;
    const bool first = (code & 64) != 0;
    const double shx = sx * a.grid.size[0], shy = sy * a.grid.size[1], shz = sz * a.grid.size[2];
    const int s1 = a.perm[t], s2 = a.perm[u];
    const double4 p1 = a.in.pos_r[s1], p2 = a.in.pos_r[s2];
    const Aux ax1 = a.in.aux[s1], ax2 = a.in.aux[s2];
    const Orient o1 = load_orient(a.cin, s1), o2 = load_orient(a.cin, s2);
    CBody A, B;
    if (first) {
      A = CBody{&a.shapes[aux_shape(ax1)], o1.b, mk(p1.x, p1.y, p1.z), o1.sc, ax1.crust, p1.w};
      B = CBody{&a.shapes[aux_shape(ax2)], o2.b, mk(p2.x + shx, p2.y + shy, p2.z + shz), o2.sc,
                ax2.crust, p2.w};
    } else {
      A = CBody{&a.shapes[aux_shape(ax2)], o2.b, mk(p2.x, p2.y, p2.z), o2.sc, ax2.crust, p2.w};
      B = CBody{&a.shapes[aux_shape(ax1)], o1.b, mk(p1.x + (-shx), p1.y + (-shy), p1.z + (-shz)),
                o1.sc, ax1.crust, p1.w};
    }
    v3 c, ov;
    double dist = 0.;
    int nit = 0;
    int rc = closest_point<SV>(A, B, c, ov, dist, nit);
    if (rc < 0) rc = closest_point_retry(A, B, c, ov, dist, nit);
    if (rc < 0) {
      atomicMin(&a.err[kErrContact], first ? pack_ids(ax1.id, ax2.id) : pack_ids(ax2.id, ax1.id));
      continue;
    }
    if (a.stats) {
      CRec& r = a.rec[item];
      r.dist = dist;
      r.nit = nit;
      r.evaluated = rc != 0;
    }
    if (rc == 0 || !(dist < 0.)) continue;
    CRec& r = a.rec[item];
    r.c[0] = c.x; r.c[1] = c.y; r.c[2] = c.z;
    r.ov[0] = ov.x; r.ov[1] = ov.y; r.ov[2] = ov.z;
    r.dist = dist;
    atomicOr(&a.mask[t], 1u << j);
    const int kin = atomicAdd(&a.in_count[u], 1);
    if (kin < kCandCap) {
      a.in_list[(size_t)u * kCandCap + kin] = make_int2(t, j);
    } else {
      atomicMin(&a.err[kErrCapacity], (unsigned long long)(uint32_t)ax2.id);
    }
  }
}

// ---------------------------------------------------------------- force
template <bool RECORD>
struct CReducer {
  const CForceArgs& a;
  int t;
  v3 X, V, W;
  double m, req;
  Aux ax;
  v3 F, T;
  int ncont, nrec;

  // One stored contact.  det_me: this particle detected it (else the partner
  // slot `other` did).  The candidate's code gives the shift applied to the
  // detected particle and whether the detecting particle is the first body.
  __device__ __forceinline__ void contact(const CRec& r, int code, int other, bool det_me) {
    const int sx = (code & 3) - 1, sy = ((code >> 2) & 3) - 1, sz = ((code >> 4) & 3) - 1;
    const bool first_det = (code & 64) != 0;
    const double shx = sx * a.grid.size[0], shy = sy * a.grid.size[1], shz = sz * a.grid.size[2];
    const int s2 = a.perm[other];
    const double4 p2 = a.in.pos_r[s2];
    const double4 vm2 = a.in.vel_m[s2];
    const double4 wr2 = a.in.om_req[s2];
    const Aux ax2 = a.in.aux[s2];
    const v3 X2 = mk(p2.x, p2.y, p2.z);
    const v3 V2 = mk(vm2.x, vm2.y, vm2.z), W2 = mk(wr2.x, wr2.y, wr2.z);
    bool meA;
    v3 xA, xB;
    if (det_me) {
      meA = first_det;
      if (meA) { xA = X; xB = mk(X2.x + shx, X2.y + shy, X2.z + shz); }
      else { xA = X2; xB = mk(X.x + (-shx), X.y + (-shy), X.z + (-shz)); }
    } else {
      meA = !first_det;
      if (meA) { xA = X; xB = mk(X2.x + (-shx), X2.y + (-shy), X2.z + (-shz)); }
      else { xA = X2; xB = mk(X.x + shx, X.y + shy, X.z + shz); }
    }
    const v3 c = mk(r.c[0], r.c[1], r.c[2]);
    const v3 ov = mk(r.ov[0], r.ov[1], r.ov[2]);
    const int matA = meA ? aux_material(ax) : aux_material(ax2);
    const int matB = meA ? aux_material(ax2) : aux_material(ax);
    const Model md = a.models[matA * kMaxMaterials + matB];
    v3 f;
    if (meA)
      f = contact_force(md, c, ov, r.dist, xA, V, W, m, req, xB, V2, W2, vm2.w, wr2.w);
    else
      f = contact_force(md, c, ov, r.dist, xA, V2, W2, vm2.w, wr2.w, xB, V, W, m, req);
    if (meA) {
      F = add(F, f);
      T = add(T, cross(sub(c, xA), f));
      ++ncont;
      if (RECORD) record(ax.id, ax2.id, c, ov, r.dist, f);
    } else {
      const v3 fn = neg(f);
      F = add(F, fn);
      T = add(T, cross(sub(c, xB), fn));
    }
  }

  // Contact with wall wi (SimpleObstacle::InterAction, SimpleObstacle.cpp:222-271)
  __device__ __forceinline__ void wall(const CRec& r, int wi) {
    const Wall& wl = a.walls[wi];
    const v3 c = mk(r.c[0], r.c[1], r.c[2]);
    const v3 ov = mk(r.ov[0], r.ov[1], r.ov[2]);
    const Model md = a.models[aux_material(ax) * kMaxMaterials + wl.material];
    const v3 zero = mk(0., 0., 0.);
    const v3 f = contact_force(md, c, ov, r.dist, X, V, W, m, req,
                               mk(wl.origin[0], wl.origin[1], wl.origin[2]), zero, zero, 1.e20,
                               wl.req);
    F = add(F, f);
    T = add(T, cross(sub(c, X), f));
    ++ncont;
    if (RECORD) record(ax.id, -1 - wi, c, ov, r.dist, f);
  }

  __device__ __forceinline__ void record(int id0, int id1, v3 c, v3 ov, double dist, v3 f) {
    if (nrec < a.rec_cap) {
      ContactRec& cr = a.crec[(size_t)t * a.rec_cap + nrec];
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

// ParticleKinematics::Move's rotation part (ParticleKinematics.cpp:80-101):
// quaternion increment from the mean angular velocity over the drift.
__device__ __forceinline__ quat rotation_increment(v3 avg, double dtd) {
  const double nOmega = norm(avg);
  if (nOmega > kEps) {
    const double cc = cos(nOmega * dtd / 2.);
    const double sn = sin(nOmega * dtd / 2.);
    return qmk(scl(avg, sn * 1. / nOmega), cc);
  }
  return qmk(mk(0., 0., 0.), 1.);
}

__device__ __forceinline__ void wrap_periodic(const Grid& g, v3& X) {
  double xo[3] = {X.x, X.y, X.z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!g.periodic[d]) continue;
    const double lo = g.origin[d], L = g.size[d];
    if (xo[d] >= lo + L) xo[d] -= L;
    else if (xo[d] < lo) xo[d] += L;
  }
  X = mk(xo[0], xo[1], xo[2]);
}

template <int MODE, bool RECORD>
__global__ void __launch_bounds__(kCBlock) k_cforce(CForceArgs a) {
  __shared__ unsigned long long sh_cnt[kCBlock / 64];
  __shared__ StatPart sh_stat[kCBlock / 64];
  const int t = logical_index(a.nblk8);
  int ncont = 0;
  StatAcc st;
  if (t < a.n) {
    if (a.stat) {
      // AppCollision::addToContactsFeatures of every closest point this
      // particle's candidates ran (its own detections, in candidate order)
      const int nc = min(a.cand_count[t], kCandCap);
      for (int j = 0; j < nc; ++j) {
        const CRec& r = a.rec[(size_t)t * kCandCap + j];
        if (r.evaluated) st.add(r.dist, r.nit);
      }
    }
    const int s = a.perm[t];
    const double4 pr = a.in.pos_r[s];
    const double4 vm = a.in.vel_m[s];
    const double4 wr = a.in.om_req[s];
    const double4 wb = a.in.wbf_I[s];
    const double4 qq = a.in.quat[s];
    const double4 in3 = a.cin.inert[s];
    CReducer<RECORD> R{a, t};
    R.ax = a.in.aux[s];
    R.X = mk(pr.x, pr.y, pr.z);
    R.V = mk(vm.x, vm.y, vm.z);
    R.W = mk(wr.x, wr.y, wr.z);
    R.m = vm.w;
    R.req = wr.w;
    const double circ = pr.w;
    // InitializeForce: weight m (1 - 0/rho) g (Particle.cpp:414-419, 1409-1416)
    const double mw = R.m * (1. - 0. / 1.);
    R.F = mk(a.gravity[0] * mw, a.gravity[1] * mw, a.gravity[2] * mw);
    R.T = mk(0., 0., 0.);
    R.ncont = 0;
    R.nrec = 0;
    if (t == 0) *a.nwork = 0;  // the candidate queue of the next step starts empty
    // own contacts in sweep order, then the received ones by increasing
    // (detecting slot, candidate index)
    uint32_t mk_bits = a.mask[t];
    while (mk_bits) {
      const int j = __ffs(mk_bits) - 1;
      mk_bits &= mk_bits - 1u;
      const int2 cd = a.cand[(size_t)t * kCandCap + j];
      if (cd.y == kWallCode)
        R.wall(a.rec[(size_t)t * kCandCap + j], cd.x);
      else
        R.contact(a.rec[(size_t)t * kCandCap + j], cd.y, cd.x, true);
    }
    const int nin = min(a.in_count[t], kCandCap);
    a.in_count[t] = 0;
    const int2* il = a.in_list + (size_t)t * kCandCap;
    long long last = -1;
    for (int e = 0; e < nin; ++e) {
      long long best = 0x7fffffffffffffffll;
      for (int f = 0; f < nin; ++f) {
        const int2 en = il[f];
        const long long k = ((long long)en.x << 8) | (long long)en.y;
        if (k > last && k < best) best = k;
      }
      if (best == 0x7fffffffffffffffll) break;
      const int src = (int)(best >> 8), idx = (int)(best & 255);
      const size_t item = (size_t)src * kCandCap + idx;
      R.contact(a.rec[item], a.cand[item].y, src, false);
      last = best;
    }
    Orient me = load_orient(a.cin, s);
    if (RECORD) a.rec_count[t] = R.nrec;
    ncont = R.ncont;
    const v3 F = R.F, T = R.T, X = R.X, V = R.V, W = R.W;
    const double m = R.m;

    // ---- tail
    v3 Xo = X, Vo = V, Wo = W;
    v3 WBF = mk(wb.x, wb.y, wb.z);
    quat q = qmk(mk(qq.x, qq.y, qq.z), qq.w);
    int ijk[3];
    clamped_cell(a.grid, X, ijk);
    int key = row_base(a.grid, a.row_rank, ijk[1], ijk[2]) + ijk[0];
    if (MODE != MODE_FORCE) {
      const quat pc = qconj(q);
      const v3 tq = qmulv(pc, qleftvec(q, T));
      const double mcf = m * 1.;
      // closing half kick (SecondOrderLeapFrog.cpp:74-94)
      Vo = add(V, scl(dvd(F, mcf), a.dt_vel));
      WBF = ssprk3_3d(tq, WBF, in3.x, in3.y, in3.z, a.dt_vel);
      Wo = qmulv(q, qleftvec(pc, WBF));
      if (MODE == MODE_MOVE) {
        // next step's kick and drift (SecondOrderLeapFrog.cpp:44-67,
        // ParticleKinematics.cpp:62-104); q and T unchanged -> same tq
        Vo = add(Vo, scl(dvd(F, mcf), a.dt_vel));
        const v3 dx = scl(Vo, a.dt_disp);
        WBF = ssprk3_3d(tq, WBF, in3.x, in3.y, in3.z, a.dt_vel);
        Wo = qmulv(q, qleftvec(pc, WBF));
        Xo = add(X, dx);
        const quat dq = rotation_increment(Wo, a.dt_disp);
        q = qmul(dq, q);
        compose_left_rotation(me.b, dq);
        if (norm(dx) > R.ax.crust)
          atomicMin(&a.err[kErrMotion], (unsigned long long)(uint32_t)R.ax.id);
        wrap_periodic(a.grid, Xo);
        int nijk[3];
        if (cell_of(a.grid, Xo, nijk)) {
          key = row_base(a.grid, a.row_rank, nijk[1], nijk[2]) + nijk[0];
        } else {
          key = 0;
          atomicMin(&a.err[kErrSim], (unsigned long long)(uint32_t)R.ax.id);
        }
      }
    }
    a.out.pos_r[t] = make_double4(Xo.x, Xo.y, Xo.z, circ);
    a.out.vel_m[t] = make_double4(Vo.x, Vo.y, Vo.z, m);
    a.out.om_req[t] = make_double4(Wo.x, Wo.y, Wo.z, R.req);
    a.out.wbf_I[t] = make_double4(WBF.x, WBF.y, WBF.z, wb.w);
    a.out.quat[t] = make_double4(q.v.x, q.v.y, q.v.z, q.w);
    a.out.aux[t] = R.ax;
    a.cout.rot0[t] = make_double4(me.b[0], me.b[1], me.b[2], me.sc[0]);
    a.cout.rot1[t] = make_double4(me.b[3], me.b[4], me.b[5], me.sc[1]);
    a.cout.rot2[t] = make_double4(me.b[6], me.b[7], me.b[8], me.sc[2]);
    a.cout.inert[t] = in3;
    a.key_out[t] = key;
    if (a.F) {
      a.F[t] = make_double4(F.x, F.y, F.z, 0.);
      a.T[t] = make_double4(T.x, T.y, T.z, 0.);
    }
  }
  unsigned long long c = (unsigned long long)ncont;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) sh_cnt[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kCBlock / 64; ++w) tot += sh_cnt[w];
    if (tot) atomicAdd(a.contacts, tot);
  }
  if (a.stat) block_stats<kCBlock>(st, a.stat, sh_stat);
}

// ---------------------------------------------------------------- integrator phases
__global__ void __launch_bounds__(kCBlock) k_cmove(State s, CState cs, const double4* Fa,
                                                   const double4* Ta, int n, Grid g,
                                                   const int32_t* row_rank, double dtv,
                                                   double dtd, int32_t* key,
                                                   unsigned long long* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pr = s.pos_r[i], vm = s.vel_m[i], wr = s.om_req[i], wb = s.wbf_I[i],
                qq = s.quat[i], in3 = cs.inert[i];
  const double4 f4 = Fa[i], t4 = Ta[i];
  const Aux ax = s.aux[i];
  const v3 F = mk(f4.x, f4.y, f4.z), T = mk(t4.x, t4.y, t4.z);
  quat q = qmk(mk(qq.x, qq.y, qq.z), qq.w);
  const quat pc = qconj(q);
  const v3 tq = qmulv(pc, qleftvec(q, T));
  const double mcf = vm.w * 1.;
  const v3 V = add(mk(vm.x, vm.y, vm.z), scl(dvd(F, mcf), dtv));
  const v3 dx = scl(V, dtd);
  const v3 WBF = ssprk3_3d(tq, mk(wb.x, wb.y, wb.z), in3.x, in3.y, in3.z, dtv);
  const v3 W = qmulv(q, qleftvec(pc, WBF));
  v3 X = add(mk(pr.x, pr.y, pr.z), dx);
  const quat dq = rotation_increment(W, dtd);
  q = qmul(dq, q);
  Orient o = load_orient(cs, i);
  compose_left_rotation(o.b, dq);
  if (norm(dx) > ax.crust) atomicMin(&err[kErrMotion], (unsigned long long)(uint32_t)ax.id);
  wrap_periodic(g, X);
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
  cs.rot0[i] = make_double4(o.b[0], o.b[1], o.b[2], o.sc[0]);
  cs.rot1[i] = make_double4(o.b[3], o.b[4], o.b[5], o.sc[1]);
  cs.rot2[i] = make_double4(o.b[6], o.b[7], o.b[8], o.sc[2]);
}

__global__ void __launch_bounds__(kCBlock) k_cadvance(State s, CState cs, const double4* Fa,
                                                      const double4* Ta, int n, double dtv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 vm = s.vel_m[i], wb = s.wbf_I[i], qq = s.quat[i], in3 = cs.inert[i];
  const double4 f4 = Fa[i], t4 = Ta[i];
  const v3 F = mk(f4.x, f4.y, f4.z), T = mk(t4.x, t4.y, t4.z);
  const quat q = qmk(mk(qq.x, qq.y, qq.z), qq.w);
  const quat pc = qconj(q);
  const v3 tq = qmulv(pc, qleftvec(q, T));
  const double mcf = vm.w * 1.;
  const v3 V = add(mk(vm.x, vm.y, vm.z), scl(dvd(F, mcf), dtv));
  const v3 WBF = ssprk3_3d(tq, mk(wb.x, wb.y, wb.z), in3.x, in3.y, in3.z, dtv);
  const v3 W = qmulv(q, qleftvec(pc, WBF));
  const double req = s.om_req[i].w;
  s.vel_m[i] = make_double4(V.x, V.y, V.z, vm.w);
  s.wbf_I[i] = make_double4(WBF.x, WBF.y, WBF.z, wb.w);
  s.om_req[i] = make_double4(W.x, W.y, W.z, req);
}

__global__ void __launch_bounds__(kCBlock) k_cunpermute(State s, CState cs, int n,
                                                        double* basis) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t id = (size_t)s.aux[i].id;
  const double4 r0 = cs.rot0[i], r1 = cs.rot1[i], r2 = cs.rot2[i];
  double* b = basis + 9 * id;
  b[0] = r0.x; b[1] = r0.y; b[2] = r0.z;
  b[3] = r1.x; b[4] = r1.y; b[5] = r1.z;
  b[6] = r2.x; b[7] = r2.y; b[8] = r2.z;
}

// ---------------------------------------------------------------- launchers
static inline int cblocks(int n) { return (n + kCBlock - 1) / kCBlock; }

void launch_cdetect(const CDetectArgs& a, hipStream_t st) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_cdetect, dim3(a.nblk8), dim3(kCBlock), 0, st, a);
}

void launch_cgjk(const CGjkArgs& a, int32_t max_work, hipStream_t st) {
  if (max_work <= 0) return;
  // grid-stride over the queue; the grid covers a typical queue in one pass
  const int blocks = std::min(std::max(cblocks(max_work), 1), 16384);
  if (a.gjk_sv)
    hipLaunchKernelGGL(k_cgjk<true>, dim3(blocks), dim3(kCBlock), 0, st, a);
  else
    hipLaunchKernelGGL(k_cgjk<false>, dim3(blocks), dim3(kCBlock), 0, st, a);
}

void launch_cforce(int mode, bool record, const CForceArgs& a, hipStream_t st) {
  if (a.n <= 0) return;
  dim3 grid(a.nblk8), block(kCBlock);
#define GDEM_LAUNCH_CFORCE(M, R) hipLaunchKernelGGL((k_cforce<M, R>), grid, block, 0, st, a)
  if (record) {
    if (mode == MODE_FORCE) GDEM_LAUNCH_CFORCE(MODE_FORCE, true);
    else if (mode == MODE_KICK) GDEM_LAUNCH_CFORCE(MODE_KICK, true);
    else GDEM_LAUNCH_CFORCE(MODE_MOVE, true);
  } else {
    if (mode == MODE_FORCE) GDEM_LAUNCH_CFORCE(MODE_FORCE, false);
    else if (mode == MODE_KICK) GDEM_LAUNCH_CFORCE(MODE_KICK, false);
    else GDEM_LAUNCH_CFORCE(MODE_MOVE, false);
  }
#undef GDEM_LAUNCH_CFORCE
}

void launch_cmove(const State& s, const CState& c, const double4* F, const double4* T, int32_t n,
                  const Grid& g, const int32_t* row_rank, double dt_vel, double dt_disp,
                  int32_t* key, unsigned long long* err, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cmove, dim3(cblocks(n)), dim3(kCBlock), 0, st, s, c, F, T, n, g, row_rank,
                     dt_vel, dt_disp, key, err);
}

void launch_cadvance(const State& s, const CState& c, const double4* F, const double4* T,
                     int32_t n, double dt_vel, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cadvance, dim3(cblocks(n)), dim3(kCBlock), 0, st, s, c, F, T, n, dt_vel);
}

void launch_cunpermute(const State& s, const CState& c, int32_t n, double* basis, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cunpermute, dim3(cblocks(n)), dim3(kCBlock), 0, st, s, c, n, basis);
}

}  // namespace gdem
