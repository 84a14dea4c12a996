// gdem_kernels.hip -- CDNA4 (gfx950) kernels of the DEM step.
//
//   binning : keys -> count (wave-aggregated int atomics) -> block scan ->
//             scatter -> per-cell sort by previous slot -> sorted position copy
//             (deterministic counting sort; replaces LinkedCell::LinkUpdate,
//             Base/src/LinkedCell.cpp:1452-1547)
//   detect  : one thread per particle in cell order sweeps the reference's
//             13-cell half stencil (LinkedCell.cpp:958-984) as 5 contiguous
//             slot ranges (cells of a row along x are consecutive) and appends
//             each contact to both partners' lists
//   force   : one thread per particle reduces its own contact segment in a
//             fixed order -- every pair force is evaluated with the
//             reference's body order so both partners get bit-identical
//             values, no floating-point atomics -- then its tail applies the
//             leap-frog kicks and drift (SecondOrderLeapFrog.cpp:44-94)
//   move / advance : stand-alone integrator phases for the per-phase ABI.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "gdem_binning.h"
#include "gdem_internal.h"
#include "gdem_sphere_step.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;
static_assert(kScanCells == kScanBlock * kScanPer, "scan block geometry");
// Workgroups of the binning kernels when a gate word may skip them: a
// grid-stride grid this size exits in about a microsecond when no rebuild is
// due, where a full-size grid costs several.
constexpr int kGateGrid = 1024;

// ---------------------------------------------------------------- keys
__global__ void __launch_bounds__(kBlock) k_keys(State s, int n, Grid g, const int32_t* row_rank,
                                                 int32_t* key, unsigned long long* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = s.pos_r[i];
  int ijk[3];
  if (!cell_of(g, mk(p.x, p.y, p.z), ijk)) {
    atomicMin(&err[kErrSim], (unsigned long long)(uint32_t)s.aux[i].id);
    key[i] = 0;
    return;
  }
  key[i] = row_base(g, row_rank, ijk[1], ijk[2]) + ijk[0];
}

// ---------------------------------------------------------------- binning
// Arrival slots of the count pass (count_pass, gdem_binning.h); with skin
// lists (words non-null) block 0 also does the evaluation's first-kernel
// duties on the skin words.
__global__ void __launch_bounds__(kBlock) k_count(int32_t* key, int32_t* slot, int32_t* cnt,
                                                  int n, SkinGate gate, SkinWords* words,
                                                  int32_t* in_count, PosSrc src, Grid g,
                                                  const int32_t* row_rank,
                                                  unsigned long long* err) {
  const bool rebuild = gate.rebuild();
  if (words && blockIdx.x == 0 && threadIdx.x == 0) skin_words_duty(gate, words, rebuild);
  if (!rebuild) return;
  count_pass(key, slot, cnt, n, words != nullptr, in_count, src, g, row_rank, err);
}

template <int BS>
__device__ __forceinline__ int block_exclusive_scan(int v, int* smem, int& total) {
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
  int prefix = wid > 0 ? smem[wid - 1] : 0;
  total = smem[BS / 64 - 1];
  return prefix + x - v;
}

// Exclusive scan of the cell counts in one launch: each block scans its
// kScanCells cells (kScanPer consecutive cells per thread) and publishes its
// total; the last block to finish (ticket) scans the totals in place
// (nb <= 8 * kScanBlock) and resets the ticket.  No block waits for another.
__global__ void __launch_bounds__(kScanBlock) k_scan(int32_t* cnt, int32_t* local, int32_t* bsum,
                                                     int32_t* ticket, int ncell, int nb,
                                                     SkinGate gate) {
  if (!gate.rebuild()) return;
  __shared__ int smem[kScanBlock / 64];
  __shared__ int last;
  const int c0 = blockIdx.x * kScanCells + threadIdx.x * kScanPer;
  int v[kScanPer];
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = c0 + k < ncell ? cnt[c0 + k] : 0;
    sum += v[k];
  }
  int total;
  int ex = block_exclusive_scan<kScanBlock>(sum, smem, total);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (c0 + k < ncell) {
      local[c0 + k] = ex;
      cnt[c0 + k] = 0;  // ready for the next count pass (no memset node in graphs)
    }
    ex += v[k];
  }
  if (threadIdx.x == 0) {
    bsum[blockIdx.x] = total;
    __threadfence();  // the total is visible device-wide before the ticket
    last = atomicAdd(ticket, 1) == nb - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // every block's total is visible to this block
  const int per = (nb + kScanBlock - 1) / kScanBlock;
  const int beg = threadIdx.x * per;
  int loc[8];
  int s = 0;
  for (int k = 0; k < per; ++k) {
    int idx = beg + k;
    loc[k] = idx < nb ? bsum[idx] : 0;
    s += loc[k];
  }
  int bex = block_exclusive_scan<kScanBlock>(s, smem, total);
  for (int k = 0; k < per; ++k) {
    int idx = beg + k;
    if (idx < nb) bsum[idx] = bex;
    bex += loc[k];
  }
  if (threadIdx.x == 0) *ticket = 0;
}

__global__ void __launch_bounds__(kBlock) k_scatter(const int32_t* key, const int32_t* slot,
                                                    const int32_t* local, const int32_t* bsum,
                                                    int32_t* perm, int n, SkinGate gate) {
  if (!gate.rebuild()) return;
  scatter_pass(key, slot, local, bsum, perm, n);
}

// One cell of k_cellsort.
__device__ __forceinline__ void sort_cell(const int32_t* local, const int32_t* bsum, int32_t* perm,
                                          int32_t* cell_start, const PosSrc& src,
                                          float4* pos_sorted, int32_t* ids_sorted, int c,
                                          int ncell, int n) {
  int beg = local[c] + bsum[c / kScanCells];
  int end = (c + 1 < ncell) ? local[c + 1] + bsum[(c + 1) / kScanCells] : n;
  for (int a = beg + 1; a < end; ++a) {
    int v = perm[a];
    int b = a - 1;
    while (b >= beg && perm[b] > v) {
      perm[b + 1] = perm[b];
      --b;
    }
    perm[b + 1] = v;
  }
  cell_start[c] = beg;
  if (c == ncell - 1) cell_start[ncell] = n;
  for (int u = beg; u < end; ++u) {
    const int s = perm[u];
    const double4 p = src.pos(s);
    pos_sorted[u] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
    if (ids_sorted) ids_sorted[u] = src.aux[s].id;
  }
}

// Orders each cell by previous slot, which makes the permutation independent of
// the order in which the count pass's atomics were served, then writes the
// cell's entries of the sorted single-precision copy of position + radius
// (read by the detection sweep's conservative pre-filter, which alone decides
// skin-list candidacy: every contact decision is made in FP64 on the
// full-precision state), and with skin lists the sorted ids.
__global__ void __launch_bounds__(kBlock) k_cellsort(const int32_t* local, const int32_t* bsum,
                                                     int32_t* perm, int32_t* cell_start,
                                                     PosSrc src, float4* pos_sorted,
                                                     int32_t* ids_sorted, int ncell, int n,
                                                     SkinGate gate) {
  if (!gate.rebuild()) return;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x)
    sort_cell(local, bsum, perm, cell_start, src, pos_sorted, ids_sorted, c, ncell, n);
}

// ---------------------------------------------------------------- detect
__global__ void __launch_bounds__(kBlock) k_detect(DetectArgs a) {
  if (!a.gate.rebuild()) return;
  const int t = logical_index(a.nblk8);
  if (t >= a.n) return;
  detect_particle(a, t);
}

// Skin lists: the candidate sweep when the lists are rebuilt (need word set),
// otherwise the screening of the existing lists for this evaluation (the force
// kernel walks every candidate after a rebuild).
__global__ void __launch_bounds__(kBlock) k_detect_screen(DetectArgs d, ForceArgs f) {
  const int t = logical_index(d.nblk8);
  if (t >= d.n) return;
  if (d.gate.rebuild())
    detect_particle(d, t);
  else
    screen_particle(f, t);
}

// ---------------------------------------------------------------- force
template <int MODE, bool RECORD, bool MEM>
__global__ void __launch_bounds__(kBlock) k_force(ForceArgs a) {
  __shared__ unsigned long long sh_cnt[kBlock / 64];
  const int t = logical_index(a.nblk8);
  __shared__ StatPart sh_stat[kBlock / 64];
  int ncont = 0;
  double drift = 0.;
  StatAcc st;
  if (t < a.n) ncont = force_particle<MODE, RECORD, MEM>(a, t, drift, st);
  if (a.stat) block_stats<kBlock>(st, a.stat, sh_stat);
  if (a.words) {
    // skin lists: path length since the rebuild, need word of the next
    // evaluation (SkinWords)
    const bool rebuilt = a.gate.rebuild();
    if (t < a.n && (MODE == MODE_MOVE || rebuilt)) {
      const float len = skin_path(rebuilt ? 0.f : a.path[t], rebuilt, drift);
      a.path[t] = len;
      if (len > a.gate.limit) atomicOr(&a.words->need[1 - a.gate.p], 1);
    }
    if (t == 0) a.words->rebuilt[1 - a.gate.p] = rebuilt ? 1 : 0;
  }
  // contact counter: wave sum, then one integer atomic per workgroup
  unsigned long long c = (unsigned long long)ncont;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) sh_cnt[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += sh_cnt[w];
    if (tot) a.contacts[blockIdx.x] += tot;  // this workgroup's slot: no atomics
  }
}

// ---------------------------------------------------------------- wall loads
// Load of every wall from one force evaluation (AllComponents::computeObstaclesLoad,
// Component/src/AllComponents.cpp:1685-1693, of the walls' torsors): each
// block sums the wall contacts of its kBlock particles in slot order, thread w
// for wall w; one block per wall then adds the block partials with a fixed
// tree.  Deterministic, no atomics.
__global__ void __launch_bounds__(kBlock) k_wall_load_partial(const int32_t* count,
                                                              const int32_t* wall,
                                                              const double4* f, const double4* tq,
                                                              int n, int lstride, int nwalls,
                                                              double4* part) {
  __shared__ int32_t s_cnt[kBlock];
  const int b0 = blockIdx.x * kBlock;
  const int i = b0 + threadIdx.x;
  s_cnt[threadIdx.x] = i < n ? count[i] : 0;
  __syncthreads();
  const int w = threadIdx.x;
  if (w >= nwalls) return;
  v3 F = mk(0., 0., 0.), T = mk(0., 0., 0.);
  for (int j = 0; j < kBlock; ++j) {
    const int nc = s_cnt[j];
    for (int k = 0; k < nc; ++k) {
      const size_t at = (size_t)k * lstride + b0 + j;
      if (wall[at] != w) continue;
      const double4 fa = f[at], ta = tq[at];
      F = add(F, mk(fa.x, fa.y, fa.z));
      T = add(T, mk(ta.x, ta.y, ta.z));
    }
  }
  const size_t o = ((size_t)blockIdx.x * nwalls + w) * 2;
  part[o] = make_double4(F.x, F.y, F.z, 0.);
  part[o + 1] = make_double4(T.x, T.y, T.z, 0.);
}

__global__ void __launch_bounds__(256) k_wall_load_final(const double4* part, int nblk, int nwalls,
                                                         double4* out) {
  __shared__ double s[6][256];
  const int w = blockIdx.x, tid = threadIdx.x;
  double acc[6] = {0., 0., 0., 0., 0., 0.};
  for (int b = tid; b < nblk; b += 256) {
    const size_t o = ((size_t)b * nwalls + w) * 2;
    const double4 fa = part[o], ta = part[o + 1];
    acc[0] += fa.x; acc[1] += fa.y; acc[2] += fa.z;
    acc[3] += ta.x; acc[4] += ta.y; acc[5] += ta.z;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k][tid] = acc[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h)
#pragma unroll
      for (int k = 0; k < 6; ++k) s[k][tid] += s[k][tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    out[2 * w] = make_double4(s[0][0], s[1][0], s[2][0], 0.);
    out[2 * w + 1] = make_double4(s[3][0], s[4][0], s[5][0], 0.);
  }
}

// ---------------------------------------------------------------- integrator phases
__global__ void __launch_bounds__(kBlock) k_move(State s, const double4* Fa, const double4* Ta,
                                                 int n, Grid g, const int32_t* row_rank,
                                                 double dtv, double dtd, int32_t* key,
                                                 unsigned long long* err, SkinGate gate,
                                                 float* path) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double drift = move_particle(s, Fa, Ta, i, g, row_rank, dtv, dtd, key, err);
  if (path) {
    // a move in place of copy p: the next evaluation reads need[p]
    const float len = skin_path(path[i], false, drift);
    path[i] = len;
    if (len > gate.limit) atomicOr(const_cast<int32_t*>(&gate.w->need[gate.p]), 1);
  }
}

// Running totals of the contact statistics from the parts of one evaluation,
// added in workgroup order; the evaluation's time is st->time + dt_step
// (Grains::Simulation advances the time before the step's force evaluation,
// Base/src/Grains.cpp:290).  The largest overlap keeps the first time it was
// reached (AppCollision.cpp:175-179).
__global__ void __launch_bounds__(256) k_stats_final(const StatPart* part, int nparts,
                                                    ContactStats* st, double dt_step) {
  // thread k adds a contiguous range of parts in order, then a fixed tree
  __shared__ StatPart sh[256];
  const int k = threadIdx.x;
  const int per = (nparts + 255) / 256;
  StatPart a{0., 0., 0ull, 0ull, 0ull};
  for (int b = k * per; b < min(nparts, (k + 1) * per); ++b) {
    const StatPart p = part[b];
    a.osum += p.osum;
    a.omax = p.omax > a.omax ? p.omax : a.omax;
    a.on += p.on;
    a.gsum += p.gsum;
    a.gn += p.gn;
  }
  sh[k] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (k < h) {
      StatPart& x = sh[k];
      const StatPart y = sh[k + h];
      x.osum += y.osum;
      x.omax = y.omax > x.omax ? y.omax : x.omax;
      x.on += y.on;
      x.gsum += y.gsum;
      x.gn += y.gn;
    }
    __syncthreads();
  }
  if (k == 0) {
    const double time = st->time + dt_step;
    st->time = time;
    st->osum = st->osum + sh[0].osum;
    st->on += sh[0].on;
    st->gsum += sh[0].gsum;
    st->gn += sh[0].gn;
    if (sh[0].omax > st->omax) {
      st->omax = sh[0].omax;
      st->tmax = time;
    }
  }
}


__global__ void __launch_bounds__(kBlock) k_advance(State s, const double4* Fa, const double4* Ta,
                                                    int n, double dtv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 vm = s.vel_m[i], wb = s.wbf_I[i], qq = s.quat[i];
  const double4 f4 = Fa[i], t4 = Ta[i];
  const v3 F = mk(f4.x, f4.y, f4.z), T = mk(t4.x, t4.y, t4.z);
  const quat q = qmk(mk(qq.x, qq.y, qq.z), qq.w);
  const quat pc = qconj(q);
  const v3 tq = qmulv(pc, qleftvec(q, T));
  const double mcf = vm.w * 1.;
  const v3 V = add(mk(vm.x, vm.y, vm.z), scl(dvd(F, mcf), dtv));
  const v3 WBF = ssprk3_sphere(tq, mk(wb.x, wb.y, wb.z), wb.w, dtv);
  const v3 W = qmulv(q, qleftvec(pc, WBF));
  const double req = s.om_req[i].w;
  s.vel_m[i] = make_double4(V.x, V.y, V.z, vm.w);
  s.wbf_I[i] = make_double4(WBF.x, WBF.y, WBF.z, wb.w);
  s.om_req[i] = make_double4(W.x, W.y, W.z, req);
}

__global__ void __launch_bounds__(kBlock) k_unpermute(State s, const double4* Fa,
                                                      const double4* Ta, int n, double* x,
                                                      double* v, double* w, double* q,
                                                      double* f, double* t) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t id = (size_t)s.aux[i].id;
  if (x) { double4 a = s.pos_r[i]; x[3 * id] = a.x; x[3 * id + 1] = a.y; x[3 * id + 2] = a.z; }
  if (v) { double4 a = s.vel_m[i]; v[3 * id] = a.x; v[3 * id + 1] = a.y; v[3 * id + 2] = a.z; }
  if (w) { double4 a = s.om_req[i]; w[3 * id] = a.x; w[3 * id + 1] = a.y; w[3 * id + 2] = a.z; }
  if (q) {
    double4 a = s.quat[i];
    q[4 * id] = a.x; q[4 * id + 1] = a.y; q[4 * id + 2] = a.z; q[4 * id + 3] = a.w;
  }
  if (f) { double4 a = Fa[i]; f[3 * id] = a.x; f[3 * id + 1] = a.y; f[3 * id + 2] = a.z; }
  if (t) { double4 a = Ta[i]; t[3 * id] = a.x; t[3 * id + 1] = a.y; t[3 * id + 2] = a.z; }
}

// Dynamic state given in id order (NULL: kept) written into the storage order
// of the copy s; the body-frame angular velocity follows from the new w and q
// as the reference's reload does (ParticleKinematics::readParticleKinematics2014,
// Component/src/ParticleKinematics.cpp:378-388).
__global__ void __launch_bounds__(kBlock) k_set_state(State s, double4* Fa, double4* Ta, int n,
                                                      const double* x, const double* v,
                                                      const double* w, const double* q,
                                                      const double* f, const double* t) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t id = (size_t)s.aux[i].id;
  if (x) {
    const double4 a = s.pos_r[i];
    s.pos_r[i] = make_double4(x[3 * id], x[3 * id + 1], x[3 * id + 2], a.w);
  }
  if (v) {
    const double4 a = s.vel_m[i];
    s.vel_m[i] = make_double4(v[3 * id], v[3 * id + 1], v[3 * id + 2], a.w);
  }
  if (q) s.quat[i] = make_double4(q[4 * id], q[4 * id + 1], q[4 * id + 2], q[4 * id + 3]);
  if (w) {
    const double4 a = s.om_req[i];
    s.om_req[i] = make_double4(w[3 * id], w[3 * id + 1], w[3 * id + 2], a.w);
  }
  if (w || q) {
    const double4 qq = s.quat[i], wr = s.om_req[i], wb = s.wbf_I[i];
    const quat qi = qmk(mk(qq.x, qq.y, qq.z), qq.w);
    const v3 wbf = qmulv(qconj(qi), qleftvec(qi, mk(wr.x, wr.y, wr.z)));
    s.wbf_I[i] = make_double4(wbf.x, wbf.y, wbf.z, wb.w);
  }
  if (f) Fa[i] = make_double4(f[3 * id], f[3 * id + 1], f[3 * id + 2], 0.);
  if (t) Ta[i] = make_double4(t[3 * id], t[3 * id + 1], t[3 * id + 2], 0.);
}

// ---------------------------------------------------------------- launchers
static inline int nblocks(int n, int bs) { return (n + bs - 1) / bs; }

void launch_keys(const State& s, int32_t n, const Grid& g, const int32_t* row_rank, int32_t* key,
                 unsigned long long* err, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_keys, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, st, s, n, g, row_rank, key,
                     err);
}

void launch_binning(Bins& b, const PosSrc& src, int32_t n, const Grid& g, const int32_t* row_rank,
                    unsigned long long* err, hipStream_t st, SkinGate gate, SkinWords* words,
                    int32_t* in_count) {
  // with skin lists the grids are capped and the kernels stride; one block at
  // least: the count pass also updates the skin words
  const bool gated = gate.w != nullptr;
  auto capped = [gated](int blocks) { return gated ? std::min(blocks, kGateGrid) : blocks; };
  const int ncell = g.ncell;
  hipLaunchKernelGGL(k_count, dim3(capped(nblocks(n > 0 ? n : 1, kBlock))), dim3(kBlock), 0, st,
                     b.key, b.slot, b.cnt, n, gate, words, in_count, src, g, row_rank, err);
  const int nb = nblocks(ncell, kScanCells);
  hipLaunchKernelGGL(k_scan, dim3(nb), dim3(kScanBlock), 0, st, b.cnt, b.local, b.bsum, b.ticket,
                     ncell, nb, gate);
  if (n > 0)
    hipLaunchKernelGGL(k_scatter, dim3(capped(nblocks(n, kBlock))), dim3(kBlock), 0, st, b.key,
                       b.slot, b.local, b.bsum, b.perm, n, gate);
  hipLaunchKernelGGL(k_cellsort, dim3(capped(nblocks(ncell, kBlock))), dim3(kBlock), 0, st,
                     b.local, b.bsum, b.perm, b.cell_start, src, b.pos, b.ids, ncell, n,
                     gate);
}

// The scan and the per-cell sort of a gated binning, for launch sequences
// that run the count and scatter passes inside kernels of their own.
// Largest list length of any particle (its outgoing or incoming entries at
// the last list build): a fixed-size tree per workgroup, then one atomic.
__global__ void __launch_bounds__(256) k_list_peak(const int32_t* out, const int32_t* in, int32_t n,
                                                   int32_t* peak) {
  __shared__ int32_t red[256];
  int32_t m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = max(m, max(out[i], in[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(peak, red[0]);
}

void launch_list_peak(const int32_t* out, const int32_t* in, int32_t n, int32_t* peak,
                      hipStream_t st) {
  if (n <= 0) return;
  const int blocks = std::min((n + 255) / 256, 1024);
  hipLaunchKernelGGL(k_list_peak, dim3(blocks), dim3(256), 0, st, out, in, n, peak);
}

void launch_scan(const Bins& b, int32_t ncell, SkinGate gate, hipStream_t st) {
  const int nb = nblocks(ncell, kScanCells);
  hipLaunchKernelGGL(k_scan, dim3(nb), dim3(kScanBlock), 0, st, b.cnt, b.local, b.bsum, b.ticket,
                     ncell, nb, gate);
}

void launch_cellsort(const Bins& b, const PosSrc& src, int32_t ncell, int32_t n, SkinGate gate,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_cellsort, dim3(std::min(nblocks(ncell, kBlock), kGateGrid)), dim3(kBlock),
                     0, st, b.local, b.bsum, b.perm, b.cell_start, src, b.pos, b.ids, ncell, n,
                     gate);
}

void launch_wall_loads(const ForceArgs& a, double4* part, double4* out, hipStream_t st) {
  if (a.nwalls <= 0) return;
  const int nb = a.n > 0 ? nblocks(a.n, kBlock) : 0;
  if (nb > 0)
    hipLaunchKernelGGL(k_wall_load_partial, dim3(nb), dim3(kBlock), 0, st, a.wl_count, a.wl_wall,
                       (const double4*)a.wl_f, (const double4*)a.wl_t, a.n, a.lstride, a.nwalls,
                       part);
  hipLaunchKernelGGL(k_wall_load_final, dim3(a.nwalls), dim3(256), 0, st, (const double4*)part, nb,
                     a.nwalls, out);
}

void launch_detect(const DetectArgs& a, hipStream_t st) {
  if (a.n <= 0) return;
  hipLaunchKernelGGL(k_detect, dim3(a.nblk8), dim3(kBlock), 0, st, a);
}

void launch_detect_screen(const DetectArgs& d, const ForceArgs& f, hipStream_t st) {
  if (d.n <= 0) return;
  hipLaunchKernelGGL(k_detect_screen, dim3(d.nblk8), dim3(kBlock), 0, st, d, f);
}

template <int M, bool R, bool MEM>
static void force_launch(const ForceArgs& a, hipStream_t st, hipEvent_t start, hipEvent_t stop) {
  const dim3 grid(a.nblk8), block(kBlock);
  if (start)
    hipExtLaunchKernelGGL((k_force<M, R, MEM>), grid, block, 0, st, start, stop, 0, a);
  else
    hipLaunchKernelGGL((k_force<M, R, MEM>), grid, block, 0, st, a);
}

template <bool R, bool MEM>
static void force_launch_mode(int mode, const ForceArgs& a, hipStream_t st, hipEvent_t start,
                              hipEvent_t stop) {
  if (mode == MODE_FORCE) force_launch<MODE_FORCE, R, MEM>(a, st, start, stop);
  else if (mode == MODE_KICK) force_launch<MODE_KICK, R, MEM>(a, st, start, stop);
  else force_launch<MODE_MOVE, R, MEM>(a, st, start, stop);
}

void launch_force(int mode, bool record, const ForceArgs& a, hipStream_t st, hipEvent_t start,
                  hipEvent_t stop, bool memory) {
  if (a.n <= 0) return;
  if (memory) {
    if (record) force_launch_mode<true, true>(mode, a, st, start, stop);
    else force_launch_mode<false, true>(mode, a, st, start, stop);
  } else {
    if (record) force_launch_mode<true, false>(mode, a, st, start, stop);
    else force_launch_mode<false, false>(mode, a, st, start, stop);
  }
}

void launch_move(const State& s, const double4* F, const double4* T, int32_t n, const Grid& g,
                 const int32_t* row_rank, double dt_vel, double dt_disp, int32_t* key,
                 unsigned long long* err, hipStream_t st, SkinGate gate, float* path) {
  if (n > 0)
    hipLaunchKernelGGL(k_move, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, st, s, F, T, n, g,
                       row_rank, dt_vel, dt_disp, key, err, gate, path);
}

void launch_stats_final(const StatPart* part, int32_t nparts, ContactStats* st, double dt_step,
                        hipStream_t stream) {
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, stream, part, nparts, st, dt_step);
}


void launch_advance(const State& s, const double4* F, const double4* T, int32_t n,
                    double dt_vel, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_advance, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, st, s, F, T, n, dt_vel);
}

void launch_unpermute(const State& s, const double4* F, const double4* T, int32_t n, double* x,
                      double* v, double* w, double* q, double* f, double* t, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_unpermute, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, st, s, F, T, n, x,
                     v, w, q, f, t);
}

void launch_set_state(const State& s, double4* F, double4* T, int32_t n, const double* x,
                      const double* v, const double* w, const double* q, const double* f,
                      const double* t, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_set_state, dim3(nblocks(n, kBlock)), dim3(kBlock), 0, st, s, F, T, n, x,
                     v, w, q, f, t);
}

}  // namespace gdem
