// gdem_batch.hip -- many independent small sphere systems in one persistent
// launch (calibration sweeps), and its C ABI (include/gdem.h, gdem_batch_*).
//
// Each workgroup takes systems from a queue (largest first) and runs every
// phase of the step for that system with workgroup barriers between phases:
// counting sort of its particles by cell key, the half-stencil detection sweep
// and the force + integrator tail.  The per-particle bodies are the ones of the
// single-system kernels (gdem_sphere_step.h) and the binning orders each cell
// by previous slot as the single-system binning does, so a system advanced
// here is bit-identical to the same system advanced alone by gdem_step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_host.h"
#include "gdem_internal.h"
#include "gdem_sphere_step.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

// Per-system record: grid, float pre-filter constants, offsets of the
// system's slices in the concatenated arrays.
struct BatchSys {
  Grid grid;
  float sizef[3];
  float margin;
  double gravity[3];
  int32_t n, off;       // particles, first particle
  int32_t cell_off;     // first cell (counts)
  int32_t cs_off;       // first cell_start entry (ncell + 1 per system)
  int32_t row_off;      // first row rank
  int32_t wall_off, nwalls;
  int32_t model_off;    // 2 x 2 model table: particle material 0, wall material 1
  int32_t substeps;     // steps of dt / substeps per step of dt
};

struct BatchArgs {
  const BatchSys* sys;
  const int32_t* order;  // queue: systems by decreasing particle count
  int32_t nsys;
  int32_t* next;         // queue head, zero at launch
  State st[2];
  int32_t cur;           // state copy holding the current state
  int32_t prime;         // 1: compute keys, bin and evaluate a0 first
  int32_t nsteps;
  double dt;
  double4* F;
  double4* T;
  int32_t* key;
  int32_t* slot;
  int32_t* perm;
  float4* pos;
  int32_t* cnt;
  int32_t* cell_start;
  const int32_t* row_rank;
  int2* out_list;
  int2* in_list;
  int32_t* out_count;
  int32_t* in_count;
  const Model* models;
  const Wall* walls;
  unsigned long long* err;       // kErrWords per system
  unsigned long long* contacts;  // per system, summed over force evaluations
};

__device__ __forceinline__ State state_at(const State& s, int off) {
  State r;
  r.pos_r = s.pos_r + off;
  r.vel_m = s.vel_m + off;
  r.om_req = s.om_req + off;
  r.wbf_I = s.wbf_I + off;
  r.quat = s.quat + off;
  r.aux = s.aux + off;
  return r;
}

template <int BS>
__device__ __forceinline__ int block_scan_excl(int v, int* smem, int& total) {
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

template <int BS>
struct BatchRunner {
  const BatchArgs& a;
  const BatchSys& sy;
  int s;
  int* smem;
  DetectArgs* sd;  // workgroup-shared launch records of the current phase
  ForceArgs* sf;
  State* sst;      // workgroup-shared views of the system's two state copies
  unsigned long long ncont;

  __device__ __forceinline__ unsigned long long* err() const { return a.err + (size_t)s * kErrWords; }

  // True once any error of this system is raised: the reference stops at the
  // exception, so the system is not advanced further (its state is then not
  // meaningful).  Read with agent-scope atomic loads (the words are written by
  // atomics) and broadcast through LDS, so the branch is uniform.
  __device__ __forceinline__ bool failed(int* flag) {
    if (threadIdx.x == 0) {
      unsigned long long all = kNoErr;
      for (int k = 0; k < kErrWords; ++k)
        all &= __hip_atomic_load(&err()[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *flag = all != kNoErr;
    }
    __syncthreads();
    const bool f = *flag != 0;
    __syncthreads();
    return f;
  }

  // Counting sort by cell key of the current copy (as launch_binning).
  __device__ __forceinline__ void bin(const State& cur) {
    const int n = sy.n, nc = sy.grid.ncell;
    int32_t* key = a.key + sy.off;
    int32_t* slot = a.slot + sy.off;
    int32_t* perm = a.perm + sy.off;
    int32_t* cnt = a.cnt + sy.cell_off;
    int32_t* cs = a.cell_start + sy.cs_off;
    for (int i = threadIdx.x; i < n; i += BS) slot[i] = atomicAdd(&cnt[key[i]], 1);
    __syncthreads();
    int carry = 0;
    for (int base = 0; base < nc; base += BS) {
      const int c = base + threadIdx.x;
      const int v = c < nc ? cnt[c] : 0;
      int total;
      const int ex = block_scan_excl<BS>(v, smem, total);
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
      const double4 p = cur.pos_r[perm[u]];
      pos[u] = make_float4((float)p.x, (float)p.y, (float)p.z, (float)p.w);
    }
    __syncthreads();
  }

  __device__ __forceinline__ void detect(const State& cur) {
    if (threadIdx.x == 0) {
      DetectArgs& d = *sd;
      d = DetectArgs{};  // fields not set below are null / zero
      d.pos = a.pos + sy.off;
      d.sizef[0] = sy.sizef[0]; d.sizef[1] = sy.sizef[1]; d.sizef[2] = sy.sizef[2];
      d.margin = sy.margin;
      d.in = cur;
      d.perm = a.perm + sy.off;
      d.cell_start = a.cell_start + sy.cs_off;
      d.row_rank = a.row_rank + sy.row_off;
      d.grid = sy.grid;
      d.n = sy.n;
      d.lstride = sy.n;
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
    for (int t = threadIdx.x; t < sy.n; t += BS) detect_particle(*sd, t);
    __syncthreads();
  }

  template <int MODE>
  __device__ __forceinline__ void force(const State& cur, const State& nxt, double dtv, double dtd, bool write_ft) {
    if (threadIdx.x == 0) {
      ForceArgs& f = *sf;
      f = ForceArgs{};  // every field not set below is null / zero (no wall loads, no history)
      f.in = cur;
      f.out = nxt;
      f.perm = a.perm + sy.off;
      f.row_rank = a.row_rank + sy.row_off;
      f.grid = sy.grid;
      f.models = a.models + sy.model_off;
      f.model_stride = 2;
      f.walls = a.walls + sy.wall_off;
      f.nwalls = sy.nwalls;
      f.n = sy.n;
      f.gravity[0] = sy.gravity[0]; f.gravity[1] = sy.gravity[1]; f.gravity[2] = sy.gravity[2];
      f.dt_vel = dtv;
      f.dt_disp = dtd;
      f.lstride = sy.n;
      f.out_list = a.out_list + (size_t)sy.off * kListCap;
      f.out_count = a.out_count + sy.off;
      f.in_list = a.in_list + (size_t)sy.off * kListCap;
      f.in_count = a.in_count + sy.off;
      f.F = write_ft ? a.F + sy.off : nullptr;
      f.T = write_ft ? a.T + sy.off : nullptr;
      f.key_out = a.key + sy.off;
      f.err = err();
      f.contacts = nullptr;
      f.rec = nullptr;
      f.rec_count = nullptr;
      f.rec_cap = 0;
      f.nblk8 = 0;
      f.gate = SkinGate{};
      f.words = nullptr;
      f.cmask = nullptr;
    }
    __syncthreads();
    double drift;
    for (int t = threadIdx.x; t < sy.n; t += BS)
      ncont += (unsigned long long)force_particle<MODE, false>(*sf, t, drift);
    __syncthreads();
  }

  __device__ __forceinline__ void run() {
    int cur = a.cur;
    if (threadIdx.x == 0) {
      sst[0] = state_at(a.st[0], sy.off);
      sst[1] = state_at(a.st[1], sy.off);
    }
    __syncthreads();
    // a stiff system advances substeps steps of dt / substeps per step
    const int sub = sy.substeps;
    const double dt = a.dt / (double)sub;
    const int nsteps = a.nsteps * sub;
    ncont = 0;
    if (a.prime) {
      // keys of the initial positions (k_keys), then Link + a0
      const State& c0 = sst[cur];
      for (int i = threadIdx.x; i < sy.n; i += BS) {
        const double4 p = c0.pos_r[i];
        int ijk[3];
        if (!cell_of(sy.grid, mk(p.x, p.y, p.z), ijk)) {
          atomicMin(&err()[kErrSim], (unsigned long long)(uint32_t)c0.aux[i].id);
          a.key[sy.off + i] = 0;
        } else {
          a.key[sy.off + i] = row_base(sy.grid, a.row_rank + sy.row_off, ijk[1], ijk[2]) + ijk[0];
        }
      }
      __syncthreads();
      bin(c0);
      detect(c0);
      force<MODE_FORCE>(c0, sst[cur ^ 1], 0., 0., true);
      cur ^= 1;
    }
    if (nsteps > 0) {
      // step 1's kick + drift, then each force evaluation closes its step and
      // opens the next one (gdem_step)
      for (int i = threadIdx.x; i < sy.n; i += BS)
        move_particle(sst[cur], a.F + sy.off, a.T + sy.off, i, sy.grid, a.row_rank + sy.row_off,
                      0.5 * dt, dt, a.key + sy.off, err());
      __syncthreads();
      for (int k = 0; k < nsteps; ++k) {
        if (failed(smem)) break;
        bin(sst[cur]);
        detect(sst[cur]);
        if (k == nsteps - 1) force<MODE_KICK>(sst[cur], sst[cur ^ 1], 0.5 * dt, dt, true);
        else force<MODE_MOVE>(sst[cur], sst[cur ^ 1], 0.5 * dt, dt, false);
        cur ^= 1;
      }
    }
    // the host takes the state from the copy that prime + nsteps evaluations
    // reach; a system in an even number of substeps ends in the other copy
    // after an odd number of steps, so its state moves there (same slot
    // order: keys, F and T stay valid)
    const int expect = a.cur ^ ((a.prime + a.nsteps) & 1);
    if (cur != expect) {
      const State& src = sst[cur];
      const State& dst = sst[expect];
      for (int i = threadIdx.x; i < sy.n; i += BS) {
        dst.pos_r[i] = src.pos_r[i];
        dst.vel_m[i] = src.vel_m[i];
        dst.om_req[i] = src.om_req[i];
        dst.wbf_I[i] = src.wbf_I[i];
        dst.quat[i] = src.quat[i];
        dst.aux[i] = src.aux[i];
      }
      __syncthreads();
    }
    // per-system contact count: wave sums, one atomic per workgroup
    unsigned long long c = ncont;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    __shared__ unsigned long long sh[BS / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long tot = 0;
      for (int w = 0; w < BS / 64; ++w) tot += sh[w];
      if (tot) atomicAdd(&a.contacts[s], tot);
    }
    __syncthreads();
  }
};

// Persistent: workgroups take systems from the queue until it is empty; every
// wave leaves the loop together when the head passes the last system.
template <int BS>
__global__ void __launch_bounds__(BS) k_batch(BatchArgs a) {
  __shared__ int s_item;
  __shared__ int smem[BS / 64];
  __shared__ DetectArgs sd;
  __shared__ ForceArgs sf;
  __shared__ BatchSys ssy;
  __shared__ State sst[2];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(a.next, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= a.nsys) break;
    const int s = a.order[item];
    if (threadIdx.x == 0) ssy = a.sys[s];
    __syncthreads();
    BatchRunner<BS> r{a, ssy, s, smem, &sd, &sf, sst, 0ull};
    r.run();
  }
}

__global__ void k_zero_int(int32_t* p) { *p = 0; }

}  // namespace gdem

using namespace gdem;

struct gdem_batch {
  int dev = 0;
  hipStream_t stream = nullptr;
  bool committed = false, primed = false;
  int32_t block = 256, workgroups = 0;
  int cur = 0;
  // host staging until commit
  std::vector<BatchSys> sys;
  std::vector<double4> pr, vm, wr, wb, qq;
  std::vector<Aux> ax;
  std::vector<int32_t> rows;
  std::vector<Model> models;
  std::vector<Wall> walls;
  int64_t ntot = 0, ncell_tot = 0, ncs_tot = 0, nmax = 0;
  // device
  BatchSys* d_sys = nullptr;
  int32_t* d_order = nullptr;
  int32_t* d_next = nullptr;
  State st[2]{};
  double4 *d_F = nullptr, *d_T = nullptr;
  int32_t *d_key = nullptr, *d_slot = nullptr, *d_perm = nullptr, *d_cnt = nullptr,
          *d_cs = nullptr, *d_rows = nullptr, *d_out_count = nullptr, *d_in_count = nullptr;
  float4* d_pos = nullptr;
  int2 *d_out_list = nullptr, *d_in_list = nullptr;
  Model* d_models = nullptr;
  Wall* d_walls = nullptr;
  unsigned long long *d_err = nullptr, *d_contacts = nullptr;
  double* d_stage = nullptr;
  // timing
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double kernel_ms = 0.;
  int64_t launches = 0, steps = 0;
  bool pending_event = false;
  // last error
  int32_t err_sys = -1, err_code = 0, err_id0 = -1, err_id1 = -1;
};

#define BHIP(expr)                                                                  \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "gdem_batch: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return GDEM_ERR_HIP;                                                          \
    }                                                                               \
  } while (0)

extern "C" int gdem_batch_create(int32_t device, gdem_batch** out) {
  if (!out) return GDEM_ERR_ARG;
  *out = nullptr;
  gdem_batch* b = new gdem_batch();
  b->dev = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return GDEM_ERR_HIP;
  }
  *out = b;
  return GDEM_OK;
}

static void free_batch_device(gdem_batch* b) {
  for (auto& s : b->st) {
    (void)hipFree(s.pos_r); (void)hipFree(s.vel_m); (void)hipFree(s.om_req); (void)hipFree(s.wbf_I);
    (void)hipFree(s.quat); (void)hipFree(s.aux);
    s = State{};
  }
  void* ptrs[] = {b->d_sys, b->d_order, b->d_next, b->d_F, b->d_T, b->d_key, b->d_slot, b->d_perm,
                  b->d_cnt, b->d_cs, b->d_rows, b->d_out_count, b->d_in_count, b->d_pos,
                  b->d_out_list, b->d_in_list, b->d_models, b->d_walls, b->d_err, b->d_contacts,
                  b->d_stage};
  for (void* p : ptrs) (void)hipFree(p);
}

extern "C" int gdem_batch_destroy(gdem_batch* b) {
  if (!b) return GDEM_OK;
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch_device(b);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GDEM_OK;
}

extern "C" int gdem_batch_add_system(gdem_batch* b, const gdem_batch_system* in, int32_t* sys_id) {
  if (!b || !in || b->committed) return GDEM_ERR_ARG;
  const int64_t n = in->n;
  if (n < 0 || b->ntot + n > 0x7fffffff) return GDEM_ERR_ARG;
  if (n > 0 && (!in->x || !in->v || !in->w || !in->q || !in->radius || !in->density || !in->crust))
    return GDEM_ERR_ARG;
  if (in->nwalls < 0 || in->nwalls > kMaxWalls) return GDEM_ERR_ARG;
  if (in->nwalls > 0 && (!in->wall_basis || !in->wall_centre || !in->wall_half)) return GDEM_ERR_ARG;
  if (in->kind != GDEM_MODEL_HOOKE && in->kind != GDEM_MODEL_HERTZ) return GDEM_ERR_ARG;
  if (!model_params_ok(in->kind, in->pp) || !model_params_ok(in->kind, in->pw)) return GDEM_ERR_ARG;
  if (in->substeps < 0 || in->substeps > 4096) return GDEM_ERR_ARG;
  BatchSys sy{};
  sy.substeps = std::max<int32_t>(in->substeps, 1);
  if (!make_grid(in->cfg, sy.grid)) return GDEM_ERR_ARG;
  sy.margin = prefilter_margin(sy.grid, sy.sizef);
  for (int d = 0; d < 3; ++d) sy.gravity[d] = in->cfg.gravity[d];
  sy.n = (int32_t)n;
  sy.off = (int32_t)b->ntot;
  sy.cell_off = (int32_t)b->ncell_tot;
  sy.cs_off = (int32_t)b->ncs_tot;
  sy.row_off = (int32_t)b->rows.size();
  sy.wall_off = (int32_t)b->walls.size();
  sy.nwalls = in->nwalls;
  sy.model_off = (int32_t)b->models.size();
  for (int64_t i = 0; i < n; ++i) {
    if (!(in->radius[i] > 0.) || !(in->density[i] > 0.) || !(in->crust[i] > 0.)) return GDEM_ERR_ARG;
  }
  const std::vector<int32_t> rank = row_morton_rank(sy.grid.nc);
  b->rows.insert(b->rows.end(), rank.begin(), rank.end());
  const Model pp = make_model(in->kind, in->pp), pw = make_model(in->kind, in->pw);
  b->models.push_back(pp);  // (0, 0)
  b->models.push_back(pw);  // (0, 1)
  b->models.push_back(pw);  // (1, 0)
  b->models.push_back(pp);  // (1, 1): wall-wall, never evaluated
  for (int k = 0; k < in->nwalls; ++k)
    b->walls.push_back(make_wall(in->wall_basis + 9 * k, in->wall_centre + 3 * k,
                                 in->wall_half + 3 * k, in->wall_crust, 1));
  for (int64_t i = 0; i < n; ++i) {
    double4 pr, vm, wr, wb, qq;
    Aux ax;
    sphere_particle(in->x + 3 * i, in->v + 3 * i, in->w + 3 * i, in->q + 4 * i, in->radius[i],
                    in->density[i], in->crust[i], 0, (int32_t)i, pr, vm, wr, wb, qq, ax);
    b->pr.push_back(pr); b->vm.push_back(vm); b->wr.push_back(wr); b->wb.push_back(wb);
    b->qq.push_back(qq); b->ax.push_back(ax);
  }
  b->ntot += n;
  b->nmax = std::max(b->nmax, n);
  b->ncell_tot += sy.grid.ncell;
  b->ncs_tot += sy.grid.ncell + 1;
  if (sys_id) *sys_id = (int32_t)b->sys.size();
  b->sys.push_back(sy);
  return GDEM_OK;
}

template <class T>
static hipError_t upload(T** d, const std::vector<T>& h) {
  hipError_t e = hipMalloc(d, sizeof(T) * std::max<size_t>(h.size(), 1));
  if (e != hipSuccess || h.empty()) return e;
  return hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

extern "C" int gdem_batch_commit(gdem_batch* b) {
  if (!b || b->committed) return GDEM_ERR_ARG;
  BHIP(hipSetDevice(b->dev));
  const size_t n = (size_t)std::max<int64_t>(b->ntot, 1);
  const size_t ns = std::max<size_t>(b->sys.size(), 1);
  std::vector<int32_t> order(b->sys.size());
  std::iota(order.begin(), order.end(), 0);
  // largest work first: particles times substeps
  auto work = [&](int32_t k) { return (int64_t)b->sys[k].n * b->sys[k].substeps; };
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return work(x) > work(y); });
  BHIP(upload(&b->d_sys, b->sys));
  BHIP(upload(&b->d_order, order));
  BHIP(upload(&b->d_rows, b->rows));
  BHIP(upload(&b->d_models, b->models));
  BHIP(upload(&b->d_walls, b->walls));
  BHIP(upload(&b->st[0].pos_r, b->pr));
  BHIP(upload(&b->st[0].vel_m, b->vm));
  BHIP(upload(&b->st[0].om_req, b->wr));
  BHIP(upload(&b->st[0].wbf_I, b->wb));
  BHIP(upload(&b->st[0].quat, b->qq));
  BHIP(upload(&b->st[0].aux, b->ax));
  State& s1 = b->st[1];
  BHIP(hipMalloc(&s1.pos_r, sizeof(double4) * n));
  BHIP(hipMalloc(&s1.vel_m, sizeof(double4) * n));
  BHIP(hipMalloc(&s1.om_req, sizeof(double4) * n));
  BHIP(hipMalloc(&s1.wbf_I, sizeof(double4) * n));
  BHIP(hipMalloc(&s1.quat, sizeof(double4) * n));
  BHIP(hipMalloc(&s1.aux, sizeof(Aux) * n));
  BHIP(hipMalloc(&b->d_next, sizeof(int32_t)));
  BHIP(hipMalloc(&b->d_F, sizeof(double4) * n));
  BHIP(hipMalloc(&b->d_T, sizeof(double4) * n));
  BHIP(hipMemset(b->d_F, 0, sizeof(double4) * n));
  BHIP(hipMemset(b->d_T, 0, sizeof(double4) * n));
  BHIP(hipMalloc(&b->d_key, sizeof(int32_t) * n));
  BHIP(hipMalloc(&b->d_slot, sizeof(int32_t) * n));
  BHIP(hipMalloc(&b->d_perm, sizeof(int32_t) * n));
  BHIP(hipMalloc(&b->d_pos, sizeof(float4) * n));
  BHIP(hipMalloc(&b->d_cnt, sizeof(int32_t) * std::max<int64_t>(b->ncell_tot, 1)));
  BHIP(hipMemset(b->d_cnt, 0, sizeof(int32_t) * std::max<int64_t>(b->ncell_tot, 1)));
  BHIP(hipMalloc(&b->d_cs, sizeof(int32_t) * std::max<int64_t>(b->ncs_tot, 1)));
  BHIP(hipMalloc(&b->d_out_list, sizeof(int2) * n * kListCap));
  BHIP(hipMalloc(&b->d_in_list, sizeof(int2) * n * kListCap));
  BHIP(hipMalloc(&b->d_out_count, sizeof(int32_t) * n));
  BHIP(hipMalloc(&b->d_in_count, sizeof(int32_t) * n));
  BHIP(hipMemset(b->d_in_count, 0, sizeof(int32_t) * n));
  BHIP(hipMalloc(&b->d_err, sizeof(unsigned long long) * kErrWords * ns));
  BHIP(hipMemset(b->d_err, 0xff, sizeof(unsigned long long) * kErrWords * ns));
  BHIP(hipMalloc(&b->d_contacts, sizeof(unsigned long long) * ns));
  BHIP(hipMemset(b->d_contacts, 0, sizeof(unsigned long long) * ns));
  BHIP(hipDeviceSynchronize());
  b->pr.clear(); b->vm.clear(); b->wr.clear(); b->wb.clear(); b->qq.clear(); b->ax.clear();
  b->committed = true;
  b->cur = 0;
  return GDEM_OK;
}

extern "C" int gdem_batch_set_launch(gdem_batch* b, int32_t block_size, int32_t workgroups) {
  if (!b) return GDEM_ERR_ARG;
  if (block_size != 64 && block_size != 128 && block_size != 256 && block_size != 512)
    return GDEM_ERR_ARG;
  if (workgroups < 0) return GDEM_ERR_ARG;
  b->block = block_size;
  b->workgroups = workgroups;
  return GDEM_OK;
}

static int batch_launch(gdem_batch* b, int prime, int32_t nsteps, double dt) {
  const int nsys = (int)b->sys.size();
  if (nsys == 0) return GDEM_OK;
  int wg = b->workgroups;
  if (wg <= 0) {
    // enough workgroups to keep every CU busy while the queue lasts
    hipDeviceProp_t prop;
    BHIP(hipGetDeviceProperties(&prop, b->dev));
    wg = prop.multiProcessorCount * (1024 / b->block) * 2;
  }
  wg = std::min(wg, nsys);
  BatchArgs a{};
  a.sys = b->d_sys;
  a.order = b->d_order;
  a.nsys = nsys;
  a.next = b->d_next;
  a.st[0] = b->st[0];
  a.st[1] = b->st[1];
  a.cur = b->cur;
  a.prime = prime;
  a.nsteps = nsteps;
  a.dt = dt;
  a.F = b->d_F;
  a.T = b->d_T;
  a.key = b->d_key;
  a.slot = b->d_slot;
  a.perm = b->d_perm;
  a.pos = b->d_pos;
  a.cnt = b->d_cnt;
  a.cell_start = b->d_cs;
  a.row_rank = b->d_rows;
  a.out_list = b->d_out_list;
  a.in_list = b->d_in_list;
  a.out_count = b->d_out_count;
  a.in_count = b->d_in_count;
  a.models = b->d_models;
  a.walls = b->d_walls;
  a.err = b->d_err;
  a.contacts = b->d_contacts;
  (void)hipGetLastError();  // an earlier failed call must not be reported as this launch's
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, b->stream, b->d_next);
  if (b->timing) {
    if (!b->ev0) {
      BHIP(hipEventCreate(&b->ev0));
      BHIP(hipEventCreate(&b->ev1));
    }
    BHIP(hipEventRecord(b->ev0, b->stream));
  }
  switch (b->block) {
    case 64: hipLaunchKernelGGL(k_batch<64>, dim3(wg), dim3(64), 0, b->stream, a); break;
    case 128: hipLaunchKernelGGL(k_batch<128>, dim3(wg), dim3(128), 0, b->stream, a); break;
    case 512: hipLaunchKernelGGL(k_batch<512>, dim3(wg), dim3(512), 0, b->stream, a); break;
    default: hipLaunchKernelGGL(k_batch<256>, dim3(wg), dim3(256), 0, b->stream, a); break;
  }
  BHIP(hipGetLastError());
  if (b->timing) {
    BHIP(hipEventRecord(b->ev1, b->stream));
    BHIP(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BHIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->kernel_ms += ms;
    b->launches++;
  }
  const int evals = (prime ? 1 : 0) + nsteps;
  b->cur = (b->cur + evals) & 1;
  b->steps += nsteps;
  return GDEM_OK;
}

extern "C" int gdem_batch_prime(gdem_batch* b) {
  if (!b || !b->committed) return GDEM_ERR_ARG;
  BHIP(hipSetDevice(b->dev));
  int rc = batch_launch(b, 1, 0, 0.);
  if (rc) return rc;
  b->primed = true;
  return GDEM_OK;
}

extern "C" int gdem_batch_step(gdem_batch* b, int32_t nsteps, double dt) {
  if (!b || !b->committed || nsteps < 0 || !(dt > 0.)) return GDEM_ERR_ARG;
  if (nsteps == 0) return GDEM_OK;
  BHIP(hipSetDevice(b->dev));
  const int rc = batch_launch(b, b->primed ? 0 : 1, nsteps, dt);
  b->primed = true;
  return rc;
}

static int batch_errors(gdem_batch* b) {
  const size_t ns = b->sys.size();
  if (ns == 0) return GDEM_OK;
  std::vector<unsigned long long> h(ns * kErrWords);
  BHIP(hipMemcpy(h.data(), b->d_err, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
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
      b->err_sys = (int32_t)s;
      b->err_code = code;
      b->err_id0 = id0;
      b->err_id1 = id1;
      BHIP(hipMemset(b->d_err, 0xff, sizeof(unsigned long long) * h.size()));
      return code;
    }
  }
  return GDEM_OK;
}

extern "C" int gdem_batch_synchronize(gdem_batch* b) {
  if (!b) return GDEM_ERR_ARG;
  BHIP(hipSetDevice(b->dev));
  BHIP(hipStreamSynchronize(b->stream));
  return batch_errors(b);
}

__global__ void k_batch_unpermute(State s, const double4* Fa, const double4* Ta, int off, int n,
                                  double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = off + i;
  const size_t id = (size_t)s.aux[g].id;
  double* x = out;
  double* v = x + 3 * (size_t)n;
  double* w = v + 3 * (size_t)n;
  double* q = w + 3 * (size_t)n;
  double* f = q + 4 * (size_t)n;
  double* t = f + 3 * (size_t)n;
  const double4 p = s.pos_r[g], vm = s.vel_m[g], wr = s.om_req[g], qq = s.quat[g], f4 = Fa[g],
                t4 = Ta[g];
  x[3 * id] = p.x; x[3 * id + 1] = p.y; x[3 * id + 2] = p.z;
  v[3 * id] = vm.x; v[3 * id + 1] = vm.y; v[3 * id + 2] = vm.z;
  w[3 * id] = wr.x; w[3 * id + 1] = wr.y; w[3 * id + 2] = wr.z;
  q[4 * id] = qq.x; q[4 * id + 1] = qq.y; q[4 * id + 2] = qq.z; q[4 * id + 3] = qq.w;
  f[3 * id] = f4.x; f[3 * id + 1] = f4.y; f[3 * id + 2] = f4.z;
  t[3 * id] = t4.x; t[3 * id + 1] = t4.y; t[3 * id + 2] = t4.z;
}

extern "C" int gdem_batch_get_state(gdem_batch* b, int32_t sys, double* x, double* v, double* w,
                                    double* q, double* F, double* T) {
  if (!b || !b->committed || sys < 0 || sys >= (int32_t)b->sys.size()) return GDEM_ERR_ARG;
  BHIP(hipSetDevice(b->dev));
  const BatchSys& sy = b->sys[(size_t)sys];
  const size_t n = (size_t)sy.n;
  if (!b->d_stage) BHIP(hipMalloc(&b->d_stage, sizeof(double) * 19 * (size_t)std::max<int64_t>(b->nmax, 1)));
  if (n > 0) {
    hipLaunchKernelGGL(k_batch_unpermute, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream,
                       b->st[b->cur], b->d_F, b->d_T, sy.off, (int)n, b->d_stage);
    BHIP(hipGetLastError());
    std::vector<double> h(19 * n);
    BHIP(hipMemcpyAsync(h.data(), b->d_stage, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                        b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    const double* src = h.data();
    if (x) memcpy(x, src, sizeof(double) * 3 * n);
    if (v) memcpy(v, src + 3 * n, sizeof(double) * 3 * n);
    if (w) memcpy(w, src + 6 * n, sizeof(double) * 3 * n);
    if (q) memcpy(q, src + 9 * n, sizeof(double) * 4 * n);
    if (F) memcpy(F, src + 13 * n, sizeof(double) * 3 * n);
    if (T) memcpy(T, src + 16 * n, sizeof(double) * 3 * n);
  }
  BHIP(hipStreamSynchronize(b->stream));
  return batch_errors(b);
}

extern "C" int gdem_batch_last_error(gdem_batch* b, int32_t* sys, int32_t* code, int32_t* id0,
                                     int32_t* id1) {
  if (!b) return GDEM_ERR_ARG;
  if (sys) *sys = b->err_sys;
  if (code) *code = b->err_code;
  if (id0) *id0 = b->err_id0;
  if (id1) *id1 = b->err_id1;
  return b->err_code;
}

extern "C" int gdem_batch_set_kernel_timing(gdem_batch* b, int32_t on) {
  if (!b) return GDEM_ERR_ARG;
  b->timing = on != 0;
  return GDEM_OK;
}

extern "C" int gdem_batch_get_counters(gdem_batch* b, int64_t* steps, int64_t* contacts,
                                       double* kernel_ms, int64_t* launches) {
  if (!b) return GDEM_ERR_ARG;
  BHIP(hipSetDevice(b->dev));
  BHIP(hipStreamSynchronize(b->stream));
  if (contacts) {
    *contacts = 0;
    const size_t ns = b->sys.size();
    if (ns && b->committed) {
      std::vector<unsigned long long> h(ns);
      BHIP(hipMemcpy(h.data(), b->d_contacts, sizeof(unsigned long long) * ns, hipMemcpyDeviceToHost));
      for (auto v : h) *contacts += (int64_t)v;
    }
  }
  if (steps) *steps = b->steps;
  if (kernel_ms) *kernel_ms = b->kernel_ms;
  if (launches) *launches = b->launches;
  return GDEM_OK;
}

extern "C" int gdem_batch_info(gdem_batch* b, int32_t* nsys, int64_t* nparticles) {
  if (!b) return GDEM_ERR_ARG;
  if (nsys) *nsys = (int32_t)b->sys.size();
  if (nparticles) *nparticles = b->ntot;
  return GDEM_OK;
}
