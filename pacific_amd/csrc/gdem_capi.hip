// gdem_capi.hip -- the C ABI declared in include/gdem.h.
//
// The host side owns one HIP stream per context and orchestrates the phases of
// Grains::Simulation (src/Grains3D/Grains/Base/src/Grains.cpp:344-364) as
// kernel launches; nothing here computes physics on the CPU except the
// per-particle constants the reference also derives once at construction
// (mass, inertia, equivalent radius: Component/src/Particle.cpp:158-164).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gdem.h"
#include "gdem_host.h"
#include "gdem_internal.h"
#include "gdem_wall_motion.h"

#pragma clang fp contract(off)

using namespace gdem;

struct TimedLaunch {
  hipEvent_t start, stop;
  int kind;
};
constexpr int kTimedKinds = 2;

struct gdem_ctx {
  gdem_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  Grid grid{};
  int32_t* d_row_rank = nullptr;
  int32_t* d_row_jk = nullptr;
  int64_t n = 0;
  State buf[2]{};
  int cur = 0;
  double4* d_F = nullptr;
  double4* d_T = nullptr;
  Bins bins{};
  int2* d_out_list = nullptr;
  int2* d_in_list = nullptr;
  int32_t* d_out_count = nullptr;
  int32_t* d_in_count = nullptr;
  Model h_models[kMaxMaterials * kMaxMaterials]{};
  Model* d_models = nullptr;
  Wall h_walls[kMaxWalls]{};
  // host copies of what the reload files describe: the particles' constants
  // in id order, material and wall names
  std::vector<double> h_radius, h_density, h_crust;
  std::vector<int32_t> h_material;
  std::string mat_name[kMaxMaterials];
  std::string wall_name[kMaxWalls];
  Wall* d_walls = nullptr;
  int nwalls = 0;
  unsigned long long* d_err = nullptr;
  unsigned long long* d_counters = nullptr;  // [0] contacts (convex kernels)
  unsigned long long* d_blk_contacts = nullptr;  // contacts counted per workgroup (spheres)
  int32_t nblk_slots = 0;
  // contact statistics (AppCollision::addToContactsFeatures), when switched on
  bool stats_on = false;
  int32_t gjk_sv = 0;  // convex narrow phase: 0 Johnson GJK, 1 signed-volume GJK
  ContactStats* d_stats = nullptr;
  StatPart* d_stat_part = nullptr;
  bool primed = false, binned = false;
  int32_t rec_cap = 0;
  ContactRec* d_rec = nullptr;
  int32_t* d_rec_count = nullptr;
  bool rec_valid = false;
  // staging for gdem_get_state
  double* d_stage = nullptr;
  size_t stage_len = 0;
  // counters / timing: events bracket kernels of kind 0 (the force kernel, or
  // the integrator kernel of the packed layout) and kind 1 (the pair kernel)
  int64_t force_evals = 0;
  int64_t prime_evals = -1;  // force_evals right after the priming evaluation
  bool timing = false;
  std::vector<TimedLaunch> ev_pool;
  size_t ev_used = 0;
  double kernel_ms[kTimedKinds] = {0., 0.};
  int64_t timed[kTimedKinds] = {0, 0};
  // last error
  int32_t err_code = 0, err_id0 = -1, err_id1 = -1;
  // convex particles: shapes, orientation state, stored contacts, wall shapes
  bool convex = false;
  CShape h_shapes[kMaxShapes]{};
  int nshapes = 0;
  CShape* d_shapes = nullptr;
  CState cbuf[2]{};
  int2* d_cand = nullptr;
  CRec* d_crec = nullptr;
  uint32_t* d_mask = nullptr;
  int32_t* d_work = nullptr;
  int32_t* d_nwork = nullptr;
  CShape h_wall_shapes[kMaxWalls]{};
  double4 h_wall_sc[kMaxWalls]{};
  CShape* d_wall_shapes = nullptr;
  double4* d_wall_sc = nullptr;
  // skin lists (spheres): requested skin (< 0: automatic), effective skin,
  // device words and reference positions
  double skin_req = -1., skin = 0., rmax = 0., crust_min = 0.;
  // skin lists: halve the skin when a list build comes within kListHeadroom
  // entries of the capacity (checked at the end of each gdem_step)
  bool skin_adapt = true;
  int32_t skin_halvings = 0;
  int32_t* d_peak = nullptr;
  SkinWords* d_skin = nullptr;
  float* d_path = nullptr;  // path length of each particle since the last rebuild
  unsigned long long* d_cmask = nullptr;
  // packed layout of the sphere step (gdem_pack.hip): used by gdem_step with
  // skin lists and memoryless models; the canonical copies buf[] are brought
  // up to date before any other call reads or writes the state
  bool packed = false;
  Pack pk[2]{};
  PackConst pkc[2]{};
  uint32_t* d_omask = nullptr;
  uint32_t* d_imask = nullptr;
  PairRec* d_prec = nullptr;
  int32_t* d_queue = nullptr;   // candidates that may touch (t * 32 + entry)
  int32_t* d_qcount = nullptr;  // its length, reset by the integrator
  int32_t* d_wsrc = nullptr;    // 1: the packed copy's world angular velocity is derived
  // history (memory) contact models: per-particle memory tables, one per
  // state copy; allocated once a memory model is set and particles exist
  bool has_memory = false;
  History hist[2]{};
  double dt_mem = 0.;
  // captured fused steps: [0] one step, [1] kGraphSteps steps in a row; one
  // graph per state-buffer parity
  bool use_graph = true;
  int64_t graph_gen = 0;
  hipGraphExec_t graph_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  double graph_dt[2][2] = {{0., 0.}, {0., 0.}};
  int64_t graph_made[2][2] = {{-1, -1}, {-1, -1}};
  // simulation time and imposed wall motions; while motions are pending,
  // gdem_step computes the walls of each of its steps on the host
  // (h_traj) and each force evaluation reads its own copy (step_walls)
  double time = 0.;
  WallKinematics kin;
  std::vector<Wall> h_traj;
  Wall* d_traj = nullptr;
  size_t traj_cap = 0;
  const Wall* step_walls = nullptr;
  // wall loads of the last force evaluation (when tracked): per-particle wall
  // contacts, block partials and totals
  bool track_loads = false;
  int32_t* d_wl_count = nullptr;
  int32_t* d_wl_wall = nullptr;
  double4* d_wl_f = nullptr;
  double4* d_wl_t = nullptr;
  double4* d_wl_part = nullptr;
  double4* d_wall_load = nullptr;
  std::vector<double> h_loads;  // host copy read by the imposed-force controller
};

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      fprintf(stderr, "gdem: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
      return GDEM_ERR_HIP;                                                   \
    }                                                                        \
  } while (0)

static int collect_errors(gdem_ctx* c);
static int adapt_skin(gdem_ctx* c);

static int set_err(gdem_ctx* c, int code, int id0, int id1) {
  c->err_code = code;
  c->err_id0 = id0;
  c->err_id1 = id1;
  return code;
}

static void free_state(State& s) {
  (void)hipFree(s.pos_r); (void)hipFree(s.vel_m); (void)hipFree(s.om_req); (void)hipFree(s.wbf_I); (void)hipFree(s.quat);
  (void)hipFree(s.aux);
  s = State{};
}

static void free_cstate(CState& s) {
  (void)hipFree(s.rot0); (void)hipFree(s.rot1); (void)hipFree(s.rot2); (void)hipFree(s.inert);
  s = CState{};
}

static int alloc_cstate(CState& s, int64_t n) {
  size_t b4 = sizeof(double4) * (size_t)std::max<int64_t>(n, 1);
  HIP_TRY(hipMalloc(&s.rot0, b4));
  HIP_TRY(hipMalloc(&s.rot1, b4));
  HIP_TRY(hipMalloc(&s.rot2, b4));
  HIP_TRY(hipMalloc(&s.inert, b4));
  return GDEM_OK;
}

static int alloc_state(State& s, int64_t n) {
  size_t b4 = sizeof(double4) * (size_t)std::max<int64_t>(n, 1);
  HIP_TRY(hipMalloc(&s.pos_r, b4));
  HIP_TRY(hipMalloc(&s.vel_m, b4));
  HIP_TRY(hipMalloc(&s.om_req, b4));
  HIP_TRY(hipMalloc(&s.wbf_I, b4));
  HIP_TRY(hipMalloc(&s.quat, b4));
  HIP_TRY(hipMalloc(&s.aux, sizeof(Aux) * (size_t)std::max<int64_t>(n, 1)));
  return GDEM_OK;
}

extern "C" const char* gdem_version(void) { return "gdem 0.1 (gfx950, fp64)"; }

extern "C" int gdem_create(const gdem_config* cfg, gdem_ctx** out) {
  if (!cfg || !out) return GDEM_ERR_ARG;
  *out = nullptr;
  gdem_ctx* c = new gdem_ctx();
  c->cfg = *cfg;
  c->dev = cfg->device;
  if (!make_grid(*cfg, c->grid)) { delete c; return GDEM_ERR_ARG; }
  const int64_t ncell = c->grid.ncell;
  if (hipSetDevice(c->dev) != hipSuccess) { delete c; return GDEM_ERR_HIP; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return GDEM_ERR_HIP; }
  std::vector<int32_t> rank = row_morton_rank(c->grid.nc);
  const int64_t nrows = (int64_t)c->grid.nc[1] * c->grid.nc[2];
  const int nb = (int)((ncell + kScanCells - 1) / kScanCells);
  c->bins.nbsum = nb;
  std::vector<int32_t> row_jk((size_t)nrows);
  for (int64_t q = 0; q < nrows; ++q) row_jk[(size_t)rank[(size_t)q]] = (int32_t)q;
  if (hipMalloc(&c->d_row_rank, sizeof(int32_t) * nrows) != hipSuccess ||
      hipMalloc(&c->d_row_jk, sizeof(int32_t) * nrows) != hipSuccess ||
      hipMalloc(&c->bins.cnt, sizeof(int32_t) * ncell) != hipSuccess ||
      hipMalloc(&c->bins.local, sizeof(int32_t) * ncell) != hipSuccess ||
      hipMalloc(&c->bins.bsum, sizeof(int32_t) * (nb + 1)) != hipSuccess ||
      hipMalloc(&c->bins.cell_start, sizeof(int32_t) * (ncell + 1)) != hipSuccess ||
      hipMalloc(&c->d_models, sizeof(c->h_models)) != hipSuccess ||
      hipMalloc(&c->d_walls, sizeof(c->h_walls)) != hipSuccess ||
      hipMalloc(&c->d_shapes, sizeof(c->h_shapes)) != hipSuccess ||
      hipMalloc(&c->d_wall_shapes, sizeof(c->h_wall_shapes)) != hipSuccess ||
      hipMalloc(&c->d_wall_sc, sizeof(c->h_wall_sc)) != hipSuccess ||
      hipMalloc(&c->d_err, sizeof(unsigned long long) * kErrWords) != hipSuccess ||
      hipMalloc(&c->d_counters, sizeof(unsigned long long) * 2) != hipSuccess) {
    gdem_destroy(c);
    return GDEM_ERR_HIP;
  }
  (void)hipMemcpy(c->d_row_rank, rank.data(), sizeof(int32_t) * nrows, hipMemcpyHostToDevice);
  (void)hipMemcpy(c->d_row_jk, row_jk.data(), sizeof(int32_t) * nrows, hipMemcpyHostToDevice);
  (void)hipMemset(c->bins.cnt, 0, sizeof(int32_t) * ncell);
  c->bins.ticket = c->bins.bsum + nb;
  (void)hipMemset(c->bins.bsum, 0, sizeof(int32_t) * (nb + 1));
  (void)hipMemset(c->d_err, 0xff, sizeof(unsigned long long) * kErrWords);
  (void)hipMemset(c->d_counters, 0, sizeof(unsigned long long) * 2);
  (void)hipMemcpy(c->d_models, c->h_models, sizeof(c->h_models), hipMemcpyHostToDevice);
  if (hipDeviceSynchronize() != hipSuccess) { gdem_destroy(c); return GDEM_ERR_HIP; }
  *out = c;
  return GDEM_OK;
}

static void free_history(gdem_ctx* c);

static void free_wall_loads(gdem_ctx* c) {
  (void)hipFree(c->d_wl_count); (void)hipFree(c->d_wl_wall); (void)hipFree(c->d_wl_f);
  (void)hipFree(c->d_wl_t); (void)hipFree(c->d_wl_part); (void)hipFree(c->d_wall_load);
  c->d_wl_count = c->d_wl_wall = nullptr;
  c->d_wl_f = c->d_wl_t = c->d_wl_part = c->d_wall_load = nullptr;
}

// Buffers of the wall-load reduction for the current particle count.
static int ensure_wall_loads(gdem_ctx* c) {
  free_wall_loads(c);
  c->graph_gen++;
  if (!c->track_loads) return GDEM_OK;
  const size_t n = (size_t)std::max<int64_t>(c->n, 1);
  const size_t nb = (n + 255) / 256;
  if (hipMalloc(&c->d_wl_count, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&c->d_wl_wall, sizeof(int32_t) * n * kWallLoadCap) != hipSuccess ||
      hipMalloc(&c->d_wl_f, sizeof(double4) * n * kWallLoadCap) != hipSuccess ||
      hipMalloc(&c->d_wl_t, sizeof(double4) * n * kWallLoadCap) != hipSuccess ||
      hipMalloc(&c->d_wl_part, sizeof(double4) * nb * kMaxWalls * 2) != hipSuccess ||
      hipMalloc(&c->d_wall_load, sizeof(double4) * kMaxWalls * 2) != hipSuccess) {
    free_wall_loads(c);
    return GDEM_ERR_HIP;
  }
  if (hipMemset(c->d_wl_count, 0, sizeof(int32_t) * n) != hipSuccess ||
      hipMemset(c->d_wall_load, 0, sizeof(double4) * kMaxWalls * 2) != hipSuccess)
    return GDEM_ERR_HIP;
  return GDEM_OK;
}

static void free_pack(gdem_ctx* c) {
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(c->pk[k].A); (void)hipFree(c->pk[k].B); (void)hipFree(c->pk[k].C);
    (void)hipFree(c->pk[k].D); (void)hipFree(c->pk[k].E);
    (void)hipFree(c->pkc[k].rr); (void)hipFree(c->pkc[k].mi); (void)hipFree(c->pkc[k].aux);
    c->pk[k] = Pack{};
    c->pkc[k] = PackConst{};
  }
  (void)hipFree(c->d_omask); (void)hipFree(c->d_imask); (void)hipFree(c->d_prec);
  (void)hipFree(c->d_queue); (void)hipFree(c->d_qcount); (void)hipFree(c->d_wsrc);
  c->d_omask = c->d_imask = nullptr;
  c->d_prec = nullptr;
  c->d_queue = c->d_qcount = c->d_wsrc = nullptr;
  c->packed = false;
}

// Capacity of one segment of the pair queue: the out entries of the
// screening workgroups that append to it (workgroups b = k mod kQueueSegs).
static int64_t queue_seg(const gdem_ctx* c) {
  const int64_t nblk = (std::max<int64_t>(c->n, 1) + 255) / 256;
  const int64_t nblk8 = (nblk + 7) / 8 * 8;
  return (nblk8 + kQueueSegs - 1) / kQueueSegs * 256 * kListCap;
}

// Buffers of the packed layout for the current particle count (allocated on
// first use: about 2 KB per particle, most of it the pair records).
static int alloc_pack(gdem_ctx* c) {
  if (c->pk[0].A) return GDEM_OK;
  const size_t n = (size_t)std::max<int64_t>(c->n, 1);
  for (int k = 0; k < 2; ++k) {
    HIP_TRY(hipMalloc(&c->pk[k].A, sizeof(double4) * n));
    HIP_TRY(hipMalloc(&c->pk[k].B, sizeof(double4) * n));
    HIP_TRY(hipMalloc(&c->pk[k].C, sizeof(double4) * n));
    HIP_TRY(hipMalloc(&c->pk[k].D, sizeof(double4) * n));
    HIP_TRY(hipMalloc(&c->pk[k].E, sizeof(double) * n));
    HIP_TRY(hipMalloc(&c->pkc[k].rr, sizeof(double2) * n));
    HIP_TRY(hipMalloc(&c->pkc[k].mi, sizeof(double2) * n));
    HIP_TRY(hipMalloc(&c->pkc[k].aux, sizeof(Aux) * n));
  }
  HIP_TRY(hipMalloc(&c->d_omask, sizeof(uint32_t) * n));
  HIP_TRY(hipMalloc(&c->d_imask, sizeof(uint32_t) * n));
  HIP_TRY(hipMalloc(&c->d_prec, sizeof(PairRec) * n * kListCap));
  HIP_TRY(hipMemset(c->d_imask, 0, sizeof(uint32_t) * n));
  // every out entry at most once per step: a segment holds the entries of
  // the screening workgroups (256 particles) that append to it
  HIP_TRY(hipMalloc(&c->d_queue, sizeof(int32_t) * (size_t)queue_seg(c) * kQueueSegs));
  HIP_TRY(hipMalloc(&c->d_qcount, sizeof(int32_t) * kQueueSegs));
  HIP_TRY(hipMemset(c->d_qcount, 0, sizeof(int32_t) * kQueueSegs));
  HIP_TRY(hipMalloc(&c->d_wsrc, sizeof(int32_t)));
  HIP_TRY(hipMemset(c->d_wsrc, 0, sizeof(int32_t)));
  return GDEM_OK;
}

// Brings the canonical state copy up to date from the packed layout (every
// call other than gdem_step reads and writes the canonical copies).
static int ensure_canon(gdem_ctx* c) {
  if (!c->packed) return GDEM_OK;
  launch_from_pack(c->pk[c->cur], c->pk[1 - c->cur], c->pkc[c->cur], c->buf[c->cur], c->d_wsrc,
                   c->bins.perm, c->d_skin, c->cur, (int32_t)c->n, c->stream);
  HIP_TRY(hipGetLastError());
  c->packed = false;
  c->graph_gen++;
  return GDEM_OK;
}

// The packed layout from the canonical copy (both constant copies filled).
static int ensure_packed(gdem_ctx* c) {
  if (c->packed) return GDEM_OK;
  const int ra = alloc_pack(c);
  if (ra) return ra;
  launch_to_pack(c->buf[c->cur], c->pk[c->cur], c->pkc[0], c->pkc[1], c->d_wsrc, (int32_t)c->n,
                 c->stream);
  HIP_TRY(hipGetLastError());
  c->packed = true;
  c->graph_gen++;
  return GDEM_OK;
}

// The packed layout serves gdem_step for spheres with skin lists, memoryless
// contact models, static walls, and neither contact recording, wall loads nor
// contact statistics (whose sums need the fixed evaluation order of the
// canonical kernels).
static bool any_rolling(const gdem_ctx* c) {
  for (const Model& m : c->h_models)
    if (model_has_rolling(m)) return true;
  return false;
}

static bool pack_eligible(const gdem_ctx* c) {
  return !c->convex && c->n > 0 && c->skin > 0. && !c->has_memory && c->rec_cap == 0 &&
         !c->track_loads && !c->stats_on && !any_rolling(c);
}

// The next evaluation rebuilds the skin lists; the drift bound restarts.
static int reset_skin_words(gdem_ctx* c) {
  if (!c->d_skin) return GDEM_OK;
  SkinWords w{};
  w.need[0] = w.need[1] = 1;
  HIP_TRY(hipMemcpy(c->d_skin, &w, sizeof(w), hipMemcpyHostToDevice));
  return GDEM_OK;
}

extern "C" int gdem_destroy(gdem_ctx* c) {
  if (!c) return GDEM_OK;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_history(c);
  free_state(c->buf[0]);
  free_state(c->buf[1]);
  (void)hipFree(c->d_F); (void)hipFree(c->d_T);
  (void)hipFree(c->bins.key); (void)hipFree(c->bins.slot); (void)hipFree(c->bins.perm); (void)hipFree(c->bins.pos);
  (void)hipFree(c->bins.ids);
  (void)hipFree(c->d_peak);
  (void)hipFree(c->d_out_list); (void)hipFree(c->d_in_list); (void)hipFree(c->d_out_count); (void)hipFree(c->d_in_count);
  (void)hipFree(c->bins.cnt); (void)hipFree(c->bins.local); (void)hipFree(c->bins.bsum); (void)hipFree(c->bins.cell_start);
  (void)hipFree(c->d_row_rank); (void)hipFree(c->d_row_jk); (void)hipFree(c->d_models); (void)hipFree(c->d_walls); (void)hipFree(c->d_err);
  (void)hipFree(c->d_counters); (void)hipFree(c->d_rec); (void)hipFree(c->d_rec_count); (void)hipFree(c->d_stage);
  free_cstate(c->cbuf[0]);
  free_cstate(c->cbuf[1]);
  (void)hipFree(c->d_cand); (void)hipFree(c->d_crec); (void)hipFree(c->d_mask); (void)hipFree(c->d_work);
  (void)hipFree(c->d_nwork); (void)hipFree(c->d_shapes);
  (void)hipFree(c->d_skin); (void)hipFree(c->d_cmask); (void)hipFree(c->d_path);
  (void)hipFree(c->d_blk_contacts); (void)hipFree(c->d_stats); (void)hipFree(c->d_stat_part);
  (void)hipFree(c->d_wall_shapes); (void)hipFree(c->d_wall_sc); (void)hipFree(c->d_traj);
  free_wall_loads(c);
  free_pack(c);
  for (auto& e : c->ev_pool) { (void)hipEventDestroy(e.start); (void)hipEventDestroy(e.stop); }
  for (auto& row : c->graph_exec)
    for (auto& g : row)
      if (g) (void)hipGraphExecDestroy(g);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GDEM_OK;
}

static void free_history(gdem_ctx* c) {
  for (History& h : c->hist) {
    (void)hipFree(h.count); (void)hipFree(h.key); (void)hipFree(h.kd); (void)hipFree(h.nrm);
    (void)hipFree(h.spr);
    h = History{};
  }
}

// Memory tables of the history models (both copies), empty: no contact has
// a memory yet.
static int ensure_history(gdem_ctx* c) {
  if (!c->has_memory || c->convex || c->n <= 0 || c->hist[0].count) return GDEM_OK;
  const size_t nn = (size_t)c->n, ne = nn * kHistCap;
  for (History& h : c->hist) {
    HIP_TRY(hipMalloc(&h.count, sizeof(int32_t) * nn));
    HIP_TRY(hipMalloc(&h.key, sizeof(int32_t) * ne));
    HIP_TRY(hipMalloc(&h.kd, sizeof(double4) * ne));
    HIP_TRY(hipMalloc(&h.nrm, sizeof(double4) * ne));
    HIP_TRY(hipMalloc(&h.spr, sizeof(double4) * ne));
    HIP_TRY(hipMemset(h.count, 0, sizeof(int32_t) * nn));
  }
  c->graph_gen++;
  return GDEM_OK;
}

// Frees the previous particle set and allocates the buffers of n particles
// (both state copies, torsors, binning, contact lists, recording).
static int alloc_particles(gdem_ctx* c, int64_t n, bool convex) {
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_state(c->buf[0]);
  free_state(c->buf[1]);
  free_cstate(c->cbuf[0]);
  free_cstate(c->cbuf[1]);
  (void)hipFree(c->d_F); (void)hipFree(c->d_T);
  (void)hipFree(c->bins.key); (void)hipFree(c->bins.slot); (void)hipFree(c->bins.perm); (void)hipFree(c->bins.pos);
  (void)hipFree(c->bins.ids);
  (void)hipFree(c->d_out_list); (void)hipFree(c->d_in_list); (void)hipFree(c->d_out_count); (void)hipFree(c->d_in_count);
  (void)hipFree(c->d_rec); (void)hipFree(c->d_rec_count);
  (void)hipFree(c->d_cand); (void)hipFree(c->d_crec); (void)hipFree(c->d_mask); (void)hipFree(c->d_work);
  (void)hipFree(c->d_nwork);
  (void)hipFree(c->d_skin); (void)hipFree(c->d_cmask); (void)hipFree(c->d_path);
  (void)hipFree(c->d_blk_contacts); (void)hipFree(c->d_stat_part);
  free_history(c);
  free_pack(c);
  c->d_skin = nullptr;
  c->d_cmask = nullptr;
  c->d_path = nullptr;
  c->d_blk_contacts = nullptr;
  c->d_stat_part = nullptr;
  c->skin = 0.;
  c->d_F = c->d_T = nullptr;
  c->bins.key = c->bins.slot = c->bins.perm = nullptr;
  c->bins.pos = nullptr;
  c->bins.ids = nullptr;
  c->d_out_list = c->d_in_list = nullptr;
  c->d_out_count = c->d_in_count = nullptr;
  c->d_rec = nullptr;
  c->d_rec_count = nullptr;
  c->d_cand = nullptr;
  c->d_crec = nullptr;
  c->d_mask = nullptr;
  c->d_work = nullptr;
  c->d_nwork = nullptr;
  c->n = n;
  c->convex = convex;
  c->rmax = 0.;
  c->graph_gen++;
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  if (alloc_state(c->buf[0], n) || alloc_state(c->buf[1], n)) return GDEM_ERR_HIP;
  if (convex) {
    if (alloc_cstate(c->cbuf[0], n) || alloc_cstate(c->cbuf[1], n)) return GDEM_ERR_HIP;
    HIP_TRY(hipMalloc(&c->d_cand, sizeof(int2) * nn * kCandCap));
    HIP_TRY(hipMalloc(&c->d_crec, sizeof(CRec) * nn * kCandCap));
    HIP_TRY(hipMalloc(&c->d_mask, sizeof(uint32_t) * nn));
    HIP_TRY(hipMalloc(&c->d_work, sizeof(int32_t) * nn * kCandCap));
    HIP_TRY(hipMalloc(&c->d_nwork, sizeof(int32_t)));
    HIP_TRY(hipMemset(c->d_nwork, 0, sizeof(int32_t)));
  } else {
    HIP_TRY(hipMalloc(&c->d_out_list, sizeof(int2) * nn * kListCap));
    HIP_TRY(hipMalloc(&c->d_skin, sizeof(SkinWords)));
    HIP_TRY(hipMalloc(&c->d_cmask, sizeof(unsigned long long) * nn));
    HIP_TRY(hipMalloc(&c->d_path, sizeof(float) * nn));
    HIP_TRY(hipMemset(c->d_path, 0, sizeof(float) * nn));
    const int rw = reset_skin_words(c);  // the first evaluation rebuilds
    if (rw) return rw;
  }
  // contact counts per workgroup of the sphere kernels, statistics parts
  // (two per workgroup: pairs, walls)
  c->nblk_slots = (int32_t)((((nn + 255) / 256 + 7) / 8) * 8);
  HIP_TRY(hipMalloc(&c->d_blk_contacts, sizeof(unsigned long long) * c->nblk_slots));
  HIP_TRY(hipMemset(c->d_blk_contacts, 0, sizeof(unsigned long long) * c->nblk_slots));
  HIP_TRY(hipMalloc(&c->d_stat_part, sizeof(StatPart) * 2 * (size_t)c->nblk_slots));
  const size_t in_cap = convex ? (size_t)kCandCap : (size_t)kListCap;
  HIP_TRY(hipMalloc(&c->d_F, sizeof(double4) * nn));
  HIP_TRY(hipMalloc(&c->d_T, sizeof(double4) * nn));
  HIP_TRY(hipMalloc(&c->bins.key, sizeof(int32_t) * nn));
  HIP_TRY(hipMalloc(&c->bins.slot, sizeof(int32_t) * nn));
  HIP_TRY(hipMalloc(&c->bins.perm, sizeof(int32_t) * nn));
  HIP_TRY(hipMalloc(&c->bins.pos, sizeof(float4) * nn));
  HIP_TRY(hipMalloc(&c->bins.ids, sizeof(int32_t) * nn));
  HIP_TRY(hipMalloc(&c->d_in_list, sizeof(int2) * nn * in_cap));
  HIP_TRY(hipMalloc(&c->d_out_count, sizeof(int32_t) * nn));
  HIP_TRY(hipMalloc(&c->d_in_count, sizeof(int32_t) * nn));
  HIP_TRY(hipMemset(c->d_in_count, 0, sizeof(int32_t) * nn));
  HIP_TRY(hipMemset(c->d_F, 0, sizeof(double4) * nn));
  HIP_TRY(hipMemset(c->d_T, 0, sizeof(double4) * nn));
  if (c->rec_cap > 0) {
    HIP_TRY(hipMalloc(&c->d_rec, sizeof(ContactRec) * nn * (size_t)c->rec_cap));
    HIP_TRY(hipMalloc(&c->d_rec_count, sizeof(int32_t) * nn));
  }
  c->primed = false;
  c->binned = false;
  c->rec_valid = false;
  if (c->track_loads && !convex) {
    const int rl = ensure_wall_loads(c);
    if (rl) return rl;
  }
  return ensure_history(c);
}

// Effective skin of the candidate lists: the requested one (automatic: 0.2
// r_max), limited so that r_i + r_j + skin stays within one cell edge (the
// half stencil then still finds every candidate); below 0.05 r_max the lists
// are off (rebuilt at every evaluation).
static void update_skin(gdem_ctx* c) {
  if (c->convex || c->d_skin == nullptr) {
    c->skin = 0.;
    return;
  }
  double edge = c->grid.edge[0];
  for (int d = 1; d < 3; ++d) edge = std::min(edge, c->grid.edge[d]);
  const double room = edge - 2. * c->rmax;
  const double want = c->skin_req < 0. ? 0.2 * c->rmax : c->skin_req;
  const double eff = std::min(want, room);
  // a skin much thinner than the particles would rebuild at almost every step
  c->skin = (want > 0. && eff >= 0.05 * c->rmax) ? eff : 0.;
  c->graph_gen++;
}

extern "C" int gdem_set_skin(gdem_ctx* c, double skin) {
  if (!c || !(skin == skin)) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int rl = ensure_canon(c);
  if (rl) return rl;
  c->skin_req = skin;
  update_skin(c);
  return reset_skin_words(c);  // rebuild at the next evaluation
}

extern "C" int gdem_get_skin(gdem_ctx* c, double* skin, int64_t* rebuilds) {
  if (!c) return GDEM_ERR_ARG;
  SkinWords w{};
  if (c->d_skin) {
    HIP_TRY(hipSetDevice(c->dev));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(&w, c->d_skin, sizeof(w), hipMemcpyDeviceToHost));
  }
  if (skin) *skin = c->skin;
  if (rebuilds) *rebuilds = w.rebuilds;
  return GDEM_OK;
}

extern "C" int gdem_set_particles(gdem_ctx* c, int64_t n, const double* x, const double* v,
                                  const double* w, const double* q, const double* radius,
                                  const double* density, const double* crust,
                                  const int32_t* material) {
  if (!c || n < 0 || n > 0x7fffffff) return GDEM_ERR_ARG;
  if (n > 0 && (!x || !v || !w || !q || !radius || !density || !crust)) return GDEM_ERR_ARG;
  int rc = alloc_particles(c, n, false);
  if (rc) return rc;
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  std::vector<double4> pr(nn), vm(nn), wr(nn), wb(nn), qq(nn);
  std::vector<Aux> ax(nn);
  for (int64_t i = 0; i < n; ++i) {
    const double rr = radius[i], rho = density[i];
    if (!(rr > 0.) || !(rho > 0.) || !(crust[i] > 0.)) return GDEM_ERR_ARG;
    int32_t mat = material ? material[i] : 0;
    if (mat < 0 || mat >= kMaxMaterials) return GDEM_ERR_ARG;
    sphere_particle(x + 3 * i, v + 3 * i, w + 3 * i, q + 4 * i, rr, rho, crust[i], mat,
                    (int32_t)i, pr[i], vm[i], wr[i], wb[i], qq[i], ax[i]);
    c->rmax = std::max(c->rmax, rr);
    c->crust_min = i == 0 ? crust[i] : std::min(c->crust_min, crust[i]);
  }
  c->h_radius.assign(radius, radius + n);
  c->h_density.assign(density, density + n);
  c->h_crust.assign(crust, crust + n);
  c->h_material.assign((size_t)n, 0);
  if (material) c->h_material.assign(material, material + n);
  State& s = c->buf[0];
  c->cur = 0;
  const size_t b4 = sizeof(double4) * (size_t)n;
  if (n > 0) {
    HIP_TRY(hipMemcpy(s.pos_r, pr.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.vel_m, vm.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.om_req, wr.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.wbf_I, wb.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.quat, qq.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.aux, ax.data(), sizeof(Aux) * (size_t)n, hipMemcpyHostToDevice));
  }
  update_skin(c);
  launch_keys(s, (int32_t)n, c->grid, c->d_row_rank, c->bins.key, c->d_err, c->stream);
  HIP_TRY(hipGetLastError());
  return GDEM_OK;
}

// ---------------------------------------------------------------- convex shapes
// Convex::bbox at identity (Geometry/src/Convex.cpp:29-41): half extent along
// axis d = (support(e_d).d - support(-e_d).d) / 2.
static void finish_shape(CShape& sh) {
  for (int d = 0; d < 3; ++d) {
    const v3 e = mk(d == 0 ? 1. : 0., d == 1 ? 1. : 0., d == 2 ? 1. : 0.);
    const double mx = dot(e, support(sh, e)), mn = dot(e, support(sh, neg(e)));
    sh.bbox_ext[d] = (mx - mn) / 2.;
  }
}

// GrainsExec::computeVolumeInertiaContrib (Base/src/GrainsExec.cpp:748-800):
// volume and inertia of the tetrahedron (0, A2, A3, A4) (F. Tonon, J. Math.
// Stat. 1 (2004) 8-11), accumulated.
static void tet_contrib(v3 A2, v3 A3, v3 A4, double& vol, double* in) {
  const double x1 = 0., x2 = A2.x, x3 = A3.x, x4 = A4.x, y1 = 0., y2 = A2.y, y3 = A3.y,
               y4 = A4.y, z1 = 0., z2 = A2.z, z3 = A3.z, z4 = A4.z;
  const double det = fabs((x2 - x1) * (y3 - y1) * (z4 - z1) + (y2 - y1) * (z3 - z1) * (x4 - x1) +
                          (z2 - z1) * (x3 - x1) * (y4 - y1) - (z2 - z1) * (y3 - y1) * (x4 - x1) -
                          (x2 - x1) * (z3 - z1) * (y4 - y1) - (y2 - y1) * (x3 - x1) * (z4 - z1));
  vol += det / 6.;
  // each moment is one left-to-right sum, in the reference's term order
  auto sq = [](double s, double a1, double a2, double a3, double a4) {
    return s + a1 * a1 + a1 * a2 + a2 * a2 + a1 * a3 + a2 * a3 + a3 * a3 + a1 * a4 + a2 * a4 +
           a3 * a4 + a4 * a4;
  };
  auto mixed = [](double a1, double a2, double a3, double a4, double b1, double b2, double b3,
                  double b4) {
    return 2. * a1 * b1 + a2 * b1 + a3 * b1 + a4 * b1 + a1 * b2 + 2. * a2 * b2 + a3 * b2 +
           a4 * b2 + a1 * b3 + a2 * b3 + 2. * a3 * b3 + a4 * b3 + a1 * b4 + a2 * b4 + a3 * b4 +
           2. * a4 * b4;
  };
  in[0] += det * sq(sq(0., y1, y2, y3, y4), z1, z2, z3, z4) / 60.;
  in[1] -= det * mixed(x1, x2, x3, x4, z1, z2, z3, z4) / 120.;
  in[2] -= det * mixed(x1, x2, x3, x4, y1, y2, y3, y4) / 120.;
  in[3] += det * sq(sq(0., x1, x2, x3, x4), z1, z2, z3, z4) / 60.;
  in[4] -= det * mixed(y1, y2, y3, y4, z1, z2, z3, z4) / 120.;
  in[5] += det * sq(sq(0., x1, x2, x3, x4), y1, y2, y3, y4) / 60.;
}

extern "C" int gdem_add_box_shape(gdem_ctx* c, const double* half, int32_t* shape_id) {
  if (!c || !half) return GDEM_ERR_ARG;
  if (c->nshapes >= kMaxShapes) return GDEM_ERR_CAPACITY;
  if (!(half[0] > 0. && half[1] > 0. && half[2] > 0.)) return GDEM_ERR_ARG;
  CShape sh{};
  sh.type = SHAPE_BOX;
  for (int d = 0; d < 3; ++d) sh.ext[d] = half[d];
  const double ex = half[0], ey = half[1], ez = half[2];
  // Box::getVolume (Box.cpp:273-279), Box::BuildInertia (:181-205),
  // Box::computeCircumscribedRadius (:226-229)
  sh.vol = 8.0 * ex * ey * ez;
  sh.inertia[0] = 8.0 * ex * ey * ez * (ey * ey + ez * ez) / 3.0;
  sh.inertia[3] = 8.0 * ex * ey * ez * (ex * ex + ez * ez) / 3.0;
  sh.inertia[5] = 8.0 * ex * ey * ez * (ey * ey + ex * ex) / 3.0;
  sh.circ = norm(mk(ex, ey, ez));
  finish_shape(sh);
  c->h_shapes[c->nshapes] = sh;
  if (shape_id) *shape_id = c->nshapes;
  c->nshapes++;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipMemcpy(c->d_shapes, c->h_shapes, sizeof(CShape) * c->nshapes, hipMemcpyHostToDevice));
  return GDEM_OK;
}

extern "C" int gdem_add_polyhedron_shape(gdem_ctx* c, int32_t nv, const double* verts,
                                         int32_t nf, const int32_t* face_size,
                                         const int32_t* face_idx, int32_t* shape_id) {
  if (!c || !verts || !face_size || !face_idx) return GDEM_ERR_ARG;
  if (c->nshapes >= kMaxShapes) return GDEM_ERR_CAPACITY;
  if (nv < 4 || nv > kMaxVerts || nf < 4) return GDEM_ERR_ARG;
  CShape sh{};
  sh.type = SHAPE_POLYHEDRON;
  sh.nv = nv;
  for (int i = 0; i < nv; ++i)
    for (int d = 0; d < 3; ++d) sh.v[i][d] = verts[3 * i + d];
  auto V = [&](int i) { return mk(sh.v[i][0], sh.v[i][1], sh.v[i][2]); };
  int off = 0;
  for (int f = 0; f < nf; ++f) {
    const int m = face_size[f];
    if (m < 3) return GDEM_ERR_ARG;
    const int32_t* fi = face_idx + off;
    for (int j = 0; j < m; ++j)
      if (fi[j] < 0 || fi[j] >= nv) return GDEM_ERR_ARG;
    // adjacency: each vertex is followed by the next vertex of the face loop
    for (int j = 0, k = m - 1; j < m; k = j++) {
      const int a = fi[k];
      if (sh.cob_n[a] >= kMaxVerts) return GDEM_ERR_ARG;
      sh.cob[a][sh.cob_n[a]++] = fi[j];
    }
    if (m == 3) {
      tet_contrib(V(fi[0]), V(fi[1]), V(fi[2]), sh.vol, sh.inertia);
    } else {
      v3 G = mk(0., 0., 0.);
      for (int j = 0; j < m; ++j) G = add(G, V(fi[j]));
      G = dvd(G, (double)m);
      for (int j = 0, k = m - 1; j < m; k = j++) tet_contrib(G, V(fi[k]), V(fi[j]), sh.vol, sh.inertia);
    }
    off += m;
  }
  // Polyhedron.cpp:362-364: the support search starts at the first vertex
  // with neighbours
  sh.first_vertex = 0;
  while (sh.first_vertex < nv && sh.cob_n[sh.first_vertex] == 0) ++sh.first_vertex;
  if (sh.first_vertex == nv) return GDEM_ERR_ARG;
  // Polyhedron::computeCircumscribedRadius (Polyhedron.cpp:201-210)
  double ray = dot(V(0), V(0));
  for (int i = 1; i < nv; ++i) {
    const double d = dot(V(i), V(i));
    if (d > ray) ray = d;
  }
  sh.circ = sqrt(ray);
  finish_shape(sh);
  c->h_shapes[c->nshapes] = sh;
  if (shape_id) *shape_id = c->nshapes;
  c->nshapes++;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipMemcpy(c->d_shapes, c->h_shapes, sizeof(CShape) * c->nshapes, hipMemcpyHostToDevice));
  return GDEM_OK;
}

// Crust scaling of a body with the given bounding-box half extents
// (RigidBodyWithCrust.cpp:51-80).
static void crust_scale(const double* ext, double crust, double* sc) {
  for (int d = 0; d < 3; ++d) sc[d] = ext[d] < kEps ? 1. : (ext[d] - crust) / ext[d];
}

extern "C" int gdem_set_convex_particles(gdem_ctx* c, int64_t n, const double* x, const double* v,
                                         const double* w, const double* q, const double* basis,
                                         const int32_t* shape, const double* density,
                                         const double* crust, const int32_t* material) {
  if (!c || n < 0 || n > 0x7fffffff) return GDEM_ERR_ARG;
  if (n > 0 && (!x || !v || !w || !q || !basis || !shape || !density || !crust))
    return GDEM_ERR_ARG;
  if (c->has_memory) return GDEM_ERR_ARG;  // memory models: spheres only
  if (any_rolling(c)) return GDEM_ERR_ARG;  // rolling resistance: spheres only
  for (int64_t i = 0; i < n; ++i) {
    if (shape[i] < 0 || shape[i] >= c->nshapes) return GDEM_ERR_ARG;
    if (!(density[i] > 0.) || !(crust[i] > 0.)) return GDEM_ERR_ARG;
    const int32_t mat = material ? material[i] : 0;
    if (mat < 0 || mat >= kMaxMaterials) return GDEM_ERR_ARG;
  }
  int rc = alloc_particles(c, n, true);
  if (rc) return rc;
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  std::vector<double4> pr(nn), vm(nn), wr(nn), wb(nn), qq(nn), r0(nn), r1(nn), r2(nn), in3(nn);
  std::vector<Aux> ax(nn);
  for (int64_t i = 0; i < n; ++i) {
    const CShape& sh = c->h_shapes[shape[i]];
    const double rho = density[i];
    const int32_t mat = material ? material[i] : 0;
    // Particle.cpp:158-164, Component::getEquivalentSphereRadius (Component.cpp:205-208)
    const double mass = rho * sh.vol;
    const double req = pow((0.75 / kPi) * sh.vol, 1. / 3.);
    double sc[3];
    crust_scale(sh.bbox_ext, crust[i], sc);
    const quat qi = qmk(mk(q[4 * i], q[4 * i + 1], q[4 * i + 2]), q[4 * i + 3]);
    const v3 wbf = qmulv(qconj(qi), qleftvec(qi, mk(w[3 * i], w[3 * i + 1], w[3 * i + 2])));
    const double* b = basis + 9 * i;
    pr[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], sh.circ);
    vm[i] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], mass);
    wr[i] = make_double4(w[3 * i], w[3 * i + 1], w[3 * i + 2], req);
    wb[i] = make_double4(wbf.x, wbf.y, wbf.z, 0.);
    qq[i] = make_double4(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
    r0[i] = make_double4(b[0], b[1], b[2], sc[0]);
    r1[i] = make_double4(b[3], b[4], b[5], sc[1]);
    r2[i] = make_double4(b[6], b[7], b[8], sc[2]);
    in3[i] = make_double4(sh.inertia[0] * rho, sh.inertia[3] * rho, sh.inertia[5] * rho, 0.);
    ax[i].id = (int32_t)i;
    ax[i].mat = mat | (shape[i] << 8);
    ax[i].crust = crust[i];
  }
  State& s = c->buf[0];
  CState& cs = c->cbuf[0];
  c->cur = 0;
  const size_t b4 = sizeof(double4) * (size_t)n;
  if (n > 0) {
    HIP_TRY(hipMemcpy(s.pos_r, pr.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.vel_m, vm.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.om_req, wr.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.wbf_I, wb.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.quat, qq.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s.aux, ax.data(), sizeof(Aux) * (size_t)n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cs.rot0, r0.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cs.rot1, r1.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cs.rot2, r2.data(), b4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(cs.inert, in3.data(), b4, hipMemcpyHostToDevice));
  }
  launch_keys(s, (int32_t)n, c->grid, c->d_row_rank, c->bins.key, c->d_err, c->stream);
  HIP_TRY(hipGetLastError());
  return GDEM_OK;
}

extern "C" int gdem_get_basis(gdem_ctx* c, double* basis) {
  if (!c || !basis) return GDEM_ERR_ARG;
  if (!c->convex) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->n;
  const size_t need = n * 9;
  if (c->stage_len < need) {
    (void)hipFree(c->d_stage);
    c->d_stage = nullptr;
    HIP_TRY(hipMalloc(&c->d_stage, sizeof(double) * std::max<size_t>(need, 1)));
    c->stage_len = need;
  }
  launch_cunpermute(c->buf[c->cur], c->cbuf[c->cur], (int32_t)n, c->d_stage, c->stream);
  HIP_TRY(hipGetLastError());
  if (n > 0)
    HIP_TRY(hipMemcpyAsync(basis, c->d_stage, sizeof(double) * need, hipMemcpyDeviceToHost,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return collect_errors(c);
}

static int add_wall_kind(gdem_ctx* c, const double* basis, const double* centre,
                         const double* half, double crust, int32_t material, int32_t kind) {
  if (c->nwalls >= kMaxWalls || material < 0 || material >= kMaxMaterials) return GDEM_ERR_ARG;
  const Wall wl = make_wall(basis, centre, half, crust, material, kind);
  const double e0 = half[0], e1 = half[1], e2 = half[2];
  // the wall as a convex body for the GJK narrow phase of convex particles:
  // Box shape, crust scaling of its extents, circumscribed radius
  CShape wsh{};
  wsh.type = SHAPE_BOX;
  for (int d = 0; d < 3; ++d) wsh.ext[d] = wsh.bbox_ext[d] = half[d];
  double wsc[3];
  crust_scale(half, crust, wsc);
  c->h_wall_shapes[c->nwalls] = wsh;
  c->h_wall_sc[c->nwalls] = make_double4(wsc[0], wsc[1], wsc[2], norm(mk(e0, e1, e2)));
  c->h_walls[c->nwalls++] = wl;
  c->graph_gen++;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipMemcpyAsync(c->d_walls, c->h_walls, sizeof(Wall) * c->nwalls, hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_wall_shapes, c->h_wall_shapes, sizeof(CShape) * c->nwalls,
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_wall_sc, c->h_wall_sc, sizeof(double4) * c->nwalls,
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GDEM_OK;
}

extern "C" int gdem_add_wall(gdem_ctx* c, const double* basis, const double* centre,
                             const double* half, double crust, int32_t material) {
  if (!c || !basis || !centre || !half) return GDEM_ERR_ARG;
  return add_wall_kind(c, basis, centre, half, crust, material, WALL_BOX);
}

extern "C" int gdem_add_rectangle_wall(gdem_ctx* c, const double* basis, const double* centre,
                                       double lx, double ly, double crust, int32_t material) {
  if (!c || !basis || !centre || !(lx > 0.) || !(ly > 0.)) return GDEM_ERR_ARG;
  const double half[3] = {lx / 2., ly / 2., 0.};
  return add_wall_kind(c, basis, centre, half, crust, material, WALL_RECTANGLE);
}

extern "C" int gdem_set_contact_model(gdem_ctx* c, int32_t a, int32_t b, int32_t kind,
                                      const double* p) {
  if (!c || !p || a < 0 || b < 0 || a >= kMaxMaterials || b >= kMaxMaterials) return GDEM_ERR_ARG;
  if (!model_params_ok(kind, p)) return GDEM_ERR_ARG;
  if (model_has_memory(kind) && c->convex) return GDEM_ERR_ARG;  // spheres only
  const Model m = make_model(kind, p);
  if (model_has_rolling(m) && c->convex) return GDEM_ERR_ARG;  // spheres only
  c->h_models[a * kMaxMaterials + b] = m;
  c->h_models[b * kMaxMaterials + a] = m;
  bool mem = false;
  for (const Model& md : c->h_models) mem = mem || model_has_memory(md.kind);
  if (mem != c->has_memory) c->graph_gen++;  // the force kernel variant changes
  c->has_memory = mem;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipMemcpyAsync(c->d_models, c->h_models, sizeof(c->h_models), hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return ensure_history(c);
}

extern "C" int gdem_set_time_step(gdem_ctx* c, double dt) {
  if (!c || !(dt > 0.)) return GDEM_ERR_ARG;
  if (dt != c->dt_mem) c->graph_gen++;
  c->dt_mem = dt;
  return GDEM_OK;
}

static PosSrc canon_src(const State& s) { return PosSrc{s.pos_r, nullptr, s.aux}; }

// Rebuild limit of the drift bound (SkinWords): skin / 2, less a relative
// 1e-6 that dwarfs the rounding of the positions and of the bound.
static float skin_limit(const gdem_ctx* c) {
  const double x = 0.5 * c->skin * (1. - 1e-6);
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, 0.f);  // rounded down
  return f;
}

static int do_binning(gdem_ctx* c) {
  // with skin lists the binning is part of the next force evaluation, which
  // runs it when a rebuild is due
  if (c->skin > 0.) return GDEM_OK;
  launch_binning(c->bins, canon_src(c->buf[c->cur]), (int32_t)c->n, c->grid, c->d_row_rank,
                 c->d_err, c->stream);
  HIP_TRY(hipGetLastError());
  c->binned = true;
  return GDEM_OK;
}

static TimedLaunch* next_event_pair(gdem_ctx* c, int kind) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return nullptr;
    c->ev_pool.push_back(TimedLaunch{a, b, 0});
  }
  TimedLaunch* e = &c->ev_pool[c->ev_used++];
  e->kind = kind;
  return e;
}

static void launch_convex_step(gdem_ctx* c, int mode, double dtv, double dtd, bool write_ft,
                               bool record, int nblk8) {
  CDetectArgs d{};
  d.pos = c->bins.pos;
  d.margin = prefilter_margin(c->grid, d.sizef);
  d.in = c->buf[c->cur];
  d.perm = c->bins.perm;
  d.cell_start = c->bins.cell_start;
  d.row_rank = c->d_row_rank;
  d.grid = c->grid;
  d.n = (int32_t)c->n;
  d.cand = c->d_cand;
  d.cand_count = c->d_out_count;
  d.mask = c->d_mask;
  d.work = c->d_work;
  d.nwork = c->d_nwork;
  d.walls = c->d_walls;
  d.nwalls = c->nwalls;
  d.err = c->d_err;
  d.nblk8 = nblk8;
  CGjkArgs g{};
  g.in = c->buf[c->cur];
  g.cin = c->cbuf[c->cur];
  g.shapes = c->d_shapes;
  g.perm = c->bins.perm;
  g.grid = c->grid;
  g.cand = c->d_cand;
  g.work = c->d_work;
  g.nwork = c->d_nwork;
  g.walls = c->d_walls;
  g.wall_shapes = c->d_wall_shapes;
  g.wall_sc = c->d_wall_sc;
  g.rec = c->d_crec;
  g.mask = c->d_mask;
  g.in_list = c->d_in_list;
  g.in_count = c->d_in_count;
  g.err = c->d_err;
  g.stats = c->stats_on ? 1 : 0;
  g.gjk_sv = c->gjk_sv;
  CForceArgs a{};
  a.in = c->buf[c->cur];
  a.out = c->buf[1 - c->cur];
  a.cin = c->cbuf[c->cur];
  a.cout = c->cbuf[1 - c->cur];
  a.shapes = c->d_shapes;
  a.perm = c->bins.perm;
  a.row_rank = c->d_row_rank;
  a.grid = c->grid;
  a.models = c->d_models;
  a.walls = c->d_walls;
  a.n = (int32_t)c->n;
  for (int dd = 0; dd < 3; ++dd) a.gravity[dd] = c->cfg.gravity[dd];
  a.dt_vel = dtv;
  a.dt_disp = dtd;
  a.cand = c->d_cand;
  a.cand_count = c->d_out_count;
  a.rec = c->d_crec;
  a.mask = c->d_mask;
  a.in_list = c->d_in_list;
  a.in_count = c->d_in_count;
  a.nwork = c->d_nwork;
  a.F = write_ft ? c->d_F : nullptr;
  a.T = write_ft ? c->d_T : nullptr;
  a.key_out = c->bins.key;
  a.err = c->d_err;
  a.contacts = c->d_counters;
  a.crec = c->d_rec;
  a.rec_count = c->d_rec_count;
  a.rec_cap = c->rec_cap;
  a.nblk8 = nblk8;
  a.stat = c->stats_on ? c->d_stat_part : nullptr;
  launch_cdetect(d, c->stream);
  // a queue of about 4 candidates per particle is covered in one pass
  launch_cgjk(g, (int32_t)std::min<int64_t>(4 * c->n, 0x7fffffff), c->stream);
  launch_cforce(mode, record, a, c->stream);
  if (c->stats_on)
    launch_stats_final(c->d_stat_part, nblk8, c->d_stats, mode == MODE_FORCE ? 0. : dtd,
                       c->stream);
}

// Kernels of one evaluation in the packed layout: gated binning, pair kernel
// (candidate sweep on a rebuild), integrator kernel.
static int pack_launches(gdem_ctx* c, int mode, double dtv, double dtd, bool write_ft) {
  const int p = c->cur;
  const SkinGate gate{c->d_skin, p, skin_limit(c)};
  const PackConst& kin = c->pkc[p];
  const PosSrc src{c->pk[p].A, kin.rr, kin.aux};
  const int nblk = (int)((c->n + 255) / 256);
  const int nblk8 = ((nblk + 7) / 8) * 8;
  PairArgs pa{};
  DetectArgs& d = pa.d;
  d.pos = c->bins.pos;
  d.ids = c->bins.ids;
  d.margin = prefilter_margin(c->grid, d.sizef);
  d.perm = c->bins.perm;
  d.cell_start = c->bins.cell_start;
  d.row_rank = c->d_row_rank;
  d.grid = c->grid;
  d.n = (int32_t)c->n;
  d.lstride = (int32_t)c->n;
  d.out_list = c->d_out_list;
  d.out_count = c->d_out_count;
  d.in_list = c->d_in_list;
  d.in_count = c->d_in_count;
  d.err = c->d_err;
  d.nblk8 = nblk8;
  d.skin = c->skin;
  d.gate = gate;
  pa.in = c->pk[p];
  pa.other = c->pk[1 - p];
  pa.wsrc = c->d_wsrc;
  pa.kin = kin;
  pa.omask = c->d_omask;
  pa.imask = c->d_imask;
  pa.rec = c->d_prec;
  pa.models = c->d_models;
  pa.model_stride = kMaxMaterials;
  pa.contacts = c->d_blk_contacts;
  pa.queue = c->d_queue;
  pa.qcount = c->d_qcount;
  pa.qseg = (int32_t)std::min<int64_t>(queue_seg(c), 0x7fffffff);
  TimedLaunch* ev = c->timing ? next_event_pair(c, 1) : nullptr;
  launch_pair(c->bins, src, c->d_skin, pa, c->stream, ev ? ev->start : nullptr,
              ev ? ev->stop : nullptr);
  HIP_TRY(hipGetLastError());
  IntegArgs a{};
  a.in = c->pk[p];
  a.out = c->pk[1 - p];
  a.wsrc = c->d_wsrc;
  a.kin = kin;
  a.kout = c->pkc[1 - p];
  a.perm = c->bins.perm;
  a.grid = c->grid;
  for (int dd = 0; dd < 3; ++dd) a.gmax[dd] = c->grid.nc[dd] * c->grid.edge[dd] * (1. - 1e-12);
  a.models = c->d_models;
  a.model_stride = kMaxMaterials;
  a.walls = c->d_walls;
  a.nwalls = c->nwalls;
  a.n = (int32_t)c->n;
  a.lstride = (int32_t)c->n;
  for (int dd = 0; dd < 3; ++dd) a.gravity[dd] = c->cfg.gravity[dd];
  a.dt_vel = dtv;
  a.dt_disp = dtd;
  a.crust_min = c->crust_min;
  a.out_list = c->d_out_list;
  a.in_list = c->d_in_list;
  a.in_count = c->d_in_count;
  a.omask = c->d_omask;
  a.imask = c->d_imask;
  a.rec = c->d_prec;
  a.F = write_ft ? c->d_F : nullptr;
  a.T = write_ft ? c->d_T : nullptr;
  a.gate = gate;
  a.words = c->d_skin;
  a.path = c->d_path;
  a.err = c->d_err;
  a.contacts = c->d_blk_contacts;
  a.nblk8 = nblk8;
  a.qcount = c->d_qcount;
  ev = c->timing ? next_event_pair(c, 0) : nullptr;
  launch_integ(mode, a, c->stream, ev ? ev->start : nullptr, ev ? ev->stop : nullptr);
  HIP_TRY(hipGetLastError());
  return GDEM_OK;
}

// Kernels of one force evaluation on the current copy (binning first when
// positions changed); no host bookkeeping, so the sequence can be captured.
static int force_launches(gdem_ctx* c, int mode, double dtv, double dtd, bool write_ft) {
  if (c->packed) return pack_launches(c, mode, dtv, dtd, write_ft);
  const bool skin = !c->convex && c->skin > 0.;
  const SkinGate gate = skin ? SkinGate{c->d_skin, c->cur, skin_limit(c)} : SkinGate{};
  if (skin) {
    // rebuild (binning + candidate sweep) only when some particle's path
    // since the last rebuild passed skin/2
    launch_binning(c->bins, canon_src(c->buf[c->cur]), (int32_t)c->n, c->grid, c->d_row_rank,
                   c->d_err, c->stream, gate, c->d_skin, c->d_in_count);
    HIP_TRY(hipGetLastError());
  } else if (!c->binned) {
    launch_binning(c->bins, canon_src(c->buf[c->cur]), (int32_t)c->n, c->grid, c->d_row_rank,
                   c->d_err, c->stream);
    HIP_TRY(hipGetLastError());
  }
  const int nblk = (int)((c->n + 255) / 256);
  const int nblk8 = ((nblk + 7) / 8) * 8;
  const bool record = c->rec_cap > 0 && c->d_rec != nullptr;
  TimedLaunch* ev = nullptr;
  if (c->convex) {
    // events bracket detection + force: the GJK narrow phase runs in detection
    if (c->timing) {
      ev = next_event_pair(c, 0);
      if (ev) HIP_TRY(hipEventRecord(ev->start, c->stream));
    }
    launch_convex_step(c, mode, dtv, dtd, write_ft, record, nblk8);
  } else {
    DetectArgs d{};
    d.pos = c->bins.pos;
    d.ids = c->bins.ids;
    d.margin = prefilter_margin(c->grid, d.sizef);
    d.in = c->buf[c->cur];
    d.perm = c->bins.perm;
    d.cell_start = c->bins.cell_start;
    d.row_rank = c->d_row_rank;
    d.grid = c->grid;
    d.n = (int32_t)c->n;
    d.lstride = (int32_t)c->n;
    d.out_list = c->d_out_list;
    d.out_count = c->d_out_count;
    d.in_list = c->d_in_list;
    d.in_count = c->d_in_count;
    d.err = c->d_err;
    d.nblk8 = nblk8;
    d.skin = skin ? c->skin : 0.;
    d.gate = gate;
    ForceArgs a{};
    a.in = c->buf[c->cur];
    a.out = c->buf[1 - c->cur];
    a.perm = c->bins.perm;
    a.row_rank = c->d_row_rank;
    a.grid = c->grid;
    a.models = c->d_models;
    a.model_stride = kMaxMaterials;
    a.walls = c->step_walls ? c->step_walls : c->d_walls;
    a.nwalls = c->nwalls;
    a.n = (int32_t)c->n;
    for (int dd = 0; dd < 3; ++dd) a.gravity[dd] = c->cfg.gravity[dd];
    a.dt_vel = dtv;
    a.dt_disp = dtd;
    a.lstride = (int32_t)c->n;
    a.out_list = c->d_out_list;
    a.out_count = c->d_out_count;
    a.in_list = c->d_in_list;
    a.in_count = c->d_in_count;
    a.F = write_ft ? c->d_F : nullptr;
    a.T = write_ft ? c->d_T : nullptr;
    a.key_out = c->bins.key;
    a.err = c->d_err;
    a.contacts = c->d_blk_contacts;
    a.stat = c->stats_on ? c->d_stat_part : nullptr;
    a.rec = c->d_rec;
    a.rec_count = c->d_rec_count;
    a.rec_cap = c->rec_cap;
    a.nblk8 = nblk8;
    if (skin) {
      a.gate = gate;
      a.words = c->d_skin;
      a.path = c->d_path;
      a.cmask = c->d_cmask;
    }
    if (c->has_memory) {
      if (!(c->dt_mem > 0.) || !c->hist[0].count) return GDEM_ERR_ARG;
      a.hin = c->hist[c->cur];
      a.hout = c->hist[1 - c->cur];
      a.dt_mem = c->dt_mem;
    }
    if (c->track_loads) {
      if (!c->d_wl_count) return GDEM_ERR_ARG;
      a.wl_count = c->d_wl_count;
      a.wl_wall = c->d_wl_wall;
      a.wl_f = c->d_wl_f;
      a.wl_t = c->d_wl_t;
    }
    if (skin)
      launch_detect_screen(d, a, c->stream);
    else
      launch_detect(d, c->stream);
    // timing: the events take the force kernel's own start and end times
    // (an event recorded between two launches would add the dispatch gap)
    if (c->timing) ev = next_event_pair(c, 0);
    launch_force(mode, record, a, c->stream, ev ? ev->start : nullptr, ev ? ev->stop : nullptr,
                 c->has_memory);
    HIP_TRY(hipGetLastError());
    if (c->track_loads) {
      launch_wall_loads(a, c->d_wl_part, c->d_wall_load, c->stream);
      HIP_TRY(hipGetLastError());
    }
    if (c->stats_on) {
      launch_stats_final(c->d_stat_part, nblk8, c->d_stats, mode == MODE_FORCE ? 0. : dtd,
                         c->stream);
      HIP_TRY(hipGetLastError());
    }
    return GDEM_OK;
  }
  HIP_TRY(hipGetLastError());
  if (ev) HIP_TRY(hipEventRecord(ev->stop, c->stream));
  return GDEM_OK;
}

static void after_force(gdem_ctx* c) {
  c->cur = 1 - c->cur;
  c->binned = false;
  c->force_evals++;
  if (c->rec_cap > 0 && c->d_rec != nullptr) c->rec_valid = true;
}

static int do_force(gdem_ctx* c, int mode, double dtv, double dtd, bool write_ft) {
  const int rc = force_launches(c, mode, dtv, dtd, write_ft);
  if (rc) return rc;
  after_force(c);
  return GDEM_OK;
}

// Steps of several fused steps per graph: fewer graph launches, and the
// work of consecutive steps is not separated by the end of a graph.
constexpr int kGraphSteps = 16;

// `steps` fused steps (binning, detection, force kernel with MODE_MOVE; 1 or
// kGraphSteps) replayed from a graph captured on first use for the current
// buffer parity and dt.  Only used without contact recording and kernel
// timing; the captured kernels clear their own counters (no memset nodes).
static int graph_force(gdem_ctx* c, double dt, int steps) {
  const int p = c->cur;
  const int k = steps > 1 ? 1 : 0;
  hipGraphExec_t& ge = c->graph_exec[k][p];
  if (!ge || c->graph_dt[k][p] != dt || c->graph_made[k][p] != c->graph_gen) {
    if (ge) {
      HIP_TRY(hipGraphExecDestroy(ge));
      ge = nullptr;
    }
    const bool binned = c->binned;
    c->binned = false;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = GDEM_OK;
    for (int s = 0; s < steps && rc == GDEM_OK; ++s) {
      rc = force_launches(c, MODE_MOVE, 0.5 * dt, dt, false);
      c->cur = 1 - c->cur;  // the next captured step reads the copy this one writes
    }
    c->cur = p;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->binned = binned;
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_TRY(e);
    const hipError_t ei = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(ei);
    c->graph_dt[k][p] = dt;
    c->graph_made[k][p] = c->graph_gen;
  }
  HIP_TRY(hipGraphLaunch(ge, c->stream));
  for (int s = 0; s < steps; ++s) after_force(c);
  return GDEM_OK;
}

extern "C" int gdem_link_update(gdem_ctx* c) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const int rl = ensure_canon(c);
  if (rl) return rl;
  return do_binning(c);
}

extern "C" int gdem_compute_forces(gdem_ctx* c) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const int rl = ensure_canon(c);
  if (rl) return rl;
  return do_force(c, MODE_FORCE, 0., 0., true);
}

static void note_move_step(gdem_ctx* c, double dtd) {
  if (dtd > 0. && dtd != c->dt_mem) {
    c->dt_mem = dtd;  // the step of the next force evaluation
    c->graph_gen++;
  }
}

extern "C" int gdem_move(gdem_ctx* c, double dtv, double dtd) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const int rl = ensure_canon(c);
  if (rl) return rl;
  note_move_step(c, dtd);
  if (c->convex)
    launch_cmove(c->buf[c->cur], c->cbuf[c->cur], c->d_F, c->d_T, (int32_t)c->n, c->grid,
                 c->d_row_rank, dtv, dtd, c->bins.key, c->d_err, c->stream);
  else if (c->skin > 0.)
    launch_move(c->buf[c->cur], c->d_F, c->d_T, (int32_t)c->n, c->grid, c->d_row_rank, dtv, dtd,
                c->bins.key, c->d_err, c->stream, SkinGate{c->d_skin, c->cur, skin_limit(c)},
                c->d_path);
  else
    launch_move(c->buf[c->cur], c->d_F, c->d_T, (int32_t)c->n, c->grid, c->d_row_rank, dtv, dtd,
                c->bins.key, c->d_err, c->stream);
  HIP_TRY(hipGetLastError());
  c->binned = false;
  return GDEM_OK;
}

extern "C" int gdem_advance_velocity(gdem_ctx* c, double dtv) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const int rl = ensure_canon(c);
  if (rl) return rl;
  if (c->convex)
    launch_cadvance(c->buf[c->cur], c->cbuf[c->cur], c->d_F, c->d_T, (int32_t)c->n, dtv,
                    c->stream);
  else
    launch_advance(c->buf[c->cur], c->d_F, c->d_T, (int32_t)c->n, dtv, c->stream);
  HIP_TRY(hipGetLastError());
  return GDEM_OK;
}

static int upload_walls(gdem_ctx* c) {
  HIP_TRY(hipMemcpyAsync(c->d_walls, c->h_walls, sizeof(Wall) * c->nwalls, hipMemcpyHostToDevice,
                         c->stream));
  // h_walls changes on the next move: the copy has to be done by then
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GDEM_OK;
}

extern "C" int gdem_prime(gdem_ctx* c) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const int rl = ensure_canon(c);
  if (rl) return rl;
  // initial wall velocities (Grains::do_before_time_stepping,
  // Base/src/Grains.cpp:123-124)
  if (c->dt_mem > 0. && c->kin.init(c->h_walls, c->nwalls, c->time, c->dt_mem)) {
    const int ru = upload_walls(c);
    if (ru) return ru;
  }
  int rc = do_binning(c);
  if (rc) return rc;
  rc = do_force(c, MODE_FORCE, 0., 0., true);
  if (rc) return rc;
  c->primed = true;
  c->prime_evals = c->force_evals;
  return GDEM_OK;
}

// gdem_step while wall motions are pending (the fused step continues after
// the first kick + drift): the walls of steps k = 1..nsteps, moved over
// [t_k - dt, t_k] as AllComponents::Move does after the particles
// (Component/src/AllComponents.cpp:151-220), are computed on the host in
// chunks and uploaded; force evaluation k reads walls at t_k.  No graph
// replay (each evaluation reads its own wall array).
// The loads of the last force evaluation into h_loads for the controller of
// the imposed forces (Obstacle::getForce in ObstacleImposedForce.cpp:232).
static int fetch_loads(gdem_ctx* c) {
  double4 h[kMaxWalls * 2];
  HIP_TRY(hipMemcpyAsync(h, c->d_wall_load, sizeof(double4) * 2 * (size_t)std::max(c->nwalls, 1),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->h_loads.assign(3 * (size_t)std::max(c->nwalls, 1), 0.);
  for (int w = 0; w < c->nwalls; ++w) {
    c->h_loads[3 * w] = h[2 * w].x;
    c->h_loads[3 * w + 1] = h[2 * w].y;
    c->h_loads[3 * w + 2] = h[2 * w].z;
  }
  c->kin.loads = c->h_loads.data();
  return GDEM_OK;
}

static int moving_wall_steps(gdem_ctx* c, int32_t nsteps, double dt) {
  const int nw = std::max(c->nwalls, 1);
  // an imposed force needs the loads of the previous evaluation: one step
  // at a time
  const bool controlled = c->kin.any_force_live();
  const int32_t chunk =
      controlled ? 1 : (int32_t)std::max<size_t>(1, (8u << 20) / (sizeof(Wall) * (size_t)nw));
  for (int32_t k0 = 0; k0 < nsteps; k0 += chunk) {
    const int32_t m = std::min(chunk, nsteps - k0);
    const size_t need = (size_t)m * (size_t)nw;
    // the previous chunk's kernels read d_traj and its copy reads h_traj
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->traj_cap < need) {
      (void)hipFree(c->d_traj);
      c->d_traj = nullptr;
      c->traj_cap = 0;
      HIP_TRY(hipMalloc(&c->d_traj, sizeof(Wall) * need));
      c->traj_cap = need;
    }
    c->h_traj.resize(need);
    if (controlled) {
      const int rl = fetch_loads(c);
      if (rl) return rl;
    }
    double t = c->time;
    for (int32_t k = 0; k < m; ++k) {
      t += dt;
      c->kin.move(c->h_walls, c->nwalls, t, dt);
      std::copy(c->h_walls, c->h_walls + c->nwalls, c->h_traj.begin() + (size_t)k * nw);
    }
    if (c->nwalls > 0)
      HIP_TRY(hipMemcpyAsync(c->d_traj, c->h_traj.data(), sizeof(Wall) * need,
                             hipMemcpyHostToDevice, c->stream));
    for (int32_t k = 0; k < m; ++k) {
      c->time += dt;
      c->step_walls = c->nwalls > 0 ? c->d_traj + (size_t)k * nw : nullptr;
      const bool last = (k0 + k == nsteps - 1);
      const int rc = do_force(c, last ? MODE_KICK : MODE_MOVE, 0.5 * dt, dt, last);
      c->step_walls = nullptr;
      if (rc) return rc;
    }
  }
  return upload_walls(c);
}

extern "C" int gdem_step(gdem_ctx* c, int32_t nsteps, double dt) {
  if (!c || nsteps < 0 || !(dt > 0.)) return GDEM_ERR_ARG;
  if (nsteps == 0) return GDEM_OK;
  HIP_TRY(hipSetDevice(c->dev));
  int rc = gdem_set_time_step(c, dt);
  if (rc) return rc;
  if (!c->primed) {
    rc = gdem_prime(c);
    if (rc) return rc;
  }
  // step 1's kick + drift; afterwards each force kernel closes its step and
  // opens the next one
  if (pack_eligible(c) && !c->kin.any_live()) {
    rc = ensure_packed(c);
    if (rc) return rc;
    note_move_step(c, dt);
    launch_move_pack(c->pk[c->cur], c->pkc[c->cur], c->d_F, c->d_T, (int32_t)c->n, c->grid,
                     0.5 * dt, dt, c->d_err, SkinGate{c->d_skin, c->cur, skin_limit(c)},
                     c->d_path, c->d_wsrc, c->stream);
    HIP_TRY(hipGetLastError());
  } else {
    rc = gdem_move(c, 0.5 * dt, dt);
    if (rc) return rc;
  }
  if (c->kin.any_live()) return moving_wall_steps(c, nsteps, dt);
  const bool graph = c->use_graph && !c->timing && c->rec_cap == 0;
  for (int32_t k = 0; k < nsteps;) {
    // the last step closes with the half kick (MODE_KICK), outside the graphs
    const int32_t open = nsteps - 1 - k;
    const int run = graph && open >= kGraphSteps ? kGraphSteps : 1;
    for (int s = 0; s < run; ++s) c->time += dt;
    const bool last = (k == nsteps - 1);
    if (!last && graph) rc = graph_force(c, dt, run);
    else rc = do_force(c, last ? MODE_KICK : MODE_MOVE, 0.5 * dt, dt, last);
    if (rc) return rc;
    k += run;
  }
  return adapt_skin(c);
}

// Largest list length of any particle at the last list build (synchronises).
static int list_peak(gdem_ctx* c, int32_t* peak) {
  *peak = 0;
  if (!c->d_out_count || !c->d_in_count || c->n <= 0) return GDEM_OK;
  if (!c->d_peak) HIP_TRY(hipMalloc(&c->d_peak, sizeof(int32_t)));
  HIP_TRY(hipMemsetAsync(c->d_peak, 0, sizeof(int32_t), c->stream));
  launch_list_peak(c->d_out_count, c->d_in_count, (int32_t)c->n, c->d_peak, c->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(peak, c->d_peak, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GDEM_OK;
}

// Degrade before the skin lists overflow: when the last list build left a
// particle within kListHeadroom entries of the capacity, halve the skin (the
// lists then hold fewer candidates; below 5 % of the largest radius the skin
// is switched off and lists hold contacts only) and rebuild at the next
// evaluation.  The physics does not depend on the skin.
static int adapt_skin(gdem_ctx* c) {
  if (!c->skin_adapt || !(c->skin > 0.) || c->convex) return GDEM_OK;
  int32_t peak = 0;
  int rc = list_peak(c, &peak);
  if (rc) return rc;
  if (peak <= kListCap - kListHeadroom) return GDEM_OK;
  // a list that overflowed has raised GDEM_ERR_CAPACITY: report it
  rc = collect_errors(c);
  if (rc) return rc;
  c->skin_halvings++;
  return gdem_set_skin(c, 0.5 * c->skin);
}

extern "C" int gdem_set_skin_adapt(gdem_ctx* c, int32_t on) {
  if (!c) return GDEM_ERR_ARG;
  c->skin_adapt = on != 0;
  return GDEM_OK;
}

extern "C" int gdem_get_list_occupancy(gdem_ctx* c, int32_t* peak, int32_t* capacity,
                                       int32_t* skin_halvings) {
  if (!c || !peak) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  const int rc = list_peak(c, peak);
  if (rc) return rc;
  if (capacity) *capacity = c->convex ? kCandCap : kListCap;
  if (skin_halvings) *skin_halvings = c->skin_halvings;
  return GDEM_OK;
}

extern "C" int gdem_set_graph(gdem_ctx* c, int32_t on) {
  if (!c) return GDEM_ERR_ARG;
  c->use_graph = on != 0;
  return GDEM_OK;
}

static int collect_errors(gdem_ctx* c) {
  unsigned long long h[kErrWords];
  HIP_TRY(hipMemcpy(h, c->d_err, sizeof(h), hipMemcpyDeviceToHost));
  int code = GDEM_OK;
  if (h[kErrMotion] != kNoErr) code = set_err(c, GDEM_ERR_MOTION, (int32_t)h[kErrMotion], -1);
  else if (h[kErrSim] != kNoErr) code = set_err(c, GDEM_ERR_SIMULATION, (int32_t)h[kErrSim], -1);
  else if (h[kErrContact] != kNoErr)
    code = set_err(c, GDEM_ERR_CONTACT, (int32_t)(h[kErrContact] >> 32),
                   (int32_t)(uint32_t)(h[kErrContact] & 0xffffffffull));
  else if (h[kErrCapacity] != kNoErr)
    code = set_err(c, GDEM_ERR_CAPACITY, (int32_t)h[kErrCapacity], -1);
  if (code != GDEM_OK) {
    HIP_TRY(hipMemset(c->d_err, 0xff, sizeof(h)));
  }
  return code;
}

extern "C" int gdem_synchronize(gdem_ctx* c) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return collect_errors(c);
}

extern "C" int gdem_get_state(gdem_ctx* c, double* x, double* v, double* w, double* q, double* F,
                              double* T) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  {
    const int rl = ensure_canon(c);
    if (rl) return rl;
  }
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->n;
  const size_t need = n * 19;
  if (c->stage_len < need) {
    (void)hipFree(c->d_stage);
    c->d_stage = nullptr;
    HIP_TRY(hipMalloc(&c->d_stage, sizeof(double) * std::max<size_t>(need, 1)));
    c->stage_len = need;
  }
  double* dx = c->d_stage;
  double* dv = dx + 3 * n;
  double* dw = dv + 3 * n;
  double* dq = dw + 3 * n;
  double* dF = dq + 4 * n;
  double* dT = dF + 3 * n;
  launch_unpermute(c->buf[c->cur], c->d_F, c->d_T, (int32_t)n, x ? dx : nullptr,
                   v ? dv : nullptr, w ? dw : nullptr, q ? dq : nullptr, F ? dF : nullptr,
                   T ? dT : nullptr, c->stream);
  HIP_TRY(hipGetLastError());
  if (n > 0) {
    if (x) HIP_TRY(hipMemcpyAsync(x, dx, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    if (v) HIP_TRY(hipMemcpyAsync(v, dv, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    if (w) HIP_TRY(hipMemcpyAsync(w, dw, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    if (q) HIP_TRY(hipMemcpyAsync(q, dq, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, c->stream));
    if (F) HIP_TRY(hipMemcpyAsync(F, dF, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    if (T) HIP_TRY(hipMemcpyAsync(T, dT, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return collect_errors(c);
}

static int ensure_stage(gdem_ctx* c, size_t need) {
  if (c->stage_len < need) {
    (void)hipFree(c->d_stage);
    c->d_stage = nullptr;
    HIP_TRY(hipMalloc(&c->d_stage, sizeof(double) * std::max<size_t>(need, 1)));
    c->stage_len = need;
  }
  return GDEM_OK;
}

extern "C" int gdem_set_state(gdem_ctx* c, const double* x, const double* v, const double* w,
                              const double* q, const double* F, const double* T) {
  if (!c || c->convex || (!F) != (!T)) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  {
    const int rl = ensure_canon(c);
    if (rl) return rl;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->n;
  if (ensure_stage(c, n * 19)) return GDEM_ERR_HIP;
  double* dx = c->d_stage;
  double* dv = dx + 3 * n;
  double* dw = dv + 3 * n;
  double* dq = dw + 3 * n;
  double* dF = dq + 4 * n;
  double* dT = dF + 3 * n;
  if (n > 0) {
    if (x) HIP_TRY(hipMemcpy(dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    if (v) HIP_TRY(hipMemcpy(dv, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    if (w) HIP_TRY(hipMemcpy(dw, w, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    if (q) HIP_TRY(hipMemcpy(dq, q, sizeof(double) * 4 * n, hipMemcpyHostToDevice));
    if (F) HIP_TRY(hipMemcpy(dF, F, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    if (T) HIP_TRY(hipMemcpy(dT, T, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  }
  launch_set_state(c->buf[c->cur], c->d_F, c->d_T, (int32_t)n, x ? dx : nullptr,
                   v ? dv : nullptr, w ? dw : nullptr, q ? dq : nullptr, F ? dF : nullptr,
                   T ? dT : nullptr, c->stream);
  HIP_TRY(hipGetLastError());
  if (x) {
    launch_keys(c->buf[c->cur], (int32_t)n, c->grid, c->d_row_rank, c->bins.key, c->d_err,
                c->stream);
    HIP_TRY(hipGetLastError());
    c->binned = false;
    const int rw = reset_skin_words(c);  // the lists are rebuilt
    if (rw) return rw;
  }
  // given F and T are the leap-frog's a0: the next step starts from them
  if (F) c->primed = true;
  else if (x || v || w || q) c->primed = false;
  c->prime_evals = -1;  // given F, T: no priming evaluation to re-run for loads
  HIP_TRY(hipStreamSynchronize(c->stream));
  return collect_errors(c);
}

// Downloads the ids of the storage order of the current copy.
static int slot_ids(gdem_ctx* c, std::vector<int32_t>& id) {
  const int rl = ensure_canon(c);
  if (rl) return rl;
  HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->n;
  std::vector<Aux> ax(n);
  if (n) HIP_TRY(hipMemcpy(ax.data(), c->buf[c->cur].aux, sizeof(Aux) * n, hipMemcpyDeviceToHost));
  id.resize(n);
  for (size_t s = 0; s < n; ++s) id[s] = ax[s].id;
  return GDEM_OK;
}

extern "C" int gdem_get_memory(gdem_ctx* c, int64_t capacity, int64_t* n_out, int32_t* owner,
                               int32_t* key, double* kd, double* nrm, double* spr) {
  if (!c || capacity < 0 || !n_out) return GDEM_ERR_ARG;
  *n_out = 0;
  if (!c->has_memory || !c->hist[0].count) return GDEM_OK;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->n, ne = n * kHistCap;
  std::vector<int32_t> id, cnt(n), hk(ne);
  std::vector<double4> hkd(ne), hnr(ne), hsp(ne);
  if (slot_ids(c, id)) return GDEM_ERR_HIP;
  const History& h = c->hist[c->cur];
  if (n) {
    HIP_TRY(hipMemcpy(cnt.data(), h.count, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hk.data(), h.key, sizeof(int32_t) * ne, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hkd.data(), h.kd, sizeof(double4) * ne, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hnr.data(), h.nrm, sizeof(double4) * ne, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hsp.data(), h.spr, sizeof(double4) * ne, hipMemcpyDeviceToHost));
  }
  // entries by (owner id, key)
  std::vector<std::pair<std::pair<int32_t, int32_t>, size_t>> e;
  for (size_t s = 0; s < n; ++s)
    for (int k = 0; k < std::min(cnt[s], kHistCap); ++k) {
      const size_t at = (size_t)k * n + s;
      e.push_back({{id[s], hk[at]}, at});
    }
  std::sort(e.begin(), e.end());
  *n_out = (int64_t)e.size();
  for (size_t j = 0; j < e.size() && (int64_t)j < capacity; ++j) {
    const size_t at = e[j].second;
    if (owner) owner[j] = e[j].first.first;
    if (key) key[j] = e[j].first.second;
    const double4 a = hkd[at], b = hnr[at], d = hsp[at];
    if (kd) { kd[3 * j] = a.x; kd[3 * j + 1] = a.y; kd[3 * j + 2] = a.z; }
    if (nrm) { nrm[3 * j] = b.x; nrm[3 * j + 1] = b.y; nrm[3 * j + 2] = b.z; }
    if (spr) { spr[3 * j] = d.x; spr[3 * j + 1] = d.y; spr[3 * j + 2] = d.z; }
  }
  return GDEM_OK;
}

extern "C" int gdem_set_memory(gdem_ctx* c, int64_t count, const int32_t* owner,
                               const int32_t* key, const double* kd, const double* nrm,
                               const double* spr) {
  if (!c || count < 0 || (count > 0 && (!owner || !key || !kd || !nrm || !spr)))
    return GDEM_ERR_ARG;
  if (!c->has_memory || !c->hist[0].count) return count == 0 ? GDEM_OK : GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->n, ne = n * kHistCap;
  std::vector<int32_t> id;
  if (slot_ids(c, id)) return GDEM_ERR_HIP;
  std::vector<int32_t> slot(n, -1), cnt(n, 0), hk(ne, 0);
  std::vector<double4> hkd(ne), hnr(ne), hsp(ne);
  for (size_t s = 0; s < n; ++s)
    if (id[s] >= 0 && (size_t)id[s] < n) slot[id[s]] = (int32_t)s;
  for (int64_t j = 0; j < count; ++j) {
    if (owner[j] < 0 || (size_t)owner[j] >= n || slot[owner[j]] < 0) return GDEM_ERR_ARG;
    if (key[j] >= 0 && (size_t)key[j] >= n) return GDEM_ERR_ARG;
    if (key[j] < 0 && -1 - key[j] >= c->nwalls) return GDEM_ERR_ARG;
    const size_t s = (size_t)slot[owner[j]];
    if (cnt[s] >= kHistCap) return GDEM_ERR_CAPACITY;
    const size_t at = (size_t)cnt[s]++ * n + s;
    hk[at] = key[j];
    hkd[at] = make_double4(kd[3 * j], kd[3 * j + 1], kd[3 * j + 2], 0.);
    hnr[at] = make_double4(nrm[3 * j], nrm[3 * j + 1], nrm[3 * j + 2], 0.);
    hsp[at] = make_double4(spr[3 * j], spr[3 * j + 1], spr[3 * j + 2], 0.);
  }
  const History& h = c->hist[c->cur];
  if (n) {
    HIP_TRY(hipMemcpy(h.count, cnt.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.key, hk.data(), sizeof(int32_t) * ne, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.kd, hkd.data(), sizeof(double4) * ne, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.nrm, hnr.data(), sizeof(double4) * ne, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h.spr, hsp.data(), sizeof(double4) * ne, hipMemcpyHostToDevice));
  }
  return GDEM_OK;
}

extern "C" int gdem_info(gdem_ctx* c, int64_t* n, int32_t* nwalls, int32_t* convex,
                         int32_t* memory_models) {
  if (!c) return GDEM_ERR_ARG;
  if (n) *n = c->n;
  if (nwalls) *nwalls = c->nwalls;
  if (convex) *convex = c->convex ? 1 : 0;
  if (memory_models) *memory_models = c->has_memory ? 1 : 0;
  return GDEM_OK;
}

// target: wall k, or composite c as -1 - c
static int add_motion(gdem_ctx* c, int32_t target, int32_t type, double t_start, double t_end,
                      const double* p) {
  if (!p || c->convex || !(t_end >= t_start)) return GDEM_ERR_ARG;
  WallMotion m;
  if (!make_wall_motion(target, type, t_start, t_end, p, m)) return GDEM_ERR_ARG;
  if ((type == GDEM_MOTION_SIN_TRANSLATION || type == GDEM_MOTION_CRENEL_TRANSLATION) &&
      !(p[4] > 0.))
    return GDEM_ERR_ARG;
  if (type == GDEM_MOTION_MULTISIN_TRANSLATION && !(p[6] > 0. && p[7] > 0. && p[8] > 0.))
    return GDEM_ERR_ARG;
  if (type >= GDEM_MOTION_SIN_TRANSLATION && type <= GDEM_MOTION_MULTISIN_TRANSLATION &&
      !(norm(mk(p[0], p[1], p[2])) > 0.))
    return GDEM_ERR_ARG;
  c->kin.insert_motion(m);
  return GDEM_OK;
}

extern "C" int gdem_add_wall_motion(gdem_ctx* c, int32_t wall, int32_t type, double t_start,
                                    double t_end, const double* p) {
  if (!c || wall < 0 || wall >= c->nwalls) return GDEM_ERR_ARG;
  return add_motion(c, wall, type, t_start, t_end, p);
}

static int add_force(gdem_ctx* c, int32_t target, double t_start, double t_end,
                     const double* p) {
  if (!(t_end >= t_start)) return GDEM_ERR_ARG;
  WallForce f;
  f.wall = target;
  f.t0 = t_start;
  f.t1 = t_end;
  f.amp = mk(p[0], p[1], p[2]);
  const double na = norm(f.amp);
  if (!(na > 0.) || !(p[4] >= 0.)) return GDEM_ERR_ARG;
  f.dir = dvd(f.amp, na);  // ObstacleImposedForce.cpp:74
  f.automass = !(p[3] > 0.);
  f.mass = f.automass ? 0. : p[3];
  f.vmax = p[4];
  if (!c->track_loads) {
    const int rc = gdem_set_wall_loads(c, 1);
    if (rc) return rc;
  }
  c->kin.fit(c->nwalls);
  c->kin.insert_force(f);
  return GDEM_OK;
}

extern "C" int gdem_add_wall_force(gdem_ctx* c, int32_t wall, double t_start, double t_end,
                                   const double* p) {
  if (!c || !p || c->convex || wall < 0 || wall >= c->nwalls) return GDEM_ERR_ARG;
  return add_force(c, wall, t_start, t_end, p);
}

extern "C" int gdem_add_composite_force(gdem_ctx* c, int32_t composite, double t_start,
                                        double t_end, const double* p) {
  if (!c || !p || c->convex || composite < 0 || composite >= (int32_t)c->kin.comps.size())
    return GDEM_ERR_ARG;
  // the automatic mass divides by the composite's crust, which is zero
  if (!(p[3] > 0.)) return GDEM_ERR_ARG;
  return add_force(c, -1 - composite, t_start, t_end, p);
}

extern "C" int gdem_add_composite_motion(gdem_ctx* c, int32_t composite, int32_t type,
                                         double t_start, double t_end, const double* p) {
  if (!c || composite < 0 || composite >= (int32_t)c->kin.comps.size()) return GDEM_ERR_ARG;
  return add_motion(c, -1 - composite, type, t_start, t_end, p);
}

// centres of all composites from their members (each top-level composite
// recomputes its subtree)
static void composite_centres(gdem_ctx* c) {
  for (int32_t k = 0; k < (int32_t)c->kin.comps.size(); ++k)
    if (c->kin.comps[k].parent < 0) c->kin.volume_centre(k, c->h_walls);
}

extern "C" int gdem_add_composite(gdem_ctx* c, int32_t parent, int32_t* id) {
  if (!c || !id || c->convex || parent < -1 || parent >= (int32_t)c->kin.comps.size() ||
      c->kin.comps.size() >= (size_t)kMaxWalls)
    return GDEM_ERR_ARG;
  Composite cp;
  cp.parent = parent;
  const int32_t k = (int32_t)c->kin.comps.size();
  c->kin.comps.push_back(cp);
  if (parent >= 0) c->kin.comps[parent].members.push_back(-1 - k);
  *id = k;
  composite_centres(c);
  return GDEM_OK;
}

extern "C" int gdem_attach_wall(gdem_ctx* c, int32_t wall, int32_t composite) {
  if (!c || wall < 0 || wall >= c->nwalls || composite < 0 ||
      composite >= (int32_t)c->kin.comps.size())
    return GDEM_ERR_ARG;
  c->kin.fit(c->nwalls);
  if (c->kin.wall_parent[wall] != -1) return GDEM_ERR_ARG;  // already a member
  c->kin.wall_parent[wall] = composite;
  c->kin.comps[composite].members.push_back(wall);
  composite_centres(c);
  return GDEM_OK;
}

extern "C" int gdem_add_cylindrical_shell(gdem_ctx* c, const double* basis, const double* centre,
                                          double height, double radius, double width, int32_t n,
                                          double crust, int32_t material, int32_t parent,
                                          int32_t* composite) {
  if (!c || !basis || !centre || !composite || n < 3 || !(height > 0.) || !(radius > 0.) ||
      !(width > 0.) || c->nwalls + n > kMaxWalls)
    return GDEM_ERR_ARG;
  std::vector<double> wb(9 * (size_t)n), wc(3 * (size_t)n), wh(3 * (size_t)n);
  cylindrical_shell_boxes(basis, centre, height, radius, width, n, wb.data(), wc.data(), wh.data());
  int32_t id = -1;
  int rc = gdem_add_composite(c, parent, &id);
  if (rc) return rc;
  for (int32_t i = 0; i < n; ++i) {
    rc = gdem_add_wall(c, wb.data() + 9 * i, wc.data() + 3 * i, wh.data() + 3 * i, crust, material);
    if (rc) return rc;
    rc = gdem_attach_wall(c, c->nwalls - 1, id);
    if (rc) return rc;
  }
  *composite = id;
  return GDEM_OK;
}

extern "C" int gdem_add_cylindrical_shell_rectangles(gdem_ctx* c, const double* basis,
                                                     const double* centre, double height,
                                                     double radius, double width, int32_t n,
                                                     double crust, int32_t material,
                                                     int32_t parent, int32_t* composite) {
  if (!c || !basis || !centre || !composite || n < 3 || !(height > 0.) || !(radius > 0.) ||
      !(width > 0.) || c->nwalls + n > kMaxWalls)
    return GDEM_ERR_ARG;
  std::vector<double> wb(9 * (size_t)n), wc(3 * (size_t)n), wh(3 * (size_t)n);
  cylindrical_shell_rectangles(basis, centre, height, radius, width, n, wb.data(), wc.data(),
                               wh.data());
  int32_t id = -1;
  int rc = gdem_add_composite(c, parent, &id);
  if (rc) return rc;
  for (int32_t i = 0; i < n; ++i) {
    rc = add_wall_kind(c, wb.data() + 9 * i, wc.data() + 3 * i, wh.data() + 3 * i, crust,
                       material, WALL_RECTANGLE);
    if (rc) return rc;
    rc = gdem_attach_wall(c, c->nwalls - 1, id);
    if (rc) return rc;
  }
  *composite = id;
  return GDEM_OK;
}

extern "C" int gdem_cylindrical_shell_rectangles(const double* basis, const double* centre,
                                                 double height, double radius, double width,
                                                 int32_t n, double* wall_basis,
                                                 double* wall_centre, double* wall_half) {
  if (!basis || !centre || !wall_basis || !wall_centre || !wall_half || n < 3) return GDEM_ERR_ARG;
  cylindrical_shell_rectangles(basis, centre, height, radius, width, n, wall_basis, wall_centre,
                               wall_half);
  return GDEM_OK;
}

extern "C" int gdem_cylindrical_shell_boxes(const double* basis, const double* centre,
                                            double height, double radius, double width, int32_t n,
                                            double* wall_basis, double* wall_centre,
                                            double* wall_half) {
  if (!basis || !centre || !wall_basis || !wall_centre || !wall_half || n < 3) return GDEM_ERR_ARG;
  cylindrical_shell_boxes(basis, centre, height, radius, width, n, wall_basis, wall_centre,
                          wall_half);
  return GDEM_OK;
}

extern "C" int gdem_set_motion_restriction(gdem_ctx* c, int32_t wall, int32_t composite,
                                           int32_t mask) {
  if (!c || mask < 0 || mask > 7) return GDEM_ERR_ARG;
  if (wall >= 0) {
    if (wall >= c->nwalls || composite >= 0) return GDEM_ERR_ARG;
    c->kin.fit(c->nwalls);
    c->kin.wall_restrict[wall] = (uint32_t)mask;
    return GDEM_OK;
  }
  if (composite < 0 || composite >= (int32_t)c->kin.comps.size()) return GDEM_ERR_ARG;
  c->kin.comps[composite].restrict_mask = (uint32_t)mask;
  return GDEM_OK;
}

extern "C" int gdem_get_composite(gdem_ctx* c, int32_t composite, double* centre, double* vel,
                                  double* omega) {
  if (!c || composite < 0 || composite >= (int32_t)c->kin.comps.size()) return GDEM_ERR_ARG;
  const Composite& cp = c->kin.comps[composite];
  const double ce[3] = {cp.centre.x, cp.centre.y, cp.centre.z};
  const double ve[3] = {cp.vel.x, cp.vel.y, cp.vel.z};
  const double om[3] = {cp.omega.x, cp.omega.y, cp.omega.z};
  if (centre) memcpy(centre, ce, sizeof(ce));
  if (vel) memcpy(vel, ve, sizeof(ve));
  if (omega) memcpy(omega, om, sizeof(om));
  return GDEM_OK;
}

extern "C" int gdem_set_wall_loads(gdem_ctx* c, int32_t on) {
  if (!c || c->convex) return GDEM_ERR_ARG;
  if (!on && !c->kin.forces.empty()) return GDEM_ERR_ARG;  // the controllers read the loads
  if ((on != 0) == c->track_loads) return GDEM_OK;  // already in that state: loads kept
  // Loads switched on after the priming evaluation: that evaluation's loads
  // (the ones the reference holds at this point) are formed again by
  // re-running it on the unchanged state.  This is exact only before the
  // first step and without contact memory (an evaluation advances the
  // tangential springs); otherwise the caller must switch loads on first.
  const bool redo = on && c->primed;
  if (redo && (c->has_memory || c->force_evals != c->prime_evals)) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  {
    const int rl = ensure_canon(c);
    if (rl) return rl;
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->track_loads = on != 0;
  const int rc = ensure_wall_loads(c);
  if (rc || !redo) return rc;
  const int rf = do_force(c, MODE_FORCE, 0., 0., true);
  if (rf) return rf;
  c->prime_evals = c->force_evals;
  return GDEM_OK;
}

extern "C" int gdem_get_wall_loads(gdem_ctx* c, double* force, double* torque) {
  if (!c || !c->track_loads || !c->d_wall_load) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  double4 h[kMaxWalls * 2];
  HIP_TRY(hipMemcpyAsync(h, c->d_wall_load, sizeof(double4) * 2 * (size_t)std::max(c->nwalls, 1),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int w = 0; w < c->nwalls; ++w) {
    if (force) { force[3 * w] = h[2 * w].x; force[3 * w + 1] = h[2 * w].y; force[3 * w + 2] = h[2 * w].z; }
    if (torque) {
      torque[3 * w] = h[2 * w + 1].x; torque[3 * w + 1] = h[2 * w + 1].y; torque[3 * w + 2] = h[2 * w + 1].z;
    }
  }
  return collect_errors(c);
}

extern "C" int gdem_set_time(gdem_ctx* c, double time) {
  if (!c) return GDEM_ERR_ARG;
  c->time = time;
  if (c->stats_on) {
    // the statistics take the time of each evaluation from the device copy
    HIP_TRY(hipSetDevice(c->dev));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(&c->d_stats->time, &time, sizeof(double), hipMemcpyHostToDevice));
  }
  return GDEM_OK;
}

extern "C" int gdem_get_time(gdem_ctx* c, double* time) {
  if (!c || !time) return GDEM_ERR_ARG;
  *time = c->time;
  return GDEM_OK;
}

extern "C" int gdem_move_walls(gdem_ctx* c, double time, double dt) {
  if (!c || !(dt > 0.)) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  c->time = time;
  if (c->kin.any_force_live()) {
    const int rl = fetch_loads(c);
    if (rl) return rl;
  }
  if (c->kin.move(c->h_walls, c->nwalls, time, dt)) return upload_walls(c);
  return GDEM_OK;
}

extern "C" int gdem_get_wall(gdem_ctx* c, int32_t wall, double* centre, double* basis,
                             double* vel, double* omega) {
  if (!c || wall < 0 || wall >= c->nwalls) return GDEM_ERR_ARG;
  const Wall& wl = c->h_walls[wall];
  if (centre) memcpy(centre, wl.origin, sizeof(wl.origin));
  if (basis) memcpy(basis, wl.basis, sizeof(wl.basis));
  if (vel) memcpy(vel, wl.vel, sizeof(wl.vel));
  if (omega) memcpy(omega, wl.omega, sizeof(wl.omega));
  return GDEM_OK;
}

extern "C" int gdem_set_contact_recording(gdem_ctx* c, int32_t cap) {
  if (!c || cap < 0 || cap > 256) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(c->d_rec);
  (void)hipFree(c->d_rec_count);
  c->d_rec = nullptr;
  c->d_rec_count = nullptr;
  c->rec_cap = cap;
  c->rec_valid = false;
  c->graph_gen++;
  if (cap > 0 && c->n > 0) {
    HIP_TRY(hipMalloc(&c->d_rec, sizeof(ContactRec) * (size_t)c->n * (size_t)cap));
    HIP_TRY(hipMalloc(&c->d_rec_count, sizeof(int32_t) * (size_t)c->n));
  }
  return GDEM_OK;
}

extern "C" int gdem_get_contacts(gdem_ctx* c, int64_t capacity, int64_t* n_out, int32_t* id0,
                                 int32_t* id1, double* point, double* ovec, double* overlap,
                                 double* force) {
  if (!c || !n_out) return GDEM_ERR_ARG;
  *n_out = 0;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int rc = collect_errors(c);
  if (rc) return rc;
  if (!c->rec_valid || c->n == 0) return GDEM_OK;
  const size_t n = (size_t)c->n, cap = (size_t)c->rec_cap;
  std::vector<int32_t> cnt(n);
  std::vector<ContactRec> rec(n * cap);
  HIP_TRY(hipMemcpy(cnt.data(), c->d_rec_count, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(rec.data(), c->d_rec, sizeof(ContactRec) * n * cap, hipMemcpyDeviceToHost));
  std::vector<ContactRec> all;
  for (size_t t = 0; t < n; ++t)
    for (int32_t k = 0; k < cnt[t] && (size_t)k < cap; ++k) all.push_back(rec[t * cap + k]);
  std::sort(all.begin(), all.end(), [](const ContactRec& a, const ContactRec& b) {
    if (a.id0 != b.id0) return a.id0 < b.id0;
    // walls (negative ids) after particles, in wall order
    const int64_t ka = a.id1 >= 0 ? a.id1 : (int64_t)0x7fffffff - a.id1;
    const int64_t kb = b.id1 >= 0 ? b.id1 : (int64_t)0x7fffffff - b.id1;
    return ka < kb;
  });
  *n_out = (int64_t)all.size();
  const size_t m = std::min<size_t>(all.size(), (size_t)std::max<int64_t>(capacity, 0));
  for (size_t k = 0; k < m; ++k) {
    const ContactRec& r = all[k];
    if (id0) id0[k] = r.id0;
    if (id1) id1[k] = r.id1;
    if (point) memcpy(point + 3 * k, r.point, sizeof(r.point));
    if (ovec) memcpy(ovec + 3 * k, r.ovec, sizeof(r.ovec));
    if (overlap) overlap[k] = r.dist;
    if (force) memcpy(force + 3 * k, r.force, sizeof(r.force));
  }
  return (int64_t)all.size() > capacity ? GDEM_ERR_CAPACITY : GDEM_OK;
}

extern "C" int gdem_last_error(gdem_ctx* c, int32_t* code, int32_t* id0, int32_t* id1) {
  if (!c) return GDEM_ERR_ARG;
  if (code) *code = c->err_code;
  if (id0) *id0 = c->err_id0;
  if (id1) *id1 = c->err_id1;
  return c->err_code;
}

extern "C" int gdem_set_kernel_timing(gdem_ctx* c, int32_t on) {
  if (!c) return GDEM_ERR_ARG;
  c->timing = on != 0;
  return GDEM_OK;
}

// Adds the elapsed times of the recorded launches to the per-kind totals.
static int fold_timings(gdem_ctx* c) {
  for (size_t k = 0; k < c->ev_used; ++k) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_pool[k].start, c->ev_pool[k].stop));
    c->kernel_ms[c->ev_pool[k].kind] += ms;
    c->timed[c->ev_pool[k].kind]++;
  }
  c->ev_used = 0;
  return GDEM_OK;
}

extern "C" int gdem_get_counters(gdem_ctx* c, int64_t* force_evals, int64_t* contacts,
                                 double* force_kernel_ms, int64_t* timed_launches) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int rt = fold_timings(c);
  if (rt) return rt;
  unsigned long long h[2];
  HIP_TRY(hipMemcpy(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost));
  if (c->d_blk_contacts) {
    std::vector<unsigned long long> blk((size_t)c->nblk_slots);
    HIP_TRY(hipMemcpy(blk.data(), c->d_blk_contacts, sizeof(unsigned long long) * blk.size(),
                      hipMemcpyDeviceToHost));
    for (unsigned long long b : blk) h[0] += b;
  }
  if (force_evals) *force_evals = c->force_evals;
  if (contacts) *contacts = (int64_t)h[0];
  if (force_kernel_ms) *force_kernel_ms = c->kernel_ms[0];
  if (timed_launches) *timed_launches = c->timed[0];
  return GDEM_OK;
}

extern "C" int gdem_get_kernel_times(gdem_ctx* c, double* ms, int64_t* launches) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int rt = fold_timings(c);
  if (rt) return rt;
  for (int k = 0; k < kTimedKinds; ++k) {
    if (ms) ms[k] = c->kernel_ms[k];
    if (launches) launches[k] = c->timed[k];
  }
  return GDEM_OK;
}

extern "C" int gdem_reset_counters(gdem_ctx* c) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemset(c->d_counters, 0, sizeof(unsigned long long) * 2));
  if (c->d_blk_contacts)
    HIP_TRY(hipMemset(c->d_blk_contacts, 0, sizeof(unsigned long long) * c->nblk_slots));
  c->force_evals = 0;
  for (int k = 0; k < kTimedKinds; ++k) {
    c->kernel_ms[k] = 0.;
    c->timed[k] = 0;
  }
  c->ev_used = 0;
  return GDEM_OK;
}

extern "C" int gdem_set_contact_stats(gdem_ctx* c, int32_t on) {
  if (!c) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (on && !c->d_stats) HIP_TRY(hipMalloc(&c->d_stats, sizeof(ContactStats)));
  if (on) {
    ContactStats z{};
    z.time = c->time;
    HIP_TRY(hipMemcpy(c->d_stats, &z, sizeof(z), hipMemcpyHostToDevice));
  }
  c->stats_on = on != 0;
  c->graph_gen++;
  return GDEM_OK;
}

extern "C" int gdem_set_narrow_phase(gdem_ctx* c, int32_t alg) {
  if (!c || (alg != GDEM_GJK && alg != GDEM_GJK_SV)) return GDEM_ERR_ARG;
  c->gjk_sv = alg == GDEM_GJK_SV ? 1 : 0;
  c->graph_gen++;
  return GDEM_OK;
}

extern "C" int gdem_get_narrow_phase(gdem_ctx* c, int32_t* alg) {
  if (!c || !alg) return GDEM_ERR_ARG;
  *alg = c->gjk_sv ? GDEM_GJK_SV : GDEM_GJK;
  return GDEM_OK;
}

extern "C" int gdem_get_contact_stats(gdem_ctx* c, double* out) {
  if (!c || !out || !c->stats_on) return GDEM_ERR_ARG;
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  ContactStats h{};
  HIP_TRY(hipMemcpy(&h, c->d_stats, sizeof(h), hipMemcpyDeviceToHost));
  out[0] = h.omax;
  out[1] = h.on ? h.osum / (double)h.on : 0.;
  out[2] = h.tmax;
  out[3] = h.gn ? (double)h.gsum / (double)h.gn : 0.;
  out[4] = (double)h.on;
  out[5] = (double)h.gn;
  return collect_errors(c);
}

extern "C" int gdem_get_particles(gdem_ctx* c, double* radius, double* density, double* crust,
                                  int32_t* material) {
  if (!c) return GDEM_ERR_ARG;
  if (c->convex || c->h_radius.size() != (size_t)c->n) return GDEM_ERR_ARG;
  const size_t n = (size_t)c->n;
  if (radius) std::copy(c->h_radius.begin(), c->h_radius.begin() + n, radius);
  if (density) std::copy(c->h_density.begin(), c->h_density.begin() + n, density);
  if (crust) std::copy(c->h_crust.begin(), c->h_crust.begin() + n, crust);
  if (material) std::copy(c->h_material.begin(), c->h_material.begin() + n, material);
  return GDEM_OK;
}

extern "C" int gdem_get_contact_model(gdem_ctx* c, int32_t a, int32_t b, int32_t* kind,
                                      double* p) {
  if (!c || a < 0 || b < 0 || a >= kMaxMaterials || b >= kMaxMaterials) return GDEM_ERR_ARG;
  const Model& m = c->h_models[a * kMaxMaterials + b];
  if (kind) *kind = m.kind;
  if (p)
    for (int k = 0; k < 8; ++k) p[k] = k < model_nparams(m.kind) ? m.p[k] : 0.;
  return GDEM_OK;
}

extern "C" int gdem_get_wall_shape(gdem_ctx* c, int32_t k, double* basis, double* centre,
                                   double* half, double* crust, int32_t* material, int32_t* kind) {
  if (!c || k < 0 || k >= c->nwalls) return GDEM_ERR_ARG;
  const Wall& w = c->h_walls[k];
  if (basis) std::copy(w.basis, w.basis + 9, basis);
  if (centre) std::copy(w.origin, w.origin + 3, centre);
  if (half) std::copy(w.ext, w.ext + 3, half);
  if (crust) *crust = w.crust;
  if (material) *material = w.material;
  if (kind) *kind = w.kind == WALL_RECTANGLE ? GDEM_WALL_RECTANGLE : GDEM_WALL_BOX;
  return GDEM_OK;
}

static int copy_name(const std::string& s, char* buf, int32_t cap) {
  if (!buf || cap <= 0) return GDEM_ERR_ARG;
  if ((int64_t)s.size() + 1 > cap) return GDEM_ERR_CAPACITY;
  memcpy(buf, s.c_str(), s.size() + 1);
  return GDEM_OK;
}

extern "C" int gdem_set_material_name(gdem_ctx* c, int32_t k, const char* name) {
  if (!c || !name || k < 0 || k >= kMaxMaterials) return GDEM_ERR_ARG;
  c->mat_name[k] = name;
  return GDEM_OK;
}

extern "C" int gdem_get_material_name(gdem_ctx* c, int32_t k, char* buf, int32_t cap) {
  if (!c || k < 0 || k >= kMaxMaterials) return GDEM_ERR_ARG;
  return copy_name(c->mat_name[k], buf, cap);
}

extern "C" int gdem_set_wall_name(gdem_ctx* c, int32_t k, const char* name) {
  if (!c || !name || k < 0 || k >= c->nwalls) return GDEM_ERR_ARG;
  c->wall_name[k] = name;
  return GDEM_OK;
}

extern "C" int gdem_get_wall_name(gdem_ctx* c, int32_t k, char* buf, int32_t cap) {
  if (!c || k < 0 || k >= c->nwalls) return GDEM_ERR_ARG;
  return copy_name(c->wall_name[k], buf, cap);
}

extern "C" int gdem_get_grid(gdem_ctx* c, int32_t* nc, double* edge) {
  if (!c) return GDEM_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    if (nc) nc[d] = c->grid.nc[d];
    if (edge) edge[d] = c->grid.edge[d];
  }
  return GDEM_OK;
}
