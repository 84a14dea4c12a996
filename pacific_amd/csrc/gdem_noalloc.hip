// gdem_noalloc.hip -- gdem_sphere_forces_noalloc: the force evaluation of a
// sphere system (gdem_prime / gdem_compute_forces on a context holding those
// particles: binning, contact detection, Hooke/Hertz forces and weight) from
// device arrays, with no device allocation.  Every buffer the kernels need
// beyond the caller's inputs and outputs lives in one file-scope device array
// (`scratch`), carved into the same buffers a context allocates.
//
// The array's address is taken once per device with hipGetSymbolAddress; host
// code uses only that address (the array's name is a host-side stand-in).
// Every call on a device shares its array: calls that overlap on one device
// race, and each call clears what must start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gdem.h"
#include "gdem_host.h"
#include "gdem_internal.h"

namespace gdem {
namespace {

constexpr int64_t kNaParticles = GDEM_NOALLOC_MAX_PARTICLES;
constexpr int64_t kNaCells = GDEM_NOALLOC_MAX_CELLS;
constexpr int64_t kNaRows = GDEM_NOALLOC_MAX_ROWS;
constexpr int kNaBlock = 256;

constexpr size_t up256(size_t b) { return (b + 255) / 256 * 256; }

// Byte offsets of the buffers in the scratch array (sized for the maxima).
struct NaLayout {
  size_t state[2][6], F, T, key, slot, perm, pos, cnt, local, bsum, cell_start, row_rank, out_list,
      out_count, in_list, in_count, models, err, contacts, total;
};

constexpr int64_t kNaScanBlocks = (kNaCells + kScanCells - 1) / kScanCells;
constexpr int64_t kNaBlk8 = ((kNaParticles + kNaBlock - 1) / kNaBlock + 7) / 8 * 8;

constexpr NaLayout na_layout() {
  NaLayout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    const size_t at = o;
    o += up256(bytes);
    return at;
  };
  for (int k = 0; k < 2; ++k) {
    for (int a = 0; a < 5; ++a) L.state[k][a] = take(sizeof(double4) * kNaParticles);
    L.state[k][5] = take(sizeof(Aux) * kNaParticles);
  }
  L.F = take(sizeof(double4) * kNaParticles);
  L.T = take(sizeof(double4) * kNaParticles);
  L.key = take(sizeof(int32_t) * kNaParticles);
  L.slot = take(sizeof(int32_t) * kNaParticles);
  L.perm = take(sizeof(int32_t) * kNaParticles);
  L.pos = take(sizeof(float4) * kNaParticles);
  L.cnt = take(sizeof(int32_t) * kNaCells);
  L.local = take(sizeof(int32_t) * kNaCells);
  L.bsum = take(sizeof(int32_t) * (kNaScanBlocks + 1));
  L.cell_start = take(sizeof(int32_t) * (kNaCells + 1));
  L.row_rank = take(sizeof(int32_t) * kNaRows);
  L.out_list = take(sizeof(int2) * kListCap * kNaParticles);
  L.out_count = take(sizeof(int32_t) * kNaParticles);
  L.in_list = take(sizeof(int2) * kListCap * kNaParticles);
  L.in_count = take(sizeof(int32_t) * kNaParticles);
  L.models = take(sizeof(Model));
  L.err = take(sizeof(unsigned long long) * kErrWords);
  L.contacts = take(sizeof(unsigned long long) * kNaBlk8);
  L.total = o;
  return L;
}

constexpr NaLayout kNa = na_layout();

// The scratch array: about 34 MB for 32,768 particles and 2^19 cells.
static __device__ __align__(256) unsigned char scratch[kNa.total];

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
unsigned char* g_na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on the current device (taken on the first
// call there).
int na_scratch(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kNaMaxDevices) return GDEM_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GDEM_ERR_HIP;
    g_na_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_na_ptr[dev];
  return GDEM_OK;
}

// Rank of row (j, k) among the rows of an nj x nk grid in the 2-D Morton
// order of (j, k) (code bit 2b = bit b of k, bit 2b + 1 = bit b of j), as
// row_morton_rank sorts them: the number of rows with a smaller code.  For
// each 1 bit p of the code, the rows that agree with it above p and have 0 at
// p form a product of two aligned ranges of j and k, clipped to the grid.
__host__ __device__ inline int32_t morton_row_rank(uint32_t j, uint32_t k, int64_t nj, int64_t nk) {
  int64_t rank = 0;
  for (int p = 61; p >= 0; --p) {
    const bool jbit = (p & 1) != 0;
    const int b = p >> 1;  // bit of j (odd p) or of k (even p) at code bit p
    if (b > 30) continue;
    const uint32_t bit = jbit ? (j >> b) & 1u : (k >> b) & 1u;
    if (!bit) continue;
    const int fj = p >> 1, fk = (p + 1) >> 1;  // free low bits of j and of k
    int64_t j0 = (int64_t)((j >> fj) << fj), k0 = (int64_t)((k >> fk) << fk);
    if (jbit) j0 &= ~((int64_t)1 << fj);
    else k0 &= ~((int64_t)1 << fk);
    const int64_t cj = (j0 + ((int64_t)1 << fj) < nj ? j0 + ((int64_t)1 << fj) : nj) - j0;
    const int64_t ck = (k0 + ((int64_t)1 << fk) < nk ? k0 + ((int64_t)1 << fk) : nk) - k0;
    if (cj > 0 && ck > 0) rank += cj * ck;
  }
  return (int32_t)rank;
}

// Clears the counts and counters that the binning, detection and force
// kernels expect at zero, sets the error words, writes the model and ranks
// the cell rows.
__global__ void __launch_bounds__(kNaBlock) k_na_prepare(int32_t* cnt, int ncell, int32_t* bsum,
                                                         int nbsum, unsigned long long* contacts,
                                                         int ncontacts, unsigned long long* err,
                                                         Model* model, Model m, int32_t* row_rank,
                                                         int nj, int nk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int c = i; c < ncell; c += stride) cnt[c] = 0;
  for (int c = i; c <= nbsum; c += stride) bsum[c] = 0;
  for (int c = i; c < ncontacts; c += stride) contacts[c] = 0ull;
  if (i < kErrWords) err[i] = kNoErr;
  if (i == 0) *model = m;
  for (int q = i; q < nj * nk; q += stride)
    row_rank[q] = morton_row_rank((uint32_t)(q / nk), (uint32_t)(q % nk), nj, nk);
}

// The state copy the evaluation reads, in particle order: position and
// radius, velocity and mass, angular velocity and equivalent-sphere radius
// as given; identity orientation (the body-frame angular velocity is then the
// world one); id = index, material 0, the given crust.
__global__ void __launch_bounds__(kNaBlock) k_na_load(State s, const double4* pos_r,
                                                      const double4* vel_m,
                                                      const double4* om_req, double crust,
                                                      int n, int32_t* in_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 w = om_req[i];
  s.pos_r[i] = pos_r[i];
  s.vel_m[i] = vel_m[i];
  s.om_req[i] = w;
  s.wbf_I[i] = make_double4(w.x, w.y, w.z, 0.);
  s.quat[i] = make_double4(0., 0., 0., 1.);
  s.aux[i] = Aux{i, 0, crust};
  in_count[i] = 0;
}

template <typename T>
T* at(unsigned char* base, size_t off) {
  return reinterpret_cast<T*>(base + off);
}

State na_state(unsigned char* b, int k) {
  State s;
  s.pos_r = at<double4>(b, kNa.state[k][0]);
  s.vel_m = at<double4>(b, kNa.state[k][1]);
  s.om_req = at<double4>(b, kNa.state[k][2]);
  s.wbf_I = at<double4>(b, kNa.state[k][3]);
  s.quat = at<double4>(b, kNa.state[k][4]);
  s.aux = at<Aux>(b, kNa.state[k][5]);
  return s;
}

}  // namespace
}  // namespace gdem

using namespace gdem;

extern "C" int gdem_sphere_forces_noalloc(const gdem_config* cfg, int32_t kind,
                                          const double* params, int64_t n, const double* pos_r,
                                          const double* vel_m, const double* om_req, double crust,
                                          double* F, double* T, void* stream, int32_t* err_ids) {
  if (!cfg || !params || n < 0 || !(crust > 0.)) return GDEM_ERR_ARG;
  if (n > 0 && (!pos_r || !vel_m || !om_req || !F || !T)) return GDEM_ERR_ARG;
  if (kind != GDEM_MODEL_HOOKE && kind != GDEM_MODEL_HERTZ) return GDEM_ERR_ARG;
  if (!model_params_ok(kind, params)) return GDEM_ERR_ARG;
  Grid g;
  if (!make_grid(*cfg, g)) return GDEM_ERR_ARG;
  // beyond the scratch array's maxima
  if (n > kNaParticles || g.ncell > kNaCells || (int64_t)g.nc[1] * g.nc[2] > kNaRows)
    return GDEM_ERR_CAPACITY;
  if (err_ids) err_ids[0] = err_ids[1] = -1;
  if (n == 0) return GDEM_OK;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return GDEM_ERR_HIP;
  if (hipSetDevice(cfg->device) != hipSuccess) return GDEM_ERR_HIP;
  int rc = GDEM_OK;
  do {
    unsigned char* base = nullptr;
    rc = na_scratch(cfg->device, &base);
    if (rc) break;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int32_t nn = (int32_t)n;
    Bins b{};
    b.key = at<int32_t>(base, kNa.key);
    b.slot = at<int32_t>(base, kNa.slot);
    b.cnt = at<int32_t>(base, kNa.cnt);
    b.local = at<int32_t>(base, kNa.local);
    b.bsum = at<int32_t>(base, kNa.bsum);
    b.perm = at<int32_t>(base, kNa.perm);
    b.cell_start = at<int32_t>(base, kNa.cell_start);
    b.pos = at<float4>(base, kNa.pos);
    b.nbsum = (g.ncell + kScanCells - 1) / kScanCells;
    b.ticket = b.bsum + b.nbsum;
    int32_t* row_rank = at<int32_t>(base, kNa.row_rank);
    int32_t* in_count = at<int32_t>(base, kNa.in_count);
    unsigned long long* err = at<unsigned long long>(base, kNa.err);
    unsigned long long* contacts = at<unsigned long long>(base, kNa.contacts);
    Model* model = at<Model>(base, kNa.models);
    const int nblk = (int)((n + 255) / 256);
    const int nblk8 = ((nblk + 7) / 8) * 8;
    const int nrows = g.nc[1] * g.nc[2];
    const int prep = (int)std::max<int64_t>(std::max<int64_t>(g.ncell, nrows), kNaBlk8);
    hipLaunchKernelGGL(k_na_prepare, dim3(std::min((prep + kNaBlock - 1) / kNaBlock, 1024)),
                       dim3(kNaBlock), 0, st, b.cnt, g.ncell, b.bsum, b.nbsum, contacts,
                       (int)kNaBlk8, err, model, make_model(kind, params), row_rank, g.nc[1],
                       g.nc[2]);
    const State s0 = na_state(base, 0), s1 = na_state(base, 1);
    hipLaunchKernelGGL(k_na_load, dim3((nn + kNaBlock - 1) / kNaBlock), dim3(kNaBlock), 0, st, s0,
                       reinterpret_cast<const double4*>(pos_r),
                       reinterpret_cast<const double4*>(vel_m),
                       reinterpret_cast<const double4*>(om_req), crust, nn, in_count);
    launch_keys(s0, nn, g, row_rank, b.key, err, st);
    launch_binning(b, PosSrc{s0.pos_r, nullptr, s0.aux}, nn, g, row_rank, err, st);
    DetectArgs d{};
    d.pos = b.pos;
    d.margin = prefilter_margin(g, d.sizef);
    d.in = s0;
    d.perm = b.perm;
    d.cell_start = b.cell_start;
    d.row_rank = row_rank;
    d.grid = g;
    d.n = nn;
    d.lstride = nn;
    d.out_list = at<int2>(base, kNa.out_list);
    d.out_count = at<int32_t>(base, kNa.out_count);
    d.in_list = at<int2>(base, kNa.in_list);
    d.in_count = in_count;
    d.err = err;
    d.nblk8 = nblk8;
    launch_detect(d, st);
    ForceArgs a{};
    a.in = s0;
    a.out = s1;
    a.perm = b.perm;
    a.row_rank = row_rank;
    a.grid = g;
    a.models = model;
    a.model_stride = 1;  // every particle has material 0
    a.n = nn;
    for (int k = 0; k < 3; ++k) a.gravity[k] = cfg->gravity[k];
    a.lstride = nn;
    a.out_list = d.out_list;
    a.out_count = d.out_count;
    a.in_list = d.in_list;
    a.in_count = in_count;
    a.F = at<double4>(base, kNa.F);
    a.T = at<double4>(base, kNa.T);
    a.key_out = b.key;
    a.err = err;
    a.contacts = contacts;
    a.nblk8 = nblk8;
    launch_force(MODE_FORCE, false, a, st);
    // forces in particle order (the evaluation wrote them in slot order)
    launch_unpermute(s1, a.F, a.T, nn, nullptr, nullptr, nullptr, nullptr, F, T, st);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
      rc = GDEM_ERR_HIP;
      break;
    }
    unsigned long long h[kErrWords];
    if (hipMemcpy(h, err, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = GDEM_ERR_HIP;
      break;
    }
    int32_t id0 = -1, id1 = -1;
    if (h[kErrSim] != kNoErr) {
      rc = GDEM_ERR_SIMULATION;
      id0 = (int32_t)h[kErrSim];
    } else if (h[kErrContact] != kNoErr) {
      rc = GDEM_ERR_CONTACT;
      id0 = (int32_t)(h[kErrContact] >> 32);
      id1 = (int32_t)(uint32_t)(h[kErrContact] & 0xffffffffull);
    } else if (h[kErrCapacity] != kNoErr) {
      rc = GDEM_ERR_CAPACITY;
      id0 = (int32_t)h[kErrCapacity];
    }
    if (err_ids) {
      err_ids[0] = id0;
      err_ids[1] = id1;
    }
  } while (false);
  (void)hipSetDevice(prev);
  return rc;
}
