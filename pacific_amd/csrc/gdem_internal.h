// gdem_internal.h -- device buffers, launch parameters and launchers shared by
// the kernels (gdem_kernels.hip) and the C ABI (gdem_capi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gdem_convex.h"
#include "gdem_device.h"

namespace gdem {

// Per-particle record kept next to the state: id, material, crust.
struct Aux {
  int32_t id;
  int32_t mat;
  double crust;
};

// One copy of the particle state, structure of double4 arrays, in the order of
// the last binning (sorted by cell key, ties by previous position).
struct State {
  double4* pos_r;   // x, y, z, radius
  double4* vel_m;   // v, mass
  double4* om_req;  // world angular velocity, equivalent-sphere radius
  double4* wbf_I;   // body-frame angular velocity, principal inertia
  double4* quat;    // orientation (x, y, z, w)
  Aux* aux;
};

// Cell key = row_rank[j * nk + k] * ni + i: cells of one row along x are
// consecutive, rows follow a 2-D Morton order of (j, k).
struct Grid {
  double origin[3], edge[3], size[3];
  int32_t nc[3];
  int32_t periodic[3];
  int32_t ncell;
};

// One recorded contact (first body's view).
struct ContactRec {
  int32_t id0, id1;
  double point[3], ovec[3], dist, force[3];
};

// Maximum contacts one particle can detect (outgoing) or receive (incoming)
// per step; exceeding it raises GDEM_ERR_CAPACITY.
constexpr int kListCap = 24;
// skin lists: entries kept free before the skin is halved
constexpr int kListHeadroom = 4;

// Cells per block of the binning scan (1,024 threads x 4 cells); the
// block totals bsum[] are indexed by cell / kScanCells.
constexpr int kScanPer = 4;
constexpr int kScanCells = 1024 * kScanPer;
static_assert(kListCap < 32, "list entries are tracked in 32-bit masks");

// Device error words, min-reduced with integer atomics so the reported
// offender does not depend on scheduling.
enum { kErrContact = 0, kErrMotion = 1, kErrSim = 2, kErrCapacity = 3, kErrWords = 4 };
constexpr unsigned long long kNoErr = ~0ull;

// Device words of the skin candidate lists.  The lists built from the
// positions of the last rebuild stay exact while every particle has moved
// less than skin/2 since then.  Each particle keeps the length of its path
// since the last rebuild (single precision rounded up, so it bounds the
// displacement); a step that takes it past the limit raises the need word of
// the next evaluation:
//   * the evaluation reading copy p rebuilds when need[p] is set; its first
//     kernel clears need[1-p] and counts the rebuild; its integrator tail
//     (path restarted on a rebuild) raises need[1-p] and one thread writes
//     rebuilt[1-p];
//   * a stand-alone move in place of copy p raises need[p].
// No kernel of an evaluation writes a word it reads, and the need words are
// written only by the few particles that cross the limit (integer OR).
struct SkinWords {
  int32_t need[2];
  int32_t rebuilt[2];  // rebuilt[p]: the evaluation that wrote copy p rebuilt
  int32_t rebuilds;
};

// Rebuild decision of the evaluation reading copy p (the same in every
// kernel of that evaluation); w == nullptr: no skin lists, the lists are
// rebuilt at every evaluation.
struct SkinGate {
  const SkinWords* w;
  int32_t p;
  float limit;  // skin / 2, rounded down
  __device__ __forceinline__ bool rebuild() const { return !w || w->need[p] != 0; }
};

// Path length of a particle after a drift of length depl, restarted when the
// lists were rebuilt from the positions before the drift.  The sum is formed
// in double precision and rounded up to single precision, so the stored
// length never falls below the true one (the double sum's rounding is far
// below a single-precision step).
__device__ __forceinline__ float skin_path(float before, bool restart, double depl) {
  const double len = restart ? depl : (double)before + depl;
  const float f = __double2float_ru(len);
  return restart ? f : nextafterf(f, __builtin_huge_valf());
}

// Full-precision position source of the binning and sweep kernels: xyz (and
// the radius in .w when rad is null), the radius in rad[i].x otherwise.
struct PosSrc {
  const double4* xyz;
  const double2* rad;
  const Aux* aux;
  __device__ __forceinline__ double4 pos(int i) const {
    const double4 a = xyz[i];
    return rad ? make_double4(a.x, a.y, a.z, rad[i].x) : a;
  }
};

// A contact-list entry: partner slot and a code packing the periodic shift to
// apply to the partner ((s+1) in 2 bits per axis), whether this particle is
// the first body of the pair (bit 6) and, in bits 7-11, the position of the
// mirror entry in the partner's other list (an out entry holds its index in
// the partner's in-list, an in entry its index in the partner's out-list).
struct DetectArgs {
  const float4* pos;      // single-precision positions + radius, sorted order
  const int32_t* ids;     // particle ids, sorted order (skin lists)
  float sizef[3];         // box size in single precision (periodic shifts)
  float margin;           // pre-filter margin, bounds the float rounding
  State in;               // state in storage order (aux via perm)
  const int32_t* perm;
  const int32_t* cell_start;
  const int32_t* row_rank;  // (j * nk + k) -> rank
  Grid grid;
  int32_t n;
  // lists [kListCap][lstride]: entry k of slot t at k * lstride + t, so the
  // k-th entries of neighbouring slots are adjacent
  int32_t lstride;
  int2* out_list;
  int32_t* out_count;     // [n]
  int2* in_list;
  int32_t* in_count;      // [n], zero on entry
  unsigned long long* err;
  int32_t nblk8;  // blocks padded to a multiple of 8 (XCD-aware mapping)
  // skin lists: skin > 0 lists candidates (every pair with gap < skin, and
  // the few within the pre-filter margin beyond) instead of contacts; the
  // sweep runs only when the gate says a rebuild is due
  double skin;
  SkinGate gate;
};

// Contact statistics (AppCollision::addToContactsFeatures, Base/src/
// AppCollision.cpp:160-190): the overlaps of the contacts (sum, count,
// largest) and the GJK iterations reported by the closest-point routines
// (sum, count of those that report some).  Each workgroup of an evaluation
// writes its part (fixed-order reductions); one kernel adds the parts in
// workgroup order to the running totals, so the figures are reproducible.
struct StatPart {
  double osum, omax;
  unsigned long long on, gsum, gn;
};
struct ContactStats {
  double osum, omax, tmax, time;  // time: of the current evaluation
  unsigned long long on, gsum, gn;
};

// Per-thread accumulator of the statistics.
struct StatAcc {
  double osum = 0., omax = 0.;
  uint32_t on = 0, gsum = 0, gn = 0;
  __device__ __forceinline__ void add(double dist, int nbIter) {
    if (nbIter) {
      gsum += (uint32_t)nbIter;
      ++gn;
    }
    if (dist < 0.) {
      const double o = fabs(dist);
      osum += o;
      ++on;
      omax = o > omax ? o : omax;
    }
  }
};

// Contact memory of the history models (HookeMemory / HertzMemory): for each
// particle (slot order, travelling with the state copies) up to kHistCap
// entries, entry-major like the lists: key (partner id, or -1 - wall) and the
// particle's own view of the contact's memory.
constexpr int kHistCap = 16;
struct History {
  int32_t* count;  // [n]
  int32_t* key;    // [kHistCap][lstride]
  double4* kd;     // tangential spring (x, y, z, -)
  double4* nrm;    // previous contact normal
  double4* spr;    // rolling-friction spring torque
};

struct ForceArgs {
  State in, out;
  const int32_t* perm;
  const int32_t* row_rank;
  Grid grid;
  const Model* models;   // pair table, models[mat_a * model_stride + mat_b]
  int32_t model_stride;
  const Wall* walls;
  int32_t nwalls;
  int32_t n;
  double gravity[3];
  double dt_vel, dt_disp;
  int32_t lstride;  // list layout as in DetectArgs
  const int2* out_list;
  const int32_t* out_count;
  const int2* in_list;
  int32_t* in_count;  // read, then cleared for the next detection
  double4* F;  // written when non-null
  double4* T;
  int32_t* key_out;
  unsigned long long* err;
  unsigned long long* contacts;  // contact counts, one per workgroup (added to)
  ContactRec* rec;               // recording (nullable)
  int32_t* rec_count;
  int32_t rec_cap;
  int32_t nblk8;  // blocks padded to a multiple of 8 (XCD-aware mapping)
  // skin lists (gate.w == nullptr: lists hold this step's contacts, state
  // is reordered through perm every evaluation)
  SkinGate gate;
  SkinWords* words;
  float* path;                // path length since the last rebuild, slot order
  unsigned long long* cmask;  // screening masks (k_detect_screen), slot order
  // history models: memory read at the input slot, written at the output slot
  History hin, hout;
  double dt_mem;        // time step of the memory update
  // wall loads (when tracked): per particle up to kWallLoadCap wall contacts
  // [k * lstride + slot]: the wall, the force on it and its torque about the
  // wall's centre (Obstacle::addForce / addTorque); null when not tracked
  int32_t* wl_count;
  int32_t* wl_wall;
  double4* wl_f;
  double4* wl_t;
  StatPart* stat;  // contact statistics of this evaluation (nullable)
};

constexpr int kWallLoadCap = 4;

enum { MODE_FORCE = 0, MODE_KICK = 1, MODE_MOVE = 2 };

// ---- convex particles
// Extra per-particle state: basis rows (row-major) with the crust scaling of
// that axis in .w, and the principal inertia.  aux.mat holds material + 256 *
// shape.
struct CState {
  double4* rot0;
  double4* rot1;
  double4* rot2;
  double4* inert;  // I0, I3, I5, 0
};

// A candidate pair of the detection sweep (circumscribed spheres overlap):
// slot j of the detecting particle t holds (partner slot, code) with the code
// of the sphere path, or (wall index, kWallCode) for a wall whose bounding box
// grown by the circumscribed radius holds the particle's centre.  The GJK pass evaluates every candidate once; a contact
// sets bit j of the detecting particle's mask and fills record (t, j).
struct CRec {
  double c[3], ov[3], dist;
  int32_t nit;        // GJK iterations the closest point reported (statistics)
  int32_t evaluated;  // the closest point ran this step (statistics)
};
constexpr int kCandCap = 32;
constexpr int kWallCode = 128;

struct CDetectArgs {
  const float4* pos;
  State in;
  const int32_t* perm;
  const int32_t* cell_start;
  const int32_t* row_rank;
  Grid grid;
  float sizef[3];
  float margin;
  int32_t n;
  int2* cand;          // [n][kCandCap]
  int32_t* cand_count; // [n]
  uint32_t* mask;      // [n], cleared here
  int32_t* work;       // candidate queue: t * kCandCap + j
  int32_t* nwork;      // queue length (zero on entry)
  const Wall* walls;
  int32_t nwalls;
  unsigned long long* err;
  int32_t nblk8;
};

struct CGjkArgs {
  State in;
  CState cin;
  const CShape* shapes;
  const int32_t* perm;
  Grid grid;
  const int2* cand;
  const int32_t* work;
  const int32_t* nwork;
  const Wall* walls;
  const CShape* wall_shapes;  // box shape of each wall
  const double4* wall_sc;     // crust scaling per axis, circumscribed radius
  CRec* rec;           // [n][kCandCap]
  uint32_t* mask;
  int2* in_list;       // [n][kCandCap]: (detecting slot, candidate index)
  int32_t* in_count;   // [n], zero on entry
  unsigned long long* err;
  int32_t stats;       // record dist and GJK iterations of every evaluated candidate
  int32_t gjk_sv;      // 1: signed-volume GJK, 0: Johnson
};

struct CForceArgs {
  State in, out;
  CState cin, cout;
  const CShape* shapes;
  const int32_t* perm;
  const int32_t* row_rank;
  Grid grid;
  const Model* models;
  const Wall* walls;
  int32_t n;
  double gravity[3];
  double dt_vel, dt_disp;
  const int2* cand;
  const int32_t* cand_count;
  const CRec* rec;
  const uint32_t* mask;
  const int2* in_list;
  int32_t* in_count;
  int32_t* nwork;      // reset to zero for the next step
  StatPart* stat;      // contact statistics, one part per workgroup (nullable)
  double4* F;
  double4* T;
  int32_t* key_out;
  unsigned long long* err;
  unsigned long long* contacts;
  ContactRec* crec;
  int32_t* rec_count;
  int32_t rec_cap;
  int32_t nblk8;
};

void launch_cgjk(const CGjkArgs& a, int32_t max_work, hipStream_t st);
void launch_cdetect(const CDetectArgs& a, hipStream_t st);
void launch_cforce(int mode, bool record, const CForceArgs& a, hipStream_t st);
void launch_cmove(const State& s, const CState& c, const double4* F, const double4* T, int32_t n,
                  const Grid& g, const int32_t* row_rank, double dt_vel, double dt_disp,
                  int32_t* key, unsigned long long* err, hipStream_t st);
void launch_cadvance(const State& s, const CState& c, const double4* F, const double4* T,
                     int32_t n, double dt_vel, hipStream_t st);
void launch_cunpermute(const State& s, const CState& c, int32_t n, double* basis, hipStream_t st);

struct Bins {
  int32_t* key;         // cell key of each particle (storage order)
  int32_t* slot;        // arrival slot within the cell (count pass)
  int32_t* cnt;         // per-cell count (zeroed by the scan)
  int32_t* local;       // block-local exclusive scan
  int32_t* bsum;        // block sums, scanned in place
  int32_t* perm;        // sorted slot -> storage index
  int32_t* cell_start;  // ncell + 1
  float4* pos;          // sorted single-precision copy of pos_r
  int32_t* ids;         // sorted particle ids (skin lists; may be null)
  int32_t* ticket;      // bsum[nbsum]: blocks of the scan pass finished
  int32_t nbsum;
};

// Launchers (all asynchronous on the given stream).
void launch_keys(const State& s, int32_t n, const Grid& g, const int32_t* row_rank, int32_t* key,
                 unsigned long long* err, hipStream_t st);
// Counting sort of the positions of src by cell key.  Without skin lists
// (gate.w null) the keys were written by the previous integrator tail.  With
// them the binning runs only when the gate says a rebuild is due; it then
// computes the keys from src itself and clears in_count for the sweep, and
// its first kernel does the evaluation's skin-word duties (SkinWords).
void launch_binning(Bins& b, const PosSrc& src, int32_t n, const Grid& g, const int32_t* row_rank,
                    unsigned long long* err, hipStream_t st, SkinGate gate = SkinGate{},
                    SkinWords* words = nullptr, int32_t* in_count = nullptr);
// The scan and the per-cell sort of a gated binning alone (the count and
// scatter passes then run elsewhere: gdem_binning.h).
void launch_scan(const Bins& b, int32_t ncell, SkinGate gate, hipStream_t st);
void launch_list_peak(const int32_t* out, const int32_t* in, int32_t n, int32_t* peak,
                      hipStream_t st);
void launch_cellsort(const Bins& b, const PosSrc& src, int32_t ncell, int32_t n, SkinGate gate,
                     hipStream_t st);
void launch_detect(const DetectArgs& a, hipStream_t st);
// With events, they record the kernel's own start and end (timing).  memory:
// the history-model variant (a.hin / a.hout used).
void launch_force(int mode, bool record, const ForceArgs& a, hipStream_t st,
                  hipEvent_t start = nullptr, hipEvent_t stop = nullptr, bool memory = false);
// Loads of a.nwalls walls from the wall contacts the force kernel stored
// (a.wl_*): block partials part[nblocks(n)][nwalls][2], totals out[nwalls][2]
// (force, torque about the wall's centre).
void launch_wall_loads(const ForceArgs& a, double4* part, double4* out, hipStream_t st);
// Skin lists: candidate sweep (d.gate set) or screening of the lists (writes
// f.cmask), in one launch.
void launch_detect_screen(const DetectArgs& d, const ForceArgs& f, hipStream_t st);
// gate (skin lists): the path lengths grow and need[gate.p] is raised when
// one passes the limit.
void launch_move(const State& s, const double4* F, const double4* T, int32_t n, const Grid& g,
                 const int32_t* row_rank, double dt_vel, double dt_disp, int32_t* key,
                 unsigned long long* err, hipStream_t st, SkinGate gate = SkinGate{},
                 float* path = nullptr);
// Adds the nparts workgroup parts of one evaluation to the running totals,
// the evaluation's time being st->time + dt_step.
void launch_stats_final(const StatPart* part, int32_t nparts, ContactStats* st, double dt_step,
                        hipStream_t stream);
void launch_advance(const State& s, const double4* F, const double4* T, int32_t n,
                    double dt_vel, hipStream_t st);
void launch_unpermute(const State& s, const double4* F, const double4* T, int32_t n,
                      double* x, double* v, double* w, double* q, double* f, double* t,
                      hipStream_t st);
// id-ordered device arrays (NULL: kept) into the storage order of s
void launch_set_state(const State& s, double4* F, double4* T, int32_t n, const double* x,
                      const double* v, const double* w, const double* q, const double* f,
                      const double* t, hipStream_t st);

// ---- sphere step in the packed layout (skin lists, memoryless models)
// Per particle, in slot order.  The four arrays that change at every step
// hold the 16 doubles the step updates (the quaternion's components ride in
// the .w lanes), two copies (ping-pong: partners read the old copy while the
// owner writes the new one).  The constants are read-only between rebuilds
// and have two copies too: a rebuild writes them in the new order, and the
// evaluation after it copies them into the other copy.
struct Pack {
  double4* A;  // x, y, z, q.x
  double4* B;  // v, q.y
  double4* C;  // world angular velocity (.w unused), stored only when the
               // W word (wsrc) is 0; otherwise derived (world_omega)
  double4* D;  // body-frame angular velocity, q.z
  double* E;   // q.w
};
struct PackConst {
  double2* rr;  // radius, equivalent-sphere radius
  double2* mi;  // mass, principal inertia
  Aux* aux;
};

// The contact of one candidate pair, evaluated once with the reference's body
// order (first body A, second body B at A's image): the force on A, the torque
// on A about its centre and the torque on B about its image's centre.  Stored
// at the detecting particle's out-list entry, [e * lstride + t].
struct PairRec {
  double f[3], ta[3], tb[3];
};

// Segments of the pair queue (workgroup b of the screening appends to
// segment b mod kQueueSegs).
constexpr int kQueueSegs = 64;

struct PairArgs {
  DetectArgs d;        // the candidate sweep of a rebuild (d.in unused)
  Pack in;
  Pack other;          // the other copy: the orientations before the last step
  int32_t* wsrc;       // W word: 1 when in.C is not stored (world_omega)
  PackConst kin;
  uint32_t* omask;     // [n] out entries in contact (cleared by the screening)
  uint32_t* imask;     // [n] in entries in contact (atomic OR, zero on entry)
  PairRec* rec;
  const Model* models;
  int32_t model_stride;
  unsigned long long* contacts;  // contact counts, one per workgroup (added to)
  int32_t* queue;      // candidates that may touch, t * 32 + entry, in segments
  int32_t* qcount;     // [kQueueSegs] segment lengths (zero on entry)
  int32_t qseg;        // capacity of one segment
};

struct IntegArgs {
  Pack in, out;
  int32_t* wsrc;       // W word, set for the copy this kernel writes
  PackConst kin, kout;
  const int32_t* perm;
  Grid grid;
  double gmax[3];      // below nc * edge by a relative 1e-12: surely inside the grid
  const Model* models;
  int32_t model_stride;
  const Wall* walls;
  int32_t nwalls;
  int32_t n, lstride;
  double gravity[3];
  double dt_vel, dt_disp;
  double crust_min;    // smallest crust: a drift below it raises no MotionError
  const int2* out_list;
  const int2* in_list;
  const int32_t* in_count;
  const uint32_t* omask;
  uint32_t* imask;     // cleared after reading
  const PairRec* rec;
  double4* F;          // written when non-null
  double4* T;
  SkinGate gate;
  SkinWords* words;
  float* path;         // path length since the last rebuild, slot order
  unsigned long long* err;
  unsigned long long* contacts;  // per workgroup (walls)
  int32_t nblk8;
  int32_t* qcount;     // the pair queue's segment lengths, cleared for the next step
};

// Both conversions keep the W word: to_pack stores the world angular
// velocity (word 0); from_pack derives it where it is not stored, from the
// other copy (other) at the slots perm gives when the last evaluation was a
// rebuild (words->rebuilt[p]).
void launch_to_pack(const State& s, const Pack& pk, const PackConst& k0, const PackConst& k1,
                    int32_t* wsrc, int32_t n, hipStream_t st);
void launch_from_pack(const Pack& pk, const Pack& other, const PackConst& k, const State& s,
                      const int32_t* wsrc, const int32_t* perm, const SkinWords* words, int p,
                      int32_t n, hipStream_t st);
// The pair stage of a packed step, the binning of a rebuild included (its
// count and scatter passes run inside the pair kernels).
void launch_pair(const Bins& b, const PosSrc& src, SkinWords* words, const PairArgs& a,
                 hipStream_t st, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
void launch_integ(int mode, const IntegArgs& a, hipStream_t st, hipEvent_t start = nullptr,
                  hipEvent_t stop = nullptr);
// Kick + drift of the packed copy in place (the opening half of gdem_step).
void launch_move_pack(const Pack& pk, const PackConst& k, const double4* F, const double4* T,
                      int32_t n, const Grid& g, double dt_vel, double dt_disp,
                      unsigned long long* err, SkinGate gate, float* path, int32_t* wsrc,
                      hipStream_t st);

}  // namespace gdem
