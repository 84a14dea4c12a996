// gdem_binning.h -- the count and scatter passes of the counting sort by
// cell key (LinkedCell::LinkUpdate, Base/src/LinkedCell.cpp:1452-1547), as
// device functions: the binning kernels call them, and the packed sphere
// step runs them inside kernels whose other branch does the work of a step
// without a rebuild.
#pragma once
#include <hip/hip_runtime.h>

#include "gdem_internal.h"
#include "gdem_sweep.h"

#pragma clang fp contract(off)

namespace gdem {

// Cell key of position p (as k_keys); -1 outside the grid.
__device__ __forceinline__ int key_of(const Grid& g, const int32_t* row_rank, double4 p) {
  int ijk[3];
  if (!cell_of(g, mk(p.x, p.y, p.z), ijk)) return -1;
  return row_base(g, row_rank, ijk[1], ijk[2]) + ijk[0];
}

// The first kernel of an evaluation with skin lists clears the need word of
// the next evaluation and counts the rebuilds.
__device__ __forceinline__ void skin_words_duty(const SkinGate& gate, SkinWords* words,
                                                bool rebuild) {
  words->need[1 - gate.p] = 0;
  if (rebuild) words->rebuilds += 1;
}

// Arrival slot of each particle in its cell (grid-stride over blocks of
// particles).  Keys arrive almost sorted (the storage order is the previous
// step's cell order), so equal keys of neighbouring lanes are merged into one
// atomic per run.  With skin lists (from_src) the keys are computed here
// from src and the incoming counts are cleared for the sweep.
__device__ __forceinline__ void count_pass(int32_t* key, int32_t* slot, int32_t* cnt, int n,
                                           bool from_src, int32_t* in_count, const PosSrc& src,
                                           const Grid& g, const int32_t* row_rank,
                                           unsigned long long* err) {
  const int lane = threadIdx.x & 63;
  // grid-stride over blocks of particles: the loop bound is uniform in the
  // block, so every lane of a wave takes part in the shuffles and ballots
  for (int b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
    const int i = b0 + threadIdx.x;
    int k = -1;
    if (i < n) {
      if (from_src) {
        k = key_of(g, row_rank, src.pos(i));
        if (k < 0) {
          atomicMin(&err[kErrSim], (unsigned long long)(uint32_t)src.aux[i].id);
          k = 0;
        }
        key[i] = k;
      } else {
        k = key[i];
      }
    }
    const int kprev = __shfl_up(k, 1, 64);
    const bool head = (lane == 0) || (k != kprev);
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int head_lane = 63 - __clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int next_head = after ? __ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (head && k >= 0) base = atomicAdd(&cnt[k], next_head - lane);
    base = __shfl(base, head_lane, 64);
    if (i < n) slot[i] = base + (lane - head_lane);
    // skin lists: the rebuild's sweep appends to cleared incoming counts
    if (in_count && i < n) in_count[i] = 0;
  }
}

// Sorted position of every particle: its cell's start plus its arrival slot
// (the per-cell sort then orders each cell by previous slot).
__device__ __forceinline__ void scatter_pass(const int32_t* key, const int32_t* slot,
                                             const int32_t* local, const int32_t* bsum,
                                             int32_t* perm, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int k = key[i];
    perm[local[k] + bsum[k / kScanCells] + slot[i]] = i;
  }
}

}  // namespace gdem
