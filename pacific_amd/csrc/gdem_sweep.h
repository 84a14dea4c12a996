// gdem_sweep.h -- cell indexing, pair codes, the XCD-aware block order and the
// 13-cell half-stencil sweep shared by the sphere and convex detection kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "gdem_internal.h"

#pragma clang fp contract(off)

namespace gdem {

__device__ __forceinline__ unsigned long long pack_ids(int32_t a, int32_t b) {
  return ((unsigned long long)(uint32_t)a << 32) | (unsigned long long)(uint32_t)b;
}

__device__ __forceinline__ int pack_code(int sx, int sy, int sz, bool first) {
  return (sx + 1) | ((sy + 1) << 2) | ((sz + 1) << 4) | (first ? 64 : 0);
}

__device__ __forceinline__ int mirror_code(int code) {
  return pack_code(1 - (code & 3), 1 - ((code >> 2) & 3), 1 - ((code >> 4) & 3), !(code & 64));
}

// Cell::GetCell (Base/src/Cell.cpp:172-184); periodic indices are clamped to
// the grid after the wrap of the position.  Returns false outside the grid.
__device__ __forceinline__ bool cell_of(const Grid& g, v3 x, int* ijk) {
  const double p[3] = {x.x, x.y, x.z};
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double t = (p[d] - g.origin[d]) / g.edge[d];
    int id;
    if (!(t == t) || t > 2.0e9 || t < -2.0e9) {
      ok = false;
      id = 0;
    } else {
      id = (int)floor(t);
      if (g.periodic[d]) {
        if (id >= g.nc[d]) id = g.nc[d] - 1;
        if (id < 0) id = 0;
      } else if (id < 0 || id >= g.nc[d]) {
        ok = false;
      }
    }
    ijk[d] = id;
  }
  return ok;
}

__device__ __forceinline__ int row_base(const Grid& g, const int32_t* row_rank, int j, int k) {
  return row_rank[j * g.nc[2] + k] * g.nc[0];
}

__device__ __forceinline__ int logical_index(int nblk8) {
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous range of cells so that neighbour reads stay in its own L2.
  const int b = blockIdx.x;
  const int per = nblk8 >> 3;
  const int lb = (b & 7) * per + (b >> 3);
  return lb * blockDim.x + threadIdx.x;
}

// LinkedCell::ComputeForces' half stencil (Base/src/LinkedCell.cpp:958-984)
// swept as 5 contiguous slot ranges: cells of a row along x are consecutive in
// the sorted order.  det.range(beg, end, sx, sy, sz, same) is called for each
// range with the periodic shift of the candidates' image.
template <class Det>
__device__ __forceinline__ void half_stencil_sweep(const Grid& g, const int32_t* row_rank,
                                                   const int32_t* cs, int t, const int* ijk,
                                                   Det& det) {
  const int i = ijk[0], j = ijk[1], k = ijk[2];
  const int ni = g.nc[0];
  const bool px = g.periodic[0] != 0;
  // own row: the rest of this cell (same cell, after self), then cell i+1
  {
    const int rb = row_base(g, row_rank, j, k);
    det.range(t + 1, cs[rb + i + 1], 0, 0, 0, true);
    if (i + 1 < ni) det.range(cs[rb + i + 1], cs[rb + i + 2], 0, 0, 0, false);
    else if (px) det.range(cs[rb], cs[rb + 1], 1, 0, 0, false);
  }
  // rows (j, k+1), (j+1, k-1), (j+1, k), (j+1, k+1): cells i-1 .. i+1
  const int drow[4][2] = {{0, 1}, {1, -1}, {1, 0}, {1, 1}};
  for (int q = 0; q < 4; ++q) {
    int jj = j + drow[q][0], kk = k + drow[q][1];
    int sy = 0, sz = 0;
    if (jj >= g.nc[1]) {
      if (!g.periodic[1]) continue;
      jj -= g.nc[1];
      sy = 1;
    }
    if (kk < 0 || kk >= g.nc[2]) {
      if (!g.periodic[2]) continue;
      if (kk < 0) { kk += g.nc[2]; sz = -1; }
      else { kk -= g.nc[2]; sz = 1; }
    }
    const int rb = row_base(g, row_rank, jj, kk);
    const int lo = max(i - 1, 0), hi = min(i + 1, ni - 1);
    if (px && i - 1 < 0) det.range(cs[rb + ni - 1], cs[rb + ni], -1, sy, sz, false);
    det.range(cs[rb + lo], cs[rb + hi + 1], 0, sy, sz, false);
    if (px && i + 1 >= ni) det.range(cs[rb], cs[rb + 1], 1, sy, sz, false);
  }
}

// Reduces the per-thread statistics of a workgroup of bs threads into its
// part (butterfly within each wave, then waves in order; every thread calls).
template <int bs>
__device__ __forceinline__ void block_stats(const StatAcc& a, StatPart* part, StatPart* sh) {
  double osum = a.osum, omax = a.omax;
  unsigned long long on = a.on, gsum = a.gsum, gn = a.gn;
  for (int o = 32; o > 0; o >>= 1) {
    osum += __shfl_xor(osum, o, 64);
    const double m = __shfl_xor(omax, o, 64);
    omax = m > omax ? m : omax;
    on += __shfl_xor(on, o, 64);
    gsum += __shfl_xor(gsum, o, 64);
    gn += __shfl_xor(gn, o, 64);
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = StatPart{osum, omax, on, gsum, gn};
  __syncthreads();
  if (threadIdx.x == 0) {
    StatPart t = sh[0];
    for (int w = 1; w < bs / 64; ++w) {
      t.osum += sh[w].osum;
      t.omax = sh[w].omax > t.omax ? sh[w].omax : t.omax;
      t.on += sh[w].on;
      t.gsum += sh[w].gsum;
      t.gn += sh[w].gn;
    }
    part[blockIdx.x] = t;
  }
}

// Next entry of an incoming contact list (entry f at il[f * stride], taken
// only where bit f of mask is set) in the order (partner slot, code, list
// position), after the entry of key `last` (-1 before the first): returns its
// position and key, or -1 when every entry has been taken.  The position
// breaks ties, so an entry that appears twice (a particle swept from a cell it
// is not binned in, after it left the grid) is still visited once per copy and
// the walk never runs past the list.
__host__ __device__ __forceinline__ int next_incoming(const int2* il, int nin, int stride,
                                                      long long last, long long& key,
                                                      uint32_t mask = 0xffffffffu) {
  long long best = 0x7fffffffffffffffll;
  int bi = -1;
  for (int f = 0; f < nin; ++f) {
    if (!((mask >> f) & 1u)) continue;
    const int2 en = il[(size_t)f * stride];
    const long long k = ((long long)en.x << 12) | ((long long)(en.y & 127) << 5) | (long long)f;
    if (k > last && k < best) {
      best = k;
      bi = f;
    }
  }
  key = best;
  return bi;
}

// Cell of a position clamped to the grid: the cell whose half stencil a
// particle sweeps (a particle outside a non-periodic grid is reported by the
// binning as a SimulationError).
__device__ __forceinline__ void clamped_cell(const Grid& g, v3 x, int* ijk) {
  cell_of(g, x, ijk);
#pragma unroll
  for (int d = 0; d < 3; ++d) ijk[d] = min(max(ijk[d], 0), g.nc[d] - 1);
}

}  // namespace gdem
