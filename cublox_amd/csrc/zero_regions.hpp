// Zeroing of recenter slab regions, shared by the recenter launches
// (kernels.hip) and the fused update's walk launch, whose extra workgroups
// clear a pending recenter's slabs beside the ray walk (integrate.hip).
//
// A region is n_outer runs of `run` contiguous voxels, `stride` apart,
// starting at `base` (the host merges a recenter's slice lists into a few
// regions, reference kernels.cu:404-499 re-laid out). A launch cuts its
// regions into chunks of about `piece` elements, one workgroup each.
#pragma once

#include "hip_common.hpp"

#include <algorithm>

namespace cublox_hip {

struct ZeroRegion {
  long long base;
  long long n_outer;
  long long stride;
  long long run;
};

// A region cut into chunks of about `piece` elements, each zeroed by one
// workgroup: the whole store set is spread evenly over the launch whatever
// the regions' shapes. Chunks of the region are numbered from chunk0.
struct ChunkedRegion {
  ZeroRegion r;
  long long chunk0;
  long long piece;     // elements per chunk (a multiple of 64)
  long long per_row;   // chunks per row (run >= piece) or rows per chunk
};

template <typename T>
__device__ __forceinline__ void zero_contiguous(CUBLOX_GLOBAL T *__restrict__ p,
                                               long long count, long long t,
                                               long long stride) {
  // Scalar head up to 16 B alignment, then 16 B stores, then scalar tail.
  const uintptr_t addr = reinterpret_cast<uintptr_t>(p);
  long long head = static_cast<long long>(((16u - (addr & 15u)) & 15u) / 4u);
  if (head > count) head = count;
  const long long n4 = (count - head) / 4;
  const long long tail0 = head + n4 * 4;
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  CUBLOX_GLOBAL u32x4 *__restrict__ p4 = (CUBLOX_GLOBAL u32x4 *)(p + head);
  const u32x4 z = {0u, 0u, 0u, 0u};
  for (long long i = t; i < n4; i += stride) p4[i] = z;
  if (t < head) p[t] = T(0);
  if (t < count - tail0) p[tail0 + t] = T(0);
}

// Zeroes one region with the rows of the region spread over the
// workgroups of this region (blk of nblk): runs of at least 256 elements
// take a workgroup per row with 16 B stores; shorter runs pack
// 256 / run rows into a workgroup, one element per lane (one division per
// lane, none per element).
template <typename T>
__device__ __forceinline__ void zero_region(CUBLOX_GLOBAL T *__restrict__ p,
                                           const ZeroRegion &r, int blk,
                                           int nblk) {
  if (r.n_outer == 1) {
    zero_contiguous(p + r.base, r.run,
                    static_cast<long long>(blk) * blockDim.x + threadIdx.x,
                    static_cast<long long>(nblk) * blockDim.x);
    return;
  }
  if (r.run >= 256) {
    for (long long o = blk; o < r.n_outer; o += nblk) {
      zero_contiguous(p + r.base + o * r.stride, r.run, threadIdx.x, blockDim.x);
    }
    return;
  }
  const int run = static_cast<int>(r.run);
  const int rows = static_cast<int>(blockDim.x) / run;
  const int my_row = static_cast<int>(threadIdx.x) / run;
  const int col = static_cast<int>(threadIdx.x) - my_row * run;
  if (my_row >= rows) return;
  for (long long o = static_cast<long long>(blk) * rows + my_row; o < r.n_outer;
       o += static_cast<long long>(nblk) * rows) {
    p[r.base + o * r.stride + col] = T(0);
  }
}

// The part of region c.r that chunk j (0-based within the region) covers.
__host__ __device__ __forceinline__ ZeroRegion chunk_part(const ChunkedRegion &c,
                                                        long long j) {
  const ZeroRegion &r = c.r;
  if (r.n_outer == 1 || r.run >= c.piece) {  // pieces of rows
    const long long row = j / c.per_row, part = j - row * c.per_row;
    const long long off = part * c.piece;
    return ZeroRegion{r.base + row * r.stride + off, 1, 0,
                      r.run - off < c.piece ? r.run - off : c.piece};
  }
  const long long row0 = j * c.per_row;  // whole short rows
  const long long rows = r.n_outer - row0 < c.per_row ? r.n_outer - row0 : c.per_row;
  return ZeroRegion{r.base + row0 * r.stride, rows, r.stride, r.run};
}

constexpr long long kChunkElems = 4096;     // at least 16 KB of float per workgroup

// Cuts a region into chunks of about `piece` elements (rows are split when
// longer, whole rows grouped when shorter); returns the chunk count.
inline long long chunk_region(const ZeroRegion &r, long long piece, long long chunk0,
                              ChunkedRegion *out) {
  out->r = r;
  out->chunk0 = chunk0;
  out->piece = piece;
  if (r.n_outer == 1 || r.run >= piece) {
    out->per_row = (r.run + piece - 1) / piece;
    return r.n_outer * out->per_row;
  }
  out->per_row = std::max<long long>(1, piece / r.run);
  return (r.n_outer + out->per_row - 1) / out->per_row;
}

// Elements per chunk so that a launch of `blocks` workgroups takes about
// one chunk each (at least kChunkElems, a multiple of 64).
inline long long chunk_piece(long long elems, long long blocks) {
  const long long p = (elems + blocks - 1) / std::max<long long>(blocks, 1);
  return std::max(kChunkElems, (p + 63) / 64 * 64);
}

// A recenter's clear staged on the host for a later launch (the bytes
// behind cublox_recenter_plan_stage): the plan's action (1: slabs, 2: a full
// reset), the voxel count, and the merged slab regions.
constexpr int kStagedRegions = 48;
struct StagedClear {
  int action;
  int n;
  long long voxels;
  ZeroRegion r[kStagedRegions];
};

} // namespace cublox_hip
