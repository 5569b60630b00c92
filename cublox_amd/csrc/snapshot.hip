// Map snapshots (no reference counterpart: the reference has no way to save
// or restore a grid). A snapshot holds the occupancy either densely or as
// the (hash, value) pairs of its nonzero voxels in ascending hash order;
// these kernels build the sparse form in three passes (count per chunk,
// scan of the chunk counts, ordered write) and scatter it back.
#include <algorithm>

#include "hip_common.hpp"

namespace cublox_hip {
namespace {

constexpr int kSnapThreads = 256;
constexpr int kSnapPer = 32;                            // voxels per thread
constexpr int kSnapChunk = kSnapThreads * kSnapPer;     // voxels per workgroup
constexpr int kScanThreads = 1024;

long long snap_chunks(int V) { return (static_cast<long long>(V) + kSnapChunk - 1) / kSnapChunk; }

// A voxel is stored when its bit pattern is nonzero (so -0.0 survives).
__global__ void __launch_bounds__(kSnapThreads)
    sparse_count_kernel(const unsigned *__restrict__ occ, int V,
                        unsigned *__restrict__ counts) {
  __shared__ int s_wave[kSnapThreads / 64];
  const long long c0 = static_cast<long long>(blockIdx.x) * kSnapChunk;
  int n = 0;
#pragma unroll 8
  for (int j = 0; j < kSnapPer; ++j) {
    const long long h = c0 + j * kSnapThreads + threadIdx.x;
    if (h < V && occ[h] != 0u) ++n;
  }
  const int incl = block_scan_incl(s_wave, n);
  if (threadIdx.x == kSnapThreads - 1) counts[blockIdx.x] = static_cast<unsigned>(incl);
}

// Exclusive scan of the chunk counts in place (one workgroup, each thread a
// contiguous run of chunks); the total goes to *total.
__global__ void __launch_bounds__(kScanThreads)
    sparse_scan_kernel(unsigned *__restrict__ counts, long long n_chunks,
                       unsigned *__restrict__ total) {
  __shared__ int s_wave[kScanThreads / 64];
  const long long per = (n_chunks + kScanThreads - 1) / kScanThreads;
  const long long b = threadIdx.x * per;
  const long long e = std::min(b + per, n_chunks);
  unsigned sum = 0;
  for (long long k = b; k < e; ++k) sum += counts[k];
  const int incl = block_scan_incl(s_wave, static_cast<int>(sum));
  unsigned run = static_cast<unsigned>(incl) - sum;
  for (long long k = b; k < e; ++k) {
    const unsigned c = counts[k];
    counts[k] = run;
    run += c;
  }
  if (threadIdx.x == kScanThreads - 1) *total = static_cast<unsigned>(incl);
}

// Ordered write: each workgroup walks its chunk in rows of kSnapThreads
// consecutive voxels; a block scan ranks the nonzero ones of each row.
__global__ void __launch_bounds__(kSnapThreads)
    sparse_write_kernel(const unsigned *__restrict__ occ, int V,
                        const unsigned *__restrict__ offsets,
                        int *__restrict__ ids, float *__restrict__ vals,
                        unsigned cap) {
  __shared__ int s_wave[kSnapThreads / 64];
  const long long c0 = static_cast<long long>(blockIdx.x) * kSnapChunk;
  unsigned base = offsets[blockIdx.x];
  for (int j = 0; j < kSnapPer; ++j) {
    const long long h = c0 + j * kSnapThreads + threadIdx.x;
    const unsigned bits = h < V ? occ[h] : 0u;
    const int nz = bits != 0u ? 1 : 0;
    const int incl = block_scan_incl(s_wave, nz);
    const unsigned row_total = static_cast<unsigned>(s_wave[kSnapThreads / 64 - 1]);
    if (nz) {
      const unsigned pos = base + static_cast<unsigned>(incl) - 1u;
      if (pos < cap) {
        ids[pos] = static_cast<int>(h);
        vals[pos] = __uint_as_float(bits);
      }
    }
    base += row_total;
    __syncthreads();  // s_wave is rewritten by the next row's scan
  }
}

__global__ void __launch_bounds__(256)
    sparse_scatter_kernel(float *__restrict__ occ, int V,
                          const int *__restrict__ ids,
                          const float *__restrict__ vals, unsigned n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = ids[i];
  if (h >= 0 && h < V) occ[h] = vals[i];
}

}  // namespace
}  // namespace cublox_hip

using namespace cublox_hip;

extern "C" {

size_t cublox_sparse_scratch_bytes(int voxel_num) {
  if (voxel_num <= 0) return 256;
  return static_cast<size_t>(snap_chunks(voxel_num)) * sizeof(unsigned) + 256;
}

int cublox_launch_sparse_compact(const float *d_occ, int voxel_num,
                                 void *d_scratch, size_t scratch_bytes,
                                 int *d_ids, float *d_vals,
                                 unsigned int capacity, unsigned int *d_count,
                                 cublox_stream s) {
  if (!d_count) return set_error(CUBLOX_ERR_INVALID, "sparse_compact: null count");
  if (voxel_num <= 0 || !d_occ) return cublox_fill_zero_async(d_count, 1, s);
  if (!d_scratch || scratch_bytes < cublox_sparse_scratch_bytes(voxel_num)) {
    return set_error(CUBLOX_ERR_CAPACITY, "sparse_compact: scratch too small");
  }
  if (capacity > 0 && (!d_ids || !d_vals)) {
    return set_error(CUBLOX_ERR_INVALID, "sparse_compact: output arrays missing");
  }
  const long long n_chunks = snap_chunks(voxel_num);
  unsigned *counts = static_cast<unsigned *>(d_scratch);
  const unsigned *bits = reinterpret_cast<const unsigned *>(d_occ);
  hipStream_t hs = as_stream(s);
  sparse_count_kernel<<<dim3(static_cast<unsigned>(n_chunks)), dim3(kSnapThreads), 0, hs>>>(
      bits, voxel_num, counts);
  int st = check_launch("sparse_count_kernel");
  if (st != CUBLOX_OK) return st;
  sparse_scan_kernel<<<dim3(1), dim3(kScanThreads), 0, hs>>>(counts, n_chunks, d_count);
  st = check_launch("sparse_scan_kernel");
  if (st != CUBLOX_OK || capacity == 0) return st;
  sparse_write_kernel<<<dim3(static_cast<unsigned>(n_chunks)), dim3(kSnapThreads), 0, hs>>>(
      bits, voxel_num, counts, d_ids, d_vals, capacity);
  return check_launch("sparse_write_kernel");
}

int cublox_launch_sparse_restore(float *d_occ, int voxel_num, const int *d_ids,
                                 const float *d_vals, unsigned int n,
                                 cublox_stream s) {
  if (voxel_num <= 0 || !d_occ) return CUBLOX_OK;
  int st = cublox_fill_zero_async(d_occ, static_cast<size_t>(voxel_num), s);
  if (st != CUBLOX_OK || n == 0) return st;
  if (!d_ids || !d_vals) return set_error(CUBLOX_ERR_INVALID, "sparse_restore: null arrays");
  sparse_scatter_kernel<<<dim3((n + 255) / 256), dim3(256), 0, as_stream(s)>>>(
      d_occ, voxel_num, d_ids, d_vals, n);
  return check_launch("sparse_scatter_kernel");
}

}  // extern "C"
