// Occupancy-update kernels for CDNA4 (gfx950, wave64) and their C-ABI
// launchers: ray cast (with the fused scan transform), apply, recenter slab
// reset, index clear and sphere fetch.
//
// Reference semantics: src/kernels.cu of the reference (lines cited per
// kernel). Arithmetic is FP32 with contraction off and IEEE division/sqrt,
// so the traversed voxel sets are bit-identical to the CPU restatement.
#include "raycast_common.hpp"
#include "zero_regions.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

namespace cublox_hip {

// ===================================================================
// Ray cast (reference rayCastUpdateKernel, kernels.cu:33-152): one ray per
// lane, one counter atomic per traversed voxel. This is the reference-
// shaped launcher; the grid's own update path is the tile-binned fused
// update in integrate.hip.
// ===================================================================

// Steps below this count are near the sensor voxel, where every ray of the
// scan passes through the same few voxels (each step moves one voxel further
// in L1 distance from the sensor voxel). Their increments are deduplicated
// across the wavefront before the global atomic.
constexpr int kNearSteps = 12;

template <bool kTransform>
__global__ void __launch_bounds__(256)
    ray_cast_kernel(const RayArgs a, const float *__restrict__ xs,
                    const float *__restrict__ ys, const float *__restrict__ zs,
                    const int n, int *__restrict__ op_cnt,
                    int *__restrict__ hit_cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  RayState r;
  if (!ray_setup<kTransform>(a, xs, ys, zs, i, r)) return;

  // kernels.cu:106-137. The hash coordinates are tracked incrementally
  // (one wrap compare per step instead of three integer remainders); every
  // lane of a wavefront is at the same step, so the near-field test is
  // uniform.
  HashCoord c = hash_coord_of(a, r.ix, r.iy, r.iz);
  const int yz = a.my * a.mz;
  for (int step = 0; step < r.max_steps; ++step) {
    if (r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) break;
    const bool inwin = abs(r.ix - a.oix) <= a.hx && abs(r.iy - a.oiy) <= a.hy &&
                       abs(r.iz - a.oiz) <= a.hz;
    const int h = inwin ? c.cx * yz + c.cy * a.mz + c.cz : -1;
    if (step < kNearSteps) {
      wave_add_unique(op_cnt, h);
    } else if (h >= 0) {
      atomic_add_i32(op_cnt + h, 1);
    }
    const int ax = ray_advance(r);
    hash_coord_step(c, ax, r, a);
  }

  // kernels.cu:141-151: the endpoint counts even when the walk did not reach
  // it; a hit only for real returns.
  if (inside_window(r.eix, r.eiy, r.eiz, a.oix, a.oiy, a.oiz, a.hx, a.hy, a.hz)) {
    const int h = global_to_hash(r.eix, r.eiy, r.eiz, a.mx, a.my, a.mz, a.hx,
                                 a.hy, a.hz);
    if (h >= 0 && h < a.vox_num) {
      atomic_add_i32(op_cnt + h, 1);
      if (!r.clipped) atomic_add_i32(hit_cnt + h, 1);
    }
  }
}

// ===================================================================
// Apply (reference applyUpdateKernel, kernels.cu:167-210)
// ===================================================================

// kernels.cu:184-200 for one voxel with op != 0. Returns the new stored
// value through *out and whether it differs from the old one.
__device__ __forceinline__ bool apply_voxel(float *__restrict__ occ,
                                            int *__restrict__ hit_cnt, int h,
                                            int op, const ApplyArgs &p,
                                            float *out) {
  const int hit = hit_cnt[h];
  const float stored0 = occ[h];
  const float stored = apply_logodds(stored0, op, hit, p);
  occ[h] = stored;
  if (hit != 0) hit_cnt[h] = 0;
  *out = stored;
  return stored != stored0;
}

__device__ __forceinline__ void record_modified(unsigned *count, int *idx,
                                                float *val, unsigned cap,
                                                int h, float v) {
  const unsigned slot = atomic_add_u32_ret(count, 1u);
  if (slot < cap) {
    idx[slot] = h;
    val[slot] = v;
  }
}

// One pass over all V counters in 16 B loads, four chunks in flight per
// thread; only voxels with op != 0 touch hit/occ.
template <bool kRecord>
__global__ void __launch_bounds__(256)
    apply_kernel(float *__restrict__ occ, int *__restrict__ op_cnt,
                 int *__restrict__ hit_cnt, const int V, const ApplyArgs p,
                 unsigned *__restrict__ mod_count, int *__restrict__ mod_idx,
                 float *__restrict__ mod_val, const unsigned mod_cap) {
  constexpr int kUnroll = 4;
  int4 *__restrict__ op4 = reinterpret_cast<int4 *>(op_cnt);
  const int n4 = V / 4;
  const int stride = gridDim.x * blockDim.x;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  for (int base = t; base < n4; base += stride * kUnroll) {
    int4 o[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int c = base + u * stride;
      o[u] = c < n4 ? op4[c] : make_int4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if ((o[u].x | o[u].y | o[u].z | o[u].w) == 0) continue;
      const int c = base + u * stride;
      const int h0 = c * 4;
      const int ov[4] = {o[u].x, o[u].y, o[u].z, o[u].w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (ov[k] == 0) continue;
        float v;
        const bool changed = apply_voxel(occ, hit_cnt, h0 + k, ov[k], p, &v);
        if (kRecord && changed) {
          record_modified(mod_count, mod_idx, mod_val, mod_cap, h0 + k, v);
        }
      }
      op4[c] = make_int4(0, 0, 0, 0);
    }
  }
  // Tail of V % 4 voxels.
  const int tail = V - n4 * 4;
  if (t < tail) {
    const int h = n4 * 4 + t;
    const int o = op_cnt[h];
    if (o != 0) {
      float v;
      const bool changed = apply_voxel(occ, hit_cnt, h, o, p, &v);
      if (kRecord && changed) {
        record_modified(mod_count, mod_idx, mod_val, mod_cap, h, v);
      }
      op_cnt[h] = 0;
    }
  }
}

// ===================================================================
// Index clear (reference clearVoxelsByIndexKernel, kernels.cu:273-284)
// ===================================================================

__global__ void __launch_bounds__(256)
    clear_by_index_kernel(float *__restrict__ occ,
                          const int *__restrict__ idx, const int n,
                          const int V) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int h = idx[t];
  if (h >= 0 && h < V) occ[h] = 0.0f;
}

// ===================================================================
// Recenter slab reset (reference kernels.cu:404-499, re-laid out)
//
// The host turns the slice lists into at most a few "regions": a region is
// n_outer runs of `run` contiguous voxels, `stride` apart, starting at
// `base`. Merged x-slices are one contiguous range; y-slices are nx runs of
// k*nz; z-slices are nx*ny runs of k. Regions travel by value.
// ===================================================================

constexpr int kMaxRegions = 48;

struct ZeroArgs {
  ChunkedRegion r[kMaxRegions];
  int n;
};

__global__ void __launch_bounds__(256)
    zero_regions_kernel(float *__restrict__ occ, int *__restrict__ op_cnt,
                        int *__restrict__ hit_cnt, const ZeroArgs z) {
  const long long j = blockIdx.x;
  int k = 0;  // the region holding chunk j (chunk0 ascending)
  while (k + 1 < z.n && z.r[k + 1].chunk0 <= j) ++k;
  const ZeroRegion part = chunk_part(z.r[k], j - z.r[k].chunk0);
  zero_region(as_global(occ), part, 0, 1);
  if (op_cnt) zero_region(as_global(op_cnt), part, 0, 1);
  if (hit_cnt) zero_region(as_global(hit_cnt), part, 0, 1);
}

// Batched recenter: regions of many grids, each with its occupancy base.
struct ZeroRegionP {
  float *occ;
  ChunkedRegion c;
};

struct RecenterBatchHeader {
  int n_regions;
  int pad;
  long long n_chunks;
  int pad2[4];
};

__global__ void __launch_bounds__(256)
    zero_regions_batch_kernel(const RecenterBatchHeader *__restrict__ hdr,
                              const ZeroRegionP *__restrict__ regions) {
  const RecenterBatchHeader h = load_const(hdr);
  for (long long j = blockIdx.x; j < h.n_chunks; j += gridDim.x) {
    int lo = 0, hi = h.n_regions - 1;  // last region with chunk0 <= j
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      const long long c0 = ((const CUBLOX_CONST long long *)&regions[mid].c.chunk0)[0];
      if (c0 <= j) {
        lo = mid;
      } else {
        hi = mid - 1;
      }
    }
    const ZeroRegionP q = load_const(regions + lo);
    zero_region(as_global(q.occ), chunk_part(q.c, j - q.c.chunk0), 0, 1);
  }
}

// Reference-shaped slab clear from a DEVICE slice list: one block row per
// slice, lanes along the contiguous z runs (x/y slabs) or across rows (z).
__global__ void __launch_bounds__(256)
    clear_slabs_list_kernel(float *__restrict__ occ, int *__restrict__ op_cnt,
                            int *__restrict__ hit_cnt, const int mx,
                            const int my, const int mz, const int hx,
                            const int hy, const int hz,
                            const int *__restrict__ slices, const int axis) {
  const int l = slices[blockIdx.y];
  long long base, n_outer, stride, run;
  if (axis == 0) {
    base = static_cast<long long>(l + hx) * my * mz;
    n_outer = 1;
    stride = 0;
    run = static_cast<long long>(my) * mz;
  } else if (axis == 1) {
    base = static_cast<long long>(l + hy) * mz;
    n_outer = mx;
    stride = static_cast<long long>(my) * mz;
    run = mz;
  } else {
    base = l + hz;
    n_outer = static_cast<long long>(mx) * my;
    stride = mz;
    run = 1;
  }
  if (l < -(axis == 0 ? hx : axis == 1 ? hy : hz) ||
      l > (axis == 0 ? hx : axis == 1 ? hy : hz)) {
    return;
  }
  const ZeroRegion r{base, n_outer, stride, run};
  zero_region(as_global(occ), r, blockIdx.x, gridDim.x);
  if (op_cnt) zero_region(as_global(op_cnt), r, blockIdx.x, gridDim.x);
  if (hit_cnt) zero_region(as_global(hit_cnt), r, blockIdx.x, gridDim.x);
}

// ===================================================================
// Sphere fetch (reference fetchOccupancySphereKernel, kernels.cu:506-562)
//
// Lanes run along z inside each (x, y) column of the disk so that occ reads
// are contiguous; kept samples are compacted with one atomic per wavefront.
// ===================================================================

// Geometry of a fetch's enumeration: the disk x z-range, row-major with z
// fastest (lanes along z read contiguous occ). The host plan keeps
// diam * diam * nz below 2^31 (kernels.cu:575-580).
struct FetchSpace {
  int r, diam, nz;
  unsigned total;
  FastDiv by_nz, by_diam;
  bool ctr;
  int qx, qy, qz;  // window origin modulo the map size (non-negative)
};

// Hash coordinate of an in-window global index g on an axis with window
// origin o, half size h, map m = 2h + 1 and q = o mod m: (g + h) mod m,
// which is global_to_local(g) + h (utils.cuh:60-70), with one compare.
__device__ __forceinline__ int in_window_coord(int g, int o, int h, int m, int q) {
  const int v = q + (g - o + h);
  return v >= m ? v - m : v;
}

__device__ __forceinline__ bool fetch_space(const cublox_fetch_cfg &c, FetchSpace &f) {
  f.r = c.disk_radius_vox;
  f.diam = 2 * f.r + 1;
  f.nz = c.gz_max - c.gz_min + 1;
  if (f.diam <= 0 || f.nz <= 0) return false;
  f.total = static_cast<unsigned>(f.diam) * f.diam * f.nz;
  f.by_nz = make_fastdiv(static_cast<unsigned>(f.nz));
  f.by_diam = make_fastdiv(static_cast<unsigned>(f.diam));
  f.ctr = c.origin_at_center != 0;
  auto pmod = [](int v, int m) {
    const int r = v % m;
    return r < 0 ? r + m : r;
  };
  f.qx = pmod(c.origin_i.x, c.map_size_i.x);
  f.qy = pmod(c.origin_i.y, c.map_size_i.y);
  f.qz = pmod(c.origin_i.z, c.map_size_i.z);
  return true;
}

// kernels.cu:520-561 for element e < total: in the window, in the sphere
// and passing the state filter; writes the voxel position and state.
__device__ __forceinline__ bool fetch_classify(const cublox_fetch_cfg &c,
                                               const float *__restrict__ occ,
                                               const FetchSpace &f, unsigned e,
                                               float &px, float &py, float &pz,
                                               unsigned char &st) {
  const unsigned rem = fast_div(e, f.by_nz);
  const int tz = static_cast<int>(e - rem * static_cast<unsigned>(f.nz));
  const unsigned tx_u = fast_div(rem, f.by_diam);
  const int ty = static_cast<int>(rem - tx_u * static_cast<unsigned>(f.diam));
  const int tx = static_cast<int>(tx_u);
  const int gx = c.center_g.x + tx - f.r;
  const int gy = c.center_g.y + ty - f.r;
  const int gz = c.gz_min + tz;
  if (!inside_window(gx, gy, gz, c.origin_i.x, c.origin_i.y, c.origin_i.z,
                     c.half_map_size_i.x, c.half_map_size_i.y, c.half_map_size_i.z)) {
    return false;
  }
  px = global_to_pos(gx, c.resolution, f.ctr);
  py = global_to_pos(gy, c.resolution, f.ctr);
  pz = global_to_pos(gz, c.resolution, f.ctr);
  const float dx = px - c.center.x;
  const float dy = py - c.center.y;
  const float dz = pz - c.center.z;
  if ((dx * dx + dy * dy) + dz * dz > c.radius_sq) return false;
  const int h =
      (in_window_coord(gx, c.origin_i.x, c.half_map_size_i.x, c.map_size_i.x, f.qx) *
           c.map_size_i.y +
       in_window_coord(gy, c.origin_i.y, c.half_map_size_i.y, c.map_size_i.y, f.qy)) *
          c.map_size_i.z +
      in_window_coord(gz, c.origin_i.z, c.half_map_size_i.z, c.map_size_i.z, f.qz);
  const float l = occ[h] + c.l_unknown;
  st = 0;
  if (l >= c.l_occupied) {
    st = 2;
  } else if (l < c.l_free) {
    st = 1;
  }
  return c.state_filter < 0 || static_cast<int>(st) == c.state_filter;
}

// Workgroup-staged form for one large query: kept samples (x, y, z, state)
// gather in LDS and each flush reserves its slots on *count with one
// atomic (a flush every kFetchStage - 256 samples at most, usually one per
// workgroup), so the shared counter is not hit once per wave. The loop is
// workgroup-uniform; flush(n, base) runs in every thread with the staged
// samples s_stage[0, n) holding slots base + k.
constexpr int kFetchThreads = 256;
constexpr int kFetchStage = 1024;

struct FetchLds {
  float4 stage[kFetchStage];
  int wave[kFetchThreads / 64];
  int misc[4];
};

template <typename Flush>
__device__ __forceinline__ void staged_flush(unsigned *__restrict__ count, int n,
                                             FetchLds &L, Flush flush) {
  if (threadIdx.x == 0) L.misc[1] = static_cast<int>(atomic_add_u32_ret(count, static_cast<unsigned>(n)));
  __syncthreads();
  flush(n, static_cast<unsigned>(L.misc[1]));
  __syncthreads();
}

template <typename Flush>
__device__ __forceinline__ void fetch_scan_staged(const cublox_fetch_cfg &c,
                                                  const float *__restrict__ occ,
                                                  unsigned *__restrict__ count,
                                                  FetchLds &L, Flush flush) {
  FetchSpace f;
  if (!fetch_space(c, f)) return;  // uniform: c is a launch argument
  const unsigned stride = gridDim.x * blockDim.x;
  int n = 0;  // staged samples (the same in every thread)
  for (unsigned b0 = blockIdx.x * blockDim.x; b0 < f.total; b0 += stride) {
    const unsigned e = b0 + threadIdx.x;
    float px = 0.f, py = 0.f, pz = 0.f;
    unsigned char st = 0;
    const bool keep = e < f.total && fetch_classify(c, occ, f, e, px, py, pz, st);
    const int incl = block_scan_incl(L.wave, keep ? 1 : 0);
    const int tot = L.wave[kFetchThreads / 64 - 1];
    if (keep) L.stage[n + incl - 1] = make_float4(px, py, pz, __int_as_float(st));
    n += tot;
    __syncthreads();
    if (n > kFetchStage - kFetchThreads) {
      staged_flush(count, n, L, flush);
      n = 0;
    }
  }
  if (n > 0) staged_flush(count, n, L, flush);
}

// Batched fetch (many queries in a launch, few samples each): one (x, y)
// column of a query's disk per half-wave. The column's window test,
// position, squared planar distance and hash base are formed once, with
// the expressions of fetch_classify; the half-wave's lanes take its z
// voxels 32 at a time, so occ reads are contiguous and no lane divides. A
// column whose squared planar distance already exceeds radius_sq holds no
// kept voxel (adding dz * dz >= 0 cannot round below it) and is skipped.
// Kept samples collect in the wave's LDS stage; the wave reserves their
// slots on the query's count with one atomic when the stage fills and once
// at its end, so no column waits on a global round trip.
constexpr int kWaveStage = 256;  // staged samples per wave

__device__ __forceinline__ void fetch_columns(const cublox_fetch_cfg &c,
                                              const float *__restrict__ occ,
                                              unsigned *__restrict__ count,
                                              float *__restrict__ out_pos,
                                              unsigned char *__restrict__ out_state,
                                              const unsigned cap, const int block,
                                              const int nblocks, float4 *stage) {
  // The enumeration's extents (fetch_space without its divisors).
  FetchSpace f;
  f.r = c.disk_radius_vox;
  f.diam = 2 * f.r + 1;
  f.nz = c.gz_max - c.gz_min + 1;
  if (f.diam <= 0 || f.nz <= 0) return;  // uniform: the query's configuration
  f.ctr = c.origin_at_center != 0;
  auto pmod = [](int v, int m) {
    const int r = v % m;
    return r < 0 ? r + m : r;
  };
  f.qx = pmod(c.origin_i.x, c.map_size_i.x);
  f.qy = pmod(c.origin_i.y, c.map_size_i.y);
  f.qz = pmod(c.origin_i.z, c.map_size_i.z);
  const int lane = __lane_id();
  const int zl = lane & 31;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const int cols = f.diam * f.diam;
  int n = 0;  // staged samples (wave-uniform)
  // Writes the staged samples to slots reserved with one atomic.
  auto flush = [&]() {
    if (n == 0) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    unsigned base = 0;
    if (lane == 0) base = atomic_add_u32_ret(count, static_cast<unsigned>(n));
    base = __shfl(base, 0);
    for (int k = lane; k < n; k += 64) {
      const unsigned slot = base + static_cast<unsigned>(k);
      if (slot >= cap) continue;
      const float4 q = stage[k];
      out_pos[3ull * slot + 0] = q.x;
      out_pos[3ull * slot + 1] = q.y;
      out_pos[3ull * slot + 2] = q.z;
      out_state[slot] = static_cast<unsigned char>(__float_as_int(q.w));
    }
    __builtin_amdgcn_wave_barrier();  // the stage is refilled after this
    n = 0;
  };
  const int waves = static_cast<int>(blockDim.x) >> 6;
  const int step = 2 * nblocks * waves;
  // The lane's column col = tx * diam + ty advances by `step` per pass
  // without a division: step = sq * diam + sr with sr < diam.
  const int col_lane = 2 * (block * waves + (static_cast<int>(threadIdx.x) >> 6)) + (lane >> 5);
  const int sq = step / f.diam;
  const int sr = step - sq * f.diam;
  int tx = col_lane / f.diam;
  int ty = col_lane - tx * f.diam;
  auto advance = [&]() {
    tx += sq;
    ty += sr;
    if (ty >= f.diam) {
      ty -= f.diam;
      ++tx;
    }
  };
  // The z terms of a lane are the same in every column: those of the
  // first 32 z voxels are formed once (deeper columns form the rest).
  struct ZTerms {
    float pz, dz2;
    int hz;
    bool in;
  };
  auto z_terms = [&](int z0) {
    const int gz = c.gz_min + z0 + zl;
    ZTerms t;
    t.pz = global_to_pos(gz, c.resolution, f.ctr);
    const float dz = t.pz - c.center.z;
    t.dz2 = dz * dz;
    t.in = z0 + zl < f.nz && abs(gz - c.origin_i.z) <= c.half_map_size_i.z;
    t.hz = in_window_coord(gz, c.origin_i.z, c.half_map_size_i.z, c.map_size_i.z, f.qz);
    return t;
  };
  const ZTerms z_first = z_terms(0);
  for (int col0 = col_lane - (lane >> 5); col0 < cols; col0 += step, advance()) {  // wave-uniform
    const int col = col0 + (lane >> 5);
    const int gx = c.center_g.x + tx - f.r;
    const int gy = c.center_g.y + ty - f.r;
    const float px = global_to_pos(gx, c.resolution, f.ctr);
    const float py = global_to_pos(gy, c.resolution, f.ctr);
    const float dx = px - c.center.x;
    const float dy = py - c.center.y;
    const float dxy = dx * dx + dy * dy;
    const bool col_in = col < cols && abs(gx - c.origin_i.x) <= c.half_map_size_i.x &&
                        abs(gy - c.origin_i.y) <= c.half_map_size_i.y &&
                        !(dxy > c.radius_sq);
    if (__ballot(col_in) == 0ull) continue;
    const int hxy =
        col_in ? (in_window_coord(gx, c.origin_i.x, c.half_map_size_i.x, c.map_size_i.x, f.qx) *
                      c.map_size_i.y +
                  in_window_coord(gy, c.origin_i.y, c.half_map_size_i.y, c.map_size_i.y, f.qy)) *
                     c.map_size_i.z
               : 0;
    for (int z0 = 0; z0 < f.nz; z0 += 32) {  // wave-uniform
      const ZTerms zt = z0 == 0 ? z_first : z_terms(z0);
      const float pz = zt.pz;
      unsigned char st = 0;
      bool keep = false;
      if (col_in && zt.in) {
        if (!(dxy + zt.dz2 > c.radius_sq)) {
          const int h = hxy + zt.hz;
          const float l = occ[h] + c.l_unknown;
          if (l >= c.l_occupied) {
            st = 2;
          } else if (l < c.l_free) {
            st = 1;
          }
          keep = c.state_filter < 0 || static_cast<int>(st) == c.state_filter;
        }
      }
      const unsigned long long m = __ballot(keep);
      if (m == 0ull) continue;
      if (keep) {
        stage[n + __popcll(m & lt_mask)] =
            make_float4(px, py, pz, __int_as_float(static_cast<int>(st)));
      }
      n += __popcll(m);
      if (n > kWaveStage - 64) flush();  // room for the next 64
    }
  }
  flush();
}

// ===================================================================
// Occupancy viz cloud (the driver's publishOccupancyCloudAsync,
// main.cpp:740-817, over the OCCUPIED fetch of main.cpp:487-491), fused
// into the fetch: a fetched sample (slot below the fetch's max_results and
// capacity) is dropped by the global-index subsample (main.cpp:71-84),
// else coloured by its distance to the robot (main.cpp:275-291), packed
// as a float (main.cpp:293-304) and appended as (x, y, z, rgb).
// ===================================================================

struct VizArgs {
  float px, py, pz;   // robot position (distances are measured from it)
  float d_max;        // colour-map range
  float inv_res;      // grid resolution inverse and binning, for the
  int center;         // subsample's global index
  int step;           // subsample step (<= 1: keep every sample)
  unsigned max_fetch; // the fetch's max_results (0: no limit)
  unsigned fetch_cap; // the fetch's output capacity
};

// main.cpp:275-304; std::clamp order kept (a NaN passes through).
__device__ __forceinline__ float viz_rgb(float d, float d_max) {
  const float dm = d_max < 1.0f ? 1.0f : d_max;  // std::max(d_max, 1.0f)
  const float q = d / dm;
  const float t = q < 0.0f ? 0.0f : (1.0f < q ? 1.0f : q);
  float r, g, b;
  if (t <= 0.5f) {
    const float k = t * 2.0f;
    r = 0.0f;
    g = 1.0f - k;
    b = k;
  } else {
    const float k = (t - 0.5f) * 2.0f;
    r = k;
    g = 0.0f;
    b = 1.0f - k;
  }
  auto u8 = [](float c) {
    const float cc = c < 0.0f ? 0.0f : (1.0f < c ? 1.0f : c);
    return static_cast<unsigned>(cc * 255.0f);
  };
  return __uint_as_float((u8(r) << 16) | (u8(g) << 8) | u8(b));
}

__global__ void __launch_bounds__(kFetchThreads)
    viz_kernel(const cublox_fetch_cfg c, const float *__restrict__ occ,
               const VizArgs v, unsigned *__restrict__ fetch_count,
               unsigned *__restrict__ viz_count, float4 *__restrict__ out,
               const unsigned out_cap) {
  __shared__ FetchLds L;
  const bool ctr = v.center != 0;
  constexpr int kPer = kFetchStage / kFetchThreads;
  fetch_scan_staged(c, occ, fetch_count, L, [&](int n, unsigned base) {
    // Fetched samples (slot below max_results and the capacity) that pass
    // the subsample; this thread's are staged k = t + j * 256.
    unsigned takes = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int k = threadIdx.x + j * kFetchThreads;
      bool take = false;
      if (k < n) {
        const unsigned slot = base + static_cast<unsigned>(k);
        take = !((v.max_fetch > 0u && slot >= v.max_fetch) || slot >= v.fetch_cap);
        if (take && v.step > 1) {
          const float4 q = L.stage[k];
          const unsigned mix = static_cast<unsigned>(pos_to_global(q.x, v.inv_res, ctr)) ^
                               static_cast<unsigned>(pos_to_global(q.y, v.inv_res, ctr)) ^
                               static_cast<unsigned>(pos_to_global(q.z, v.inv_res, ctr));
          take = mix % static_cast<unsigned>(v.step) == 0u;
        }
      }
      takes |= (take ? 1u : 0u) << j;
    }
    const int mine = __popc(takes);
    const int incl = block_scan_incl(L.wave, mine);
    const int total = L.wave[kFetchThreads / 64 - 1];
    if (threadIdx.x == 0 && total > 0) {
      L.misc[2] = static_cast<int>(atomic_add_u32_ret(viz_count, static_cast<unsigned>(total)));
    }
    __syncthreads();
    unsigned idx = static_cast<unsigned>(L.misc[2]) + static_cast<unsigned>(incl - mine);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (!((takes >> j) & 1u)) continue;
      const float4 q = L.stage[threadIdx.x + j * kFetchThreads];
      if (idx < out_cap) {
        const float dx = q.x - v.px, dy = q.y - v.py, dz = q.z - v.pz;
        const float d = sqrtf((dx * dx + dy * dy) + dz * dz);
        out[idx] = make_float4(q.x, q.y, q.z, viz_rgb(d, v.d_max));
      }
      ++idx;
    }
  });
}

__global__ void __launch_bounds__(kFetchThreads)
    fetch_kernel(const cublox_fetch_cfg c, const float *__restrict__ occ,
                 unsigned *__restrict__ count, float *__restrict__ out_pos,
                 unsigned char *__restrict__ out_state, const unsigned cap,
                 const unsigned max_results) {
  __shared__ FetchLds L;
  fetch_scan_staged(c, occ, count, L, [&](int n, unsigned base) {
    for (int k = threadIdx.x; k < n; k += kFetchThreads) {
      const unsigned slot = base + static_cast<unsigned>(k);
      if ((max_results > 0u && slot >= max_results) || slot >= cap) continue;
      const float4 q = L.stage[k];
      out_pos[3ull * slot + 0] = q.x;
      out_pos[3ull * slot + 1] = q.y;
      out_pos[3ull * slot + 2] = q.z;
      out_state[slot] = static_cast<unsigned char>(__float_as_int(q.w));
    }
  });
}

__global__ void __launch_bounds__(256)
    fetch_batch_kernel(const cublox_fetch_cfg *__restrict__ cfgs,
                       const float *__restrict__ occ,
                       unsigned *__restrict__ counts,
                       float *__restrict__ out_pos,
                       unsigned char *__restrict__ out_state,
                       const unsigned per_query_cap) {
  __shared__ float4 s_stage[4][kWaveStage];  // 256 threads: four waves
  const int q = blockIdx.y;
  const cublox_fetch_cfg c = cfgs[q];
  const size_t off = static_cast<size_t>(q) * per_query_cap;
  fetch_columns(c, occ, counts + q, out_pos + 3 * off, out_state + off, per_query_cap,
                blockIdx.x, gridDim.x, s_stage[threadIdx.x >> 6]);
}

constexpr int kMaxGridY = 65535;  // queries per batched launch

// Dense gather of batched fetch results: query q's first min(count, cap)
// samples move from its capacity segment to offsets[q].
__global__ void __launch_bounds__(256)
    fetch_gather_kernel(const unsigned *__restrict__ counts,
                        const unsigned *__restrict__ offsets,
                        const unsigned cap, const float *__restrict__ pos_seg,
                        const unsigned char *__restrict__ st_seg,
                        float *__restrict__ pos_out,
                        unsigned char *__restrict__ st_out) {
  const int q = blockIdx.y;
  const unsigned n = min(counts[q], cap);
  const size_t src = static_cast<size_t>(q) * cap;
  const size_t dst = offsets[q];
  for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += gridDim.x * blockDim.x) {
    pos_out[3 * (dst + j) + 0] = pos_seg[3 * (src + j) + 0];
    pos_out[3 * (dst + j) + 1] = pos_seg[3 * (src + j) + 1];
    pos_out[3 * (dst + j) + 2] = pos_seg[3 * (src + j) + 2];
    st_out[dst + j] = st_seg[src + j];
  }
}

// ===================================================================
// Counter census: nnz(op), sum(op), sum(hit) with wave reductions.
// ===================================================================

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ void __launch_bounds__(256)
    census_kernel(const int *__restrict__ op, const int *__restrict__ hit,
                  const int n, unsigned long long *__restrict__ out) {
  unsigned long long nnz = 0, sop = 0, shit = 0, nhit = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const int o = op[i];
    const int h = hit[i];
    nnz += o != 0;
    sop += static_cast<unsigned long long>(o);
    shit += static_cast<unsigned long long>(h);
    nhit += h != 0;
  }
  nnz = wave_sum_u64(nnz);
  sop = wave_sum_u64(sop);
  shit = wave_sum_u64(shit);
  nhit = wave_sum_u64(nhit);
  if (__lane_id() == 0) {
    __hip_atomic_fetch_add(out + 0, nnz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(out + 1, sop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(out + 2, shit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(out + 3, nhit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ===================================================================
// Host-side recenter plan (reference planRecenterKernel, kernels.cu:313-398)
// ===================================================================

// kernels.cu:303-311
inline int normalize_slice(int x, int a, int b) {
  const int range = b - a + 1;
  int y = (x - a) % range;
  if (y < 0) y += range;
  return y + a;
}

struct HostPlan {
  int action = 0;
  int new_origin[3] = {0, 0, 0};
  std::vector<int> slices[3];
};

HostPlan plan_recenter(const cublox_recenter_cfg &c) {
  HostPlan p;
  const float dx = c.pos.x - c.origin_f.x;
  const float dy = c.pos.y - c.origin_f.y;
  const float dz = c.pos.z - c.origin_f.z;
  if ((dx * dx + dy * dy) + dz * dz < c.recenter_threshold_sq) return p;
  const bool ctr = c.origin_at_center != 0;
  const int no[3] = {pos_to_global(c.pos.x, c.inv_resolution, ctr),
                     pos_to_global(c.pos.y, c.inv_resolution, ctr),
                     pos_to_global(c.pos.z, c.inv_resolution, ctr)};
  const int og[3] = {c.origin_i.x, c.origin_i.y, c.origin_i.z};
  const int ms[3] = {c.map_size_i.x, c.map_size_i.y, c.map_size_i.z};
  const int hs[3] = {c.half_map_size_i.x, c.half_map_size_i.y,
                     c.half_map_size_i.z};
  int max_shift = 0, max_map = 0;
  for (int a = 0; a < 3; ++a) {
    max_shift = std::max(max_shift, std::abs(no[a] - og[a]));
    max_map = std::max(max_map, ms[a]);
    p.new_origin[a] = no[a];
  }
  // kernels.cu:344-347: any shift >= the largest map extent resets all.
  if (max_shift >= max_map) {
    p.action = 2;
    return p;
  }
  p.action = 1;
  for (int a = 0; a < 3; ++a) {
    const int s = no[a] - og[a];
    if (s == 0) continue;
    const int h = hs[a];
    const int min_l = (-h + og[a]) % ms[a];
    std::vector<int> &out = p.slices[a];
    if (s > 0) {
      // Exits on the low side, then the fixed-local "entering" slices
      // (kernels.cu:377-384), which do not depend on the origin.
      for (int k = 0; k < s; ++k) out.push_back(normalize_slice(min_l + k, -h, h));
      for (int k = 0; k < s; ++k) out.push_back(normalize_slice(h - k, -h, h));
    } else {
      for (int k = -1; k >= s; --k) out.push_back(normalize_slice(min_l + k, -h, h));
      for (int k = 0; k < -s; ++k) out.push_back(normalize_slice(-h + k, -h, h));
    }
  }
  return p;
}

constexpr long long kSingleChunkBlocks = 2048;  // one grid's event: ~8 workgroups per CU

// Sorted unique slice values, merged into inclusive [lo, hi] runs.
std::vector<std::pair<int, int>> merge_runs(std::vector<int> v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  std::vector<std::pair<int, int>> runs;
  for (int x : v) {
    if (!runs.empty() && runs.back().second + 1 == x) {
      runs.back().second = x;
    } else {
      runs.emplace_back(x, x);
    }
  }
  return runs;
}

// The zero regions of a slab-clearing recenter plan: merged runs of the
// exit and entering slices per axis (x: contiguous; y: one run per x row;
// z: one run per column).
std::vector<ZeroRegion> recenter_regions(const cublox_recenter_cfg &cfg) {
  const HostPlan p = plan_recenter(cfg);
  const long long mx = cfg.map_size_i.x, my = cfg.map_size_i.y,
                  mz = cfg.map_size_i.z;
  const int hx = cfg.half_map_size_i.x, hy = cfg.half_map_size_i.y,
            hz = cfg.half_map_size_i.z;
  std::vector<ZeroRegion> regions;
  for (const auto &run : merge_runs(p.slices[0])) {
    regions.push_back(ZeroRegion{(run.first + hx) * my * mz, 1, 0,
                                 (run.second - run.first + 1) * my * mz});
  }
  for (const auto &run : merge_runs(p.slices[1])) {
    regions.push_back(ZeroRegion{(run.first + hy) * mz, mx, my * mz,
                                 (run.second - run.first + 1) * mz});
  }
  for (const auto &run : merge_runs(p.slices[2])) {
    regions.push_back(ZeroRegion{run.first + hz, mx * my, mz,
                                 static_cast<long long>(run.second - run.first + 1)});
  }
  return regions;
}

constexpr unsigned kBatchChunksPerRegion = 32;  // batch launch: blocks per region slot

int launch_regions(float *occ, int *op, int *hit,
                   const std::vector<ZeroRegion> &regions, hipStream_t s) {
  for (size_t i0 = 0; i0 < regions.size(); i0 += kMaxRegions) {
    ZeroArgs z{};
    z.n = static_cast<int>(std::min<size_t>(kMaxRegions, regions.size() - i0));
    long long elems = 0;
    for (int k = 0; k < z.n; ++k) elems += regions[i0 + k].n_outer * regions[i0 + k].run;
    const long long piece = chunk_piece(elems, kSingleChunkBlocks);
    long long chunks = 0;
    for (int k = 0; k < z.n; ++k) {
      chunks += chunk_region(regions[i0 + k], piece, chunks, &z.r[k]);
    }
    if (chunks <= 0) continue;
    zero_regions_kernel<<<dim3(static_cast<unsigned>(chunks)), dim3(256), 0, s>>>(
        occ, op, hit, z);
    const int st = check_launch("zero_regions_kernel");
    if (st != CUBLOX_OK) return st;
  }
  return CUBLOX_OK;
}

} // namespace cublox_hip

using namespace cublox_hip;

// Each link of the index chain of one case (utils.cuh:24-166), with the
// device functions the kernels use.
__global__ void __launch_bounds__(256)
    index_chain_kernel(const cublox_index_case *__restrict__ cases, int n,
                       cublox_index_chain *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cublox_index_case c = cases[i];
  const int l[3] = {c.local.x, c.local.y, c.local.z};
  const int m[3] = {c.map_size_i.x, c.map_size_i.y, c.map_size_i.z};
  const int h[3] = {c.half_map_size_i.x, c.half_map_size_i.y, c.half_map_size_i.z};
  const int o[3] = {c.origin_i.x, c.origin_i.y, c.origin_i.z};
  const bool ctr = c.origin_at_center != 0;
  int g[3], lb[3], gp[3];
  float p[3];
  for (int a = 0; a < 3; ++a) {
    g[a] = local_to_global(l[a], m[a], h[a], o[a]);
    lb[a] = global_to_local(g[a], m[a], h[a]);
    p[a] = global_to_pos(g[a], c.resolution, ctr);
    gp[a] = pos_to_global(p[a], c.inv_resolution, ctr);
  }
  // utils.cuh:131-136 on the local index itself (global_to_hash of any
  // global coordinate with these local indices gives the same id).
  const int hid = (l[0] + h[0]) * m[1] * m[2] + (l[1] + h[1]) * m[2] + (l[2] + h[2]);
  int lh[3];
  hash_to_local(hid, m[1], m[2], h[0], h[1], h[2], lh[0], lh[1], lh[2]);
  cublox_index_chain r;
  r.global = cublox_int3{g[0], g[1], g[2]};
  r.local_back = cublox_int3{lb[0], lb[1], lb[2]};
  r.hash_id = hid;
  r.local_from_hash = cublox_int3{lh[0], lh[1], lh[2]};
  r.global_from_hash = cublox_int3{local_to_global(lh[0], m[0], h[0], o[0]),
                                   local_to_global(lh[1], m[1], h[1], o[1]),
                                   local_to_global(lh[2], m[2], h[2], o[2])};
  r.pos = cublox_float3{p[0], p[1], p[2]};
  r.global_from_pos = cublox_int3{gp[0], gp[1], gp[2]};
  out[i] = r;
}

extern "C" {

int cublox_launch_ray_cast_update(const cublox_raycast_cfg *cfg,
                                  const float *d_x, const float *d_y,
                                  const float *d_z, int n, cublox_stream s,
                                  int *d_op_cnt, int *d_hit_cnt) {
  if (!cfg || n <= 0 || !d_x || !d_y || !d_z || !d_op_cnt || !d_hit_cnt) {
    return CUBLOX_OK;  // kernels.cu:223-225: no-op
  }
  const RayArgs a = make_ray_args(*cfg);
  const int blocks = (n + 255) / 256;
  ray_cast_kernel<false><<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
      a, d_x, d_y, d_z, n, d_op_cnt, d_hit_cnt);
  return check_launch("ray_cast_kernel");
}

int cublox_launch_ray_cast_update_sensor(const cublox_raycast_cfg *cfg,
                                         const float T[12], const float *d_xs,
                                         const float *d_ys, const float *d_zs,
                                         int n, cublox_stream s,
                                         int *d_op_cnt, int *d_hit_cnt) {
  if (!cfg || !T || n <= 0 || !d_xs || !d_ys || !d_zs || !d_op_cnt ||
      !d_hit_cnt) {
    return CUBLOX_OK;
  }
  RayArgs a = make_ray_args(*cfg);
  for (int k = 0; k < 12; ++k) a.T[k] = T[k];
  // The sensor origin is the transform's translation (main.cpp:453).
  a.ox = T[3];
  a.oy = T[7];
  a.oz = T[11];
  const int blocks = (n + 255) / 256;
  ray_cast_kernel<true><<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
      a, d_xs, d_ys, d_zs, n, d_op_cnt, d_hit_cnt);
  return check_launch("ray_cast_kernel<transform>");
}

int cublox_launch_apply_update(float *d_occ, int *d_op_cnt, int *d_hit_cnt,
                               int voxel_num, float l_hit, float l_miss,
                               float l_min, float l_max, float l_unknown,
                               cublox_stream s, unsigned int *d_mod_count,
                               int *d_mod_idx, float *d_mod_val,
                               unsigned int mod_cap) {
  if (voxel_num <= 0 || !d_occ || !d_op_cnt || !d_hit_cnt) return CUBLOX_OK;
  const ApplyArgs p{l_hit, l_miss, l_min, l_max, l_unknown};
  // kernels.cu:255-262: record only when the buffers can hold every voxel.
  const bool record = d_mod_count && d_mod_idx && d_mod_val &&
                      static_cast<unsigned>(voxel_num) <= mod_cap;
  const int n4 = voxel_num / 4;
  int blocks = (n4 + 255) / 256;
  blocks = std::max(1, std::min(blocks, 4096));
  if (record) {
    apply_kernel<true><<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
        d_occ, d_op_cnt, d_hit_cnt, voxel_num, p, d_mod_count, d_mod_idx,
        d_mod_val, mod_cap);
  } else {
    apply_kernel<false><<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
        d_occ, d_op_cnt, d_hit_cnt, voxel_num, p, nullptr, nullptr, nullptr,
        0u);
  }
  return check_launch("apply_kernel");
}

int cublox_launch_clear_voxels_by_index(float *d_occ, const int *d_indices,
                                        int n, int voxel_num,
                                        cublox_stream s) {
  if (n <= 0 || !d_occ || !d_indices || voxel_num <= 0) return CUBLOX_OK;
  clear_by_index_kernel<<<dim3((n + 255) / 256), dim3(256), 0,
                          as_stream(s)>>>(d_occ, d_indices, n, voxel_num);
  return check_launch("clear_by_index_kernel");
}

int cublox_plan_recenter(const cublox_recenter_cfg *cfg, int *slices_x,
                         int *slices_y, int *slices_z,
                         cublox_recenter_result *result) {
  if (!cfg || !result) return set_error(CUBLOX_ERR_INVALID, "null argument");
  const HostPlan p = plan_recenter(*cfg);
  result->action = p.action;
  result->new_origin_i = cublox_int3{p.new_origin[0], p.new_origin[1],
                                     p.new_origin[2]};
  int *outs[3] = {slices_x, slices_y, slices_z};
  const long long ms[3] = {cfg->map_size_i.x, cfg->map_size_i.y,
                           cfg->map_size_i.z};
  long long keep = 1;
  for (int a = 0; a < 3; ++a) {
    result->n_slices[a] = static_cast<int>(p.slices[a].size());
    if (outs[a]) std::copy(p.slices[a].begin(), p.slices[a].end(), outs[a]);
    std::vector<int> u = p.slices[a];
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    keep *= ms[a] - static_cast<long long>(u.size());
  }
  const long long V = ms[0] * ms[1] * ms[2];
  result->zeroed_voxels = p.action == 0 ? 0 : p.action == 2 ? V : V - keep;
  return CUBLOX_OK;
}

int cublox_launch_recenter(float *d_occ, int *d_op_cnt, int *d_hit_cnt,
                           const cublox_recenter_cfg *cfg, cublox_stream s,
                           cublox_recenter_result *result) {
  if (!d_occ || !cfg || !result) return CUBLOX_OK;  // kernels.cu:633-636
  const int st = cublox_plan_recenter(cfg, nullptr, nullptr, nullptr, result);
  if (st != CUBLOX_OK) return st;
  if (result->action == 0) return CUBLOX_OK;
  const size_t V = static_cast<size_t>(cfg->map_vox_num);
  hipStream_t hs = as_stream(s);
  if (result->action == 2) {
    // kernels.cu:652-663: full reset.
    int r = cublox_fill_zero_async(d_occ, V, s);
    if (r == CUBLOX_OK && d_op_cnt) r = cublox_fill_zero_async(d_op_cnt, V, s);
    if (r == CUBLOX_OK && d_hit_cnt) r = cublox_fill_zero_async(d_hit_cnt, V, s);
    return r;
  }
  return launch_regions(d_occ, d_op_cnt, d_hit_cnt, recenter_regions(*cfg), hs);
}

size_t cublox_recenter_plan_bytes(void) { return sizeof(StagedClear); }

int cublox_recenter_plan_stage(const cublox_recenter_cfg *cfg, void *plan,
                               cublox_recenter_result *result) {
  if (!cfg || !plan || !result) {
    return set_error(CUBLOX_ERR_INVALID, "recenter_plan_stage: null argument");
  }
  const int st = cublox_plan_recenter(cfg, nullptr, nullptr, nullptr, result);
  if (st != CUBLOX_OK) return st;
  StagedClear *p = static_cast<StagedClear *>(plan);
  p->action = result->action;
  p->n = 0;
  p->voxels = cfg->map_vox_num;
  if (result->action != 1) return CUBLOX_OK;
  const std::vector<ZeroRegion> regions = recenter_regions(*cfg);
  if (regions.size() > static_cast<size_t>(kStagedRegions)) {
    // Merged runs of at most two slice ranges per axis: never this many.
    p->action = 2;
    return CUBLOX_OK;
  }
  p->n = static_cast<int>(regions.size());
  for (int k = 0; k < p->n; ++k) p->r[k] = regions[static_cast<size_t>(k)];
  return CUBLOX_OK;
}

int cublox_launch_recenter_plan(const void *plan, float *d_occ, int *d_op_cnt,
                                int *d_hit_cnt, cublox_stream s) {
  if (!plan || !d_occ) return CUBLOX_OK;
  const StagedClear *p = static_cast<const StagedClear *>(plan);
  if (p->action == 2) {
    const size_t V = static_cast<size_t>(p->voxels);
    int r = cublox_fill_zero_async(d_occ, V, s);
    if (r == CUBLOX_OK && d_op_cnt) r = cublox_fill_zero_async(d_op_cnt, V, s);
    if (r == CUBLOX_OK && d_hit_cnt) r = cublox_fill_zero_async(d_hit_cnt, V, s);
    return r;
  }
  if (p->action != 1 || p->n <= 0) return CUBLOX_OK;
  return launch_regions(d_occ, d_op_cnt, d_hit_cnt,
                        std::vector<ZeroRegion>(p->r, p->r + p->n), as_stream(s));
}

size_t cublox_recenter_batch_staging_bytes(int max_regions) {
  static_assert(sizeof(RecenterBatchHeader) == 32, "header size");
  return sizeof(RecenterBatchHeader) +
         static_cast<size_t>(std::max(max_regions, 1)) * sizeof(ZeroRegionP);
}

int cublox_recenter_batch_stage(const cublox_recenter_item *items, int n,
                                int max_regions, void *h_staging,
                                cublox_recenter_result *results) {
  if (n < 0 || (n > 0 && (!items || !results)) || !h_staging) {
    return set_error(CUBLOX_ERR_INVALID, "recenter_batch: bad arguments");
  }
  auto *hdr = static_cast<RecenterBatchHeader *>(h_staging);
  auto *regs = reinterpret_cast<ZeroRegionP *>(hdr + 1);
  std::vector<std::pair<float *, ZeroRegion>> all;
  for (int k = 0; k < n; ++k) {
    const cublox_recenter_item &it = items[k];
    results[k] = cublox_recenter_result{};
    if (!it.d_occ) continue;
    const int st = cublox_plan_recenter(&it.cfg, nullptr, nullptr, nullptr,
                                        results + k);
    if (st != CUBLOX_OK) return st;
    if (results[k].action == 0) continue;
    std::vector<ZeroRegion> rs;
    if (results[k].action == 2) {
      rs.push_back(ZeroRegion{0, 1, 0, static_cast<long long>(it.cfg.map_vox_num)});
    } else {
      rs = recenter_regions(it.cfg);
    }
    for (const ZeroRegion &r : rs) {
      if (static_cast<int>(all.size()) >= max_regions) {
        return set_error(CUBLOX_ERR_CAPACITY, "recenter_batch: too many regions");
      }
      all.emplace_back(it.d_occ, r);
    }
  }
  long long elems = 0;
  for (const auto &pr : all) elems += pr.second.n_outer * pr.second.run;
  const long long piece =
      chunk_piece(elems, static_cast<long long>(max_regions) * kBatchChunksPerRegion);
  long long chunks = 0;
  for (size_t k = 0; k < all.size(); ++k) {
    regs[k].occ = all[k].first;
    chunks += chunk_region(all[k].second, piece, chunks, &regs[k].c);
  }
  *hdr = RecenterBatchHeader{static_cast<int>(all.size()), 0, chunks, {0, 0, 0, 0}};
  return CUBLOX_OK;
}

int cublox_recenter_batch_upload(const void *h_staging, void *d_area, int max_regions,
                                 cublox_stream s) {
  if (!h_staging || !d_area || max_regions <= 0) {
    return set_error(CUBLOX_ERR_INVALID, "recenter_batch: bad arguments");
  }
  return check(hipMemcpyAsync(d_area, h_staging,
                              cublox_recenter_batch_staging_bytes(max_regions),
                              hipMemcpyHostToDevice, as_stream(s)),
               "hipMemcpyAsync(recenter batch)");
}

int cublox_launch_recenter_batch_uploaded(const void *d_area, int max_regions,
                                          cublox_stream s) {
  if (!d_area || max_regions <= 0) {
    return set_error(CUBLOX_ERR_INVALID, "recenter_batch: bad arguments");
  }
  const auto *hdr = static_cast<const RecenterBatchHeader *>(d_area);
  zero_regions_batch_kernel<<<dim3(kBatchChunksPerRegion * static_cast<unsigned>(max_regions)),
                              dim3(256), 0, as_stream(s)>>>(
      hdr, reinterpret_cast<const ZeroRegionP *>(hdr + 1));
  return check_launch("zero_regions_batch_kernel");
}

int cublox_launch_recenter_batch_staged(const void *h_staging, void *d_area,
                                        int max_regions, cublox_stream s) {
  const int st = cublox_recenter_batch_upload(h_staging, d_area, max_regions, s);
  if (st != CUBLOX_OK) return st;
  return cublox_launch_recenter_batch_uploaded(d_area, max_regions, s);
}

int cublox_launch_clear_recenter_slabs_for_axis(
    float *d_occ, int *d_op_cnt, int *d_hit_cnt, cublox_int3 ms,
    cublox_int3 hs, const int *d_slices, int n_slices, int axis,
    cublox_stream s) {
  if (n_slices <= 0 || !d_occ || !d_slices) return CUBLOX_OK;  // :599-601
  if (axis < 0 || axis > 2) return CUBLOX_OK;
  const long long work = axis == 0   ? (static_cast<long long>(ms.y) * ms.z + 3) / 4
                         : axis == 1 ? static_cast<long long>(ms.x) * ms.z
                                     : static_cast<long long>(ms.x) * ms.y;
  long long bx = (work + 255) / 256;
  bx = std::max<long long>(1, std::min<long long>(bx, 1024));
  clear_slabs_list_kernel<<<dim3(static_cast<unsigned>(bx),
                                 static_cast<unsigned>(n_slices)),
                            dim3(256), 0, as_stream(s)>>>(
      d_occ, d_op_cnt, d_hit_cnt, ms.x, ms.y, ms.z, hs.x, hs.y, hs.z, d_slices,
      axis);
  return check_launch("clear_slabs_list_kernel");
}

int cublox_fetch_plan(cublox_int3 origin_i, cublox_int3 map_size_i,
                      cublox_int3 half, float res, float inv_res, int center,
                      float l_free, float l_occupied, float l_unknown,
                      cublox_float3 c, float radius, int state_filter,
                      int max_results, cublox_fetch_cfg *cfg,
                      long long *est_voxels, unsigned int *output_cap) {
  if (!cfg) return 0;
  if (!(radius > 0.0f)) return 0;
  const bool ctr = center != 0;
  // Grid::maxHorizontalInWindowRadius (Grid.cpp:67-78)
  const int cgx = pos_to_global(c.x, inv_res, ctr);
  const int cgy = pos_to_global(c.y, inv_res, ctr);
  const int cgz = pos_to_global(c.z, inv_res, ctr);
  const int mgx = half.x - std::abs(cgx - origin_i.x);
  const int mgy = half.y - std::abs(cgy - origin_i.y);
  float margin = 0.0f;
  if (mgx >= 0 && mgy >= 0) margin = static_cast<float>(std::min(mgx, mgy)) * res;
  // OccupancyGrid.cpp:248-290
  const float r = std::min(radius, margin);
  if (!(r > 0.0f)) return 0;
  int gmin[3] = {pos_to_global(c.x - r, inv_res, ctr),
                 pos_to_global(c.y - r, inv_res, ctr),
                 pos_to_global(c.z - r, inv_res, ctr)};
  int gmax[3] = {pos_to_global(c.x + r, inv_res, ctr),
                 pos_to_global(c.y + r, inv_res, ctr),
                 pos_to_global(c.z + r, inv_res, ctr)};
  const int o[3] = {origin_i.x, origin_i.y, origin_i.z};
  const int h[3] = {half.x, half.y, half.z};
  for (int a = 0; a < 3; ++a) {
    gmin[a] = std::max(gmin[a], o[a] - h[a]);
    gmax[a] = std::min(gmax[a], o[a] + h[a]);
    if (gmax[a] < gmin[a]) return 0;
  }
  const int disk_r = static_cast<int>(std::ceil(static_cast<double>(r * inv_res)));
  const int nz = gmax[2] - gmin[2] + 1;
  const int diam = 2 * disk_r + 1;
  const long long est = static_cast<long long>(diam) * diam * nz;
  if (est <= 0 || est > static_cast<long long>(INT_MAX)) return 0;
  cfg->center_g = cublox_int3{cgx, cgy, cgz};
  cfg->origin_i = origin_i;
  cfg->map_size_i = map_size_i;
  cfg->half_map_size_i = half;
  cfg->gz_min = gmin[2];
  cfg->gz_max = gmax[2];
  cfg->disk_radius_vox = disk_r;
  cfg->center = c;
  cfg->radius_sq = r * r;
  cfg->resolution = res;
  cfg->l_free = l_free;
  cfg->l_occupied = l_occupied;
  cfg->l_unknown = l_unknown;
  cfg->origin_at_center = center;
  cfg->state_filter = state_filter;
  if (est_voxels) *est_voxels = est;
  if (output_cap) {
    *output_cap = max_results > 0 ? static_cast<unsigned>(max_results) : 500000u;
  }
  return 1;
}

int cublox_launch_fetch_occupancy_around(const cublox_fetch_cfg *cfg,
                                         const float *d_occ,
                                         unsigned int *d_count,
                                         float *d_pos, unsigned char *d_state,
                                         unsigned int cap,
                                         unsigned int max_results,
                                         cublox_stream s) {
  if (!cfg || cfg->disk_radius_vox < 0 || cfg->gz_min > cfg->gz_max ||
      !d_occ || !d_count || !d_pos || !d_state || cap == 0) {
    return CUBLOX_OK;  // kernels.cu:570-574
  }
  const long long diam = 2LL * cfg->disk_radius_vox + 1;
  const long long total = diam * diam * (cfg->gz_max - cfg->gz_min + 1);
  if (total > INT_MAX) {
    return set_error(CUBLOX_ERR_INVALID, "fetch: more than 2^31 - 1 voxels to test");
  }
  long long blocks = (total + 255) / 256;
  blocks = std::max<long long>(1, std::min<long long>(blocks, 4096));
  fetch_kernel<<<dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                 as_stream(s)>>>(*cfg, d_occ, d_count, d_pos, d_state, cap,
                                 max_results);
  return check_launch("fetch_kernel");
}

int cublox_launch_viz_cloud(const cublox_fetch_cfg *cfg, const float *d_occ,
                            const cublox_viz_cfg *viz,
                            unsigned int fetch_capacity,
                            unsigned int *d_fetch_count,
                            unsigned int *d_viz_count, float *d_out,
                            unsigned int out_capacity, cublox_stream s) {
  if (!cfg || !viz || !d_occ || !d_fetch_count || !d_viz_count || !d_out) {
    return set_error(CUBLOX_ERR_INVALID, "viz_cloud: null argument");
  }
  int st = cublox_fill_zero_async(d_fetch_count, 1, s);
  if (st == CUBLOX_OK) st = cublox_fill_zero_async(d_viz_count, 1, s);
  if (st != CUBLOX_OK) return st;
  if (cfg->disk_radius_vox < 0 || cfg->gz_min > cfg->gz_max || fetch_capacity == 0) {
    return CUBLOX_OK;  // kernels.cu:570-574: nothing fetched
  }
  VizArgs v{};
  v.px = viz->robot_pos.x;
  v.py = viz->robot_pos.y;
  v.pz = viz->robot_pos.z;
  v.d_max = viz->max_viz_radius;
  v.inv_res = viz->inv_resolution;
  v.center = viz->origin_at_center;
  v.step = viz->subsample;
  v.max_fetch = viz->max_points > 0 ? static_cast<unsigned>(viz->max_points) : 0u;
  v.fetch_cap = fetch_capacity;
  const long long diam = 2LL * cfg->disk_radius_vox + 1;
  const long long total = diam * diam * (cfg->gz_max - cfg->gz_min + 1);
  if (total > INT_MAX) {
    return set_error(CUBLOX_ERR_INVALID, "viz_cloud: more than 2^31 - 1 voxels to test");
  }
  long long blocks = (total + 255) / 256;
  blocks = std::max<long long>(1, std::min<long long>(blocks, 4096));
  viz_kernel<<<dim3(static_cast<unsigned>(blocks)), dim3(256), 0, as_stream(s)>>>(
      *cfg, d_occ, v, d_fetch_count, d_viz_count, reinterpret_cast<float4 *>(d_out),
      out_capacity);
  return check_launch("viz_kernel");
}

int cublox_launch_fetch_occupancy_batch(const cublox_fetch_cfg *d_cfgs,
                                        int n_queries, int max_disk_diam,
                                        int max_nz, const float *d_occ,
                                        unsigned int *d_counts,
                                        float *d_pos,
                                        unsigned char *d_state,
                                        unsigned int per_query_cap,
                                        cublox_stream s) {
  if (n_queries <= 0 || !d_cfgs || !d_occ || !d_counts || !d_pos ||
      !d_state || per_query_cap == 0 || max_disk_diam <= 0 || max_nz <= 0) {
    return CUBLOX_OK;
  }
  const long long total =
      static_cast<long long>(max_disk_diam) * max_disk_diam * max_nz;
  if (total > INT_MAX) {
    return set_error(CUBLOX_ERR_INVALID, "fetch batch: more than 2^31 - 1 voxels per query");
  }
  // A few workgroups per query (eight disk columns per pass of a 256-thread
  // workgroup), each striding over the columns about sixteen times: a
  // workgroup per few hundred voxels made the dispatch of tens of
  // thousands of short workgroups the cost. Queries go in launches of at
  // most 65,535 (the grid's y extent).
  const long long cols = static_cast<long long>(max_disk_diam) * max_disk_diam;
  long long blocks = (cols + 127) / 128;
  blocks = std::max<long long>(1, std::min<long long>(blocks, 64));
  for (int q0 = 0; q0 < n_queries; q0 += kMaxGridY) {
    const int nq = std::min(kMaxGridY, n_queries - q0);
    const size_t off = static_cast<size_t>(q0) * per_query_cap;
    fetch_batch_kernel<<<dim3(static_cast<unsigned>(blocks), static_cast<unsigned>(nq)),
                         dim3(256), 0, as_stream(s)>>>(
        d_cfgs + q0, d_occ, d_counts + q0, d_pos + 3 * off, d_state + off, per_query_cap);
    const int st = check_launch("fetch_batch_kernel");
    if (st != CUBLOX_OK) return st;
  }
  return CUBLOX_OK;
}

int cublox_launch_fetch_gather(const unsigned int *d_counts,
                               const unsigned int *d_offsets, int n_queries,
                               unsigned int per_query_cap,
                               unsigned int max_count, const float *d_pos_seg,
                               const unsigned char *d_state_seg,
                               float *d_pos_out, unsigned char *d_state_out,
                               cublox_stream s) {
  if (n_queries <= 0 || per_query_cap == 0 || max_count == 0) return CUBLOX_OK;
  const unsigned m = std::min(max_count, per_query_cap);
  const unsigned bx = std::max(1u, std::min((m + 255u) / 256u, 64u));
  for (int q0 = 0; q0 < n_queries; q0 += kMaxGridY) {
    const int nq = std::min(kMaxGridY, n_queries - q0);
    const size_t off = static_cast<size_t>(q0) * per_query_cap;
    fetch_gather_kernel<<<dim3(bx, static_cast<unsigned>(nq)), dim3(256), 0,
                          as_stream(s)>>>(d_counts + q0, d_offsets + q0, per_query_cap,
                                          d_pos_seg + 3 * off, d_state_seg + off,
                                          d_pos_out, d_state_out);
    const int st = check_launch("fetch_gather_kernel");
    if (st != CUBLOX_OK) return st;
  }
  return CUBLOX_OK;
}

int cublox_counter_census(const int *d_op, const int *d_hit, int n,
                          unsigned long long *d_out, cublox_stream s) {
  if (n <= 0 || !d_op || !d_hit || !d_out) return CUBLOX_OK;
  int blocks = std::max(1, std::min((n + 255) / 256, 2048));
  census_kernel<<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(d_op, d_hit, n, d_out);
  return check_launch("census_kernel");
}

int cublox_launch_index_chain(const cublox_index_case *d_cases, int n,
                              cublox_index_chain *d_out, cublox_stream s) {
  if (n <= 0 || !d_cases || !d_out) return CUBLOX_OK;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  index_chain_kernel<<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(d_cases, n, d_out);
  return check_launch("index_chain_kernel");
}

} // extern "C"
