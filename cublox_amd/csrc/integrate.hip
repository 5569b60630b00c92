// Tile-binned fused scan update: ray cast + clamped log-odds fold straight
// into the occupancy buffer, without per-voxel global counters.
//
// Same results as the reference's raycast-then-apply pair (kernels.cu:33-210)
// — the same per-voxel op/hit counts folded once per scan — organised for
// CDNA4, where device-scope atomics are the scarce resource (measured on
// MI355X: ~27 G/s scattered, ~90 M/s on one address) and LDS is plentiful:
//
//   walk        one ray per lane. Steps < K (the near field: step s visits a
//               voxel at L1 distance s from the sensor voxel) count into an
//               LDS cube around the sensor voxel, flushed once per
//               workgroup. Beyond, the walk is cut into "segments" at every
//               entry into a 16x16x32 tile of the hash-ordered grid: each
//               segment carries the exact DDA state at the tile entry plus
//               the ray's constants and goes to the ray's own slot range.
//   bin_count   per workgroup, an LDS hash table aggregates its segments by
//               tile; one global atomic per (workgroup, tile).
//   tile_scan   one workgroup: tile offsets, and the work items — one per
//               tile, or one per 1024-segment chunk of a hot tile — ordered
//               by decreasing cost (log2 classes).
//   bin_place   the same aggregation; one returning atomic per (workgroup,
//               tile) reserves a range; segments are copied to tile order.
//   tile_apply  one workgroup per work item: resume every segment inside
//               the tile with packed 16+16-bit LDS counters (32 KB), then
//               fold the tile's counts into occ with all loads in flight.
//               Chunks of a hot tile store their counts to scratch slabs.
//   hot_fold    sums each split tile's slabs and folds, 8 workgroups a tile.
//   near_apply  fold the near-field cube into occ and clear it.
//
// The walk's t_max recurrence is a running FP32 sum, so a segment resumed
// from its carried state is bit-identical to the uninterrupted walk. Every
// voxel's visits happen at one L1 distance from the sensor voxel, so the
// near-field cube and the tiles own disjoint voxel sets and each voxel is
// folded exactly once.
#include "raycast_common.hpp"

#include <algorithm>

namespace cublox_hip {

constexpr int TX = 16, TY = 16, TZ = 32;
constexpr int kTileVox = TX * TY * TZ;  // 8192
constexpr int kWalkBlock = 256;
constexpr int kMaxNearK = 12;
constexpr int kMaxCube = 2 * kMaxNearK - 1;  // 23
constexpr int kMaxCubeCells = kMaxCube * kMaxCube * kMaxCube;
constexpr int kBinRays = 128;
constexpr int kHash = 4096;
constexpr int kHashProbes = 64;
constexpr int kChunk = 1024;      // records per chunk of a hot tile
constexpr int kMaxSlabs = 2048;   // scratch slabs (chunk counts), 32 KB each
constexpr int kMaxSplit = kMaxSlabs / 2;  // a split tile has >= 2 chunks
constexpr int kPass = 65535;      // records per LDS pass: 16-bit counts
constexpr int kApplyThreads = 256;
constexpr int kPerThread = kTileVox / kApplyThreads;  // 32 voxels
constexpr int kHotSlices = 8;     // hot_fold workgroups per split tile
constexpr int kCostBuckets = 18;  // log2 classes of item cost

struct Segment {  // 64 B, self-contained; step < 0 marks an endpoint record
  int step;
  int ix, iy, iz;
  float tMx, tMy, tMz;
  float tDx, tDy, tDz;
  int eix, eiy, eiz;
  int max_steps;
  int flags;  // bits 0..5: sx+1, sy+1, sz+1 (2 bits each); bit 6: clipped
  int tile;
};

struct Item {  // one tile_apply work unit: records [beg, end) of a tile
  int tile, beg, end;
  int slab;  // < 0: accumulate and fold; else a hot-tile chunk's slab
};

struct HotTile {  // a tile split into chunks, folded by hot_fold
  int tile, slab0, chunks, pad;
};

struct IntegrateArgs {
  RayArgs a;
  int n;
  int ntx, nty, ntz, n_tiles;
  int near_k;   // K
  int cube;     // 2K - 1
  int max_seg;  // per-ray segment slots
  ApplyArgs p;
};

struct Workspace {
  int *seg_count;     // n: segments written by each ray
  Segment *segs;      // n * max_seg, per-ray slot ranges
  Segment *sorted;    // n * max_seg, tile order
  int *tile_count;    // n_tiles (zero between launches)
  int *tile_offset;   // n_tiles + 1
  int *tile_cursor;   // n_tiles
  Item *items;        // max_items, costliest first
  HotTile *hot;       // kMaxSplit
  unsigned *scratch;  // kMaxSlabs * kTileVox: chunk counts, op | hit << 16
  int *misc;          // [0] items, [1] segments, [2] overflow flag, [3] split tiles
  int *near_op;       // cube^3 (zero between launches)
  int *near_hit;      // cube^3 (zero between launches)
  long long seg_capacity;
  int max_items;
};

__device__ __forceinline__ int tile_of(const IntegrateArgs &g, int cx, int cy,
                                       int cz) {
  return ((cx / TX) * g.nty + (cy / TY)) * g.ntz + (cz / TZ);
}

__device__ __forceinline__ int cube_index(const IntegrateArgs &g, int dx,
                                          int dy, int dz) {
  const int k1 = g.near_k - 1;
  return ((dx + k1) * g.cube + (dy + k1)) * g.cube + (dz + k1);
}

// ----------------------------------------------------------------------
// walk
// ----------------------------------------------------------------------
template <bool kTransform>
__global__ void __launch_bounds__(kWalkBlock)
    walk_kernel(const IntegrateArgs g, const float *__restrict__ xs,
                const float *__restrict__ ys, const float *__restrict__ zs,
                Workspace w) {
  __shared__ int s_near[kMaxCubeCells];
  const int cells = g.cube * g.cube * g.cube;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) s_near[c] = 0;
  __syncthreads();

  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const RayArgs &a = g.a;
  RayState r;
  const bool valid = i < g.n && ray_setup<kTransform>(a, xs, ys, zs, i, r);
  if (valid) {
    const int flags = (r.sx + 1) | ((r.sy + 1) << 2) | ((r.sz + 1) << 4) |
                      (r.clipped ? 64 : 0);
    Segment *__restrict__ mine = w.segs + static_cast<size_t>(i) * g.max_seg;
    int nseg = 0;
    const int s0x = r.ix, s0y = r.iy, s0z = r.iz;  // sensor voxel
    HashCoord hc = hash_coord_of(a, r.ix, r.iy, r.iz);
    int cur = -1;
    const int K = g.near_k;
    for (int step = 0; step < r.max_steps; ++step) {
      if (r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) break;
      const bool inwin = abs(r.ix - a.oix) <= a.hx &&
                         abs(r.iy - a.oiy) <= a.hy && abs(r.iz - a.oiz) <= a.hz;
      if (step < K) {
        if (inwin) {
          atomicAdd(s_near + cube_index(g, r.ix - s0x, r.iy - s0y, r.iz - s0z), 1);
        }
      } else {
        const int t = inwin ? tile_of(g, hc.cx, hc.cy, hc.cz) : -1;
        if (t != cur && t >= 0) {
          if (nseg < g.max_seg) {
            Segment sg;
            sg.step = step;
            sg.ix = r.ix;
            sg.iy = r.iy;
            sg.iz = r.iz;
            sg.tMx = r.tMx;
            sg.tMy = r.tMy;
            sg.tMz = r.tMz;
            sg.tDx = r.tDx;
            sg.tDy = r.tDy;
            sg.tDz = r.tDz;
            sg.eix = r.eix;
            sg.eiy = r.eiy;
            sg.eiz = r.eiz;
            sg.max_steps = r.max_steps;
            sg.flags = flags;
            sg.tile = t;
            mine[nseg] = sg;
          } else {
            w.misc[2] = 1;
          }
          ++nseg;
        }
        cur = t;
      }
      const int ax = ray_advance(r);
      hash_coord_step(hc, ax, r, a);
    }
    // Endpoint (kernels.cu:141-151): op += 1, hit += 1 unless clipped.
    if (inside_window(r.eix, r.eiy, r.eiz, a.oix, a.oiy, a.oiz, a.hx, a.hy,
                      a.hz)) {
      const int ddx = r.eix - s0x, ddy = r.eiy - s0y, ddz = r.eiz - s0z;
      if (abs(ddx) + abs(ddy) + abs(ddz) < K) {
        const int c = cube_index(g, ddx, ddy, ddz);
        atomicAdd(s_near + c, 1);
        if (!r.clipped) atomic_add_i32(w.near_hit + c, 1);
      } else {
        const HashCoord e = hash_coord_of(a, r.eix, r.eiy, r.eiz);
        if (nseg < g.max_seg) {
          Segment sg;
          sg.step = -1;
          sg.ix = r.eix;
          sg.iy = r.eiy;
          sg.iz = r.eiz;
          sg.flags = flags;
          sg.tile = tile_of(g, e.cx, e.cy, e.cz);
          mine[nseg] = sg;
        } else {
          w.misc[2] = 1;
        }
        ++nseg;
      }
    }
    w.seg_count[i] = min(nseg, g.max_seg);
  } else if (i < g.n) {
    w.seg_count[i] = 0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < cells; c += blockDim.x) {
    const int v = s_near[c];
    if (v) atomic_add_i32(w.near_op + c, v);
  }
}

// ----------------------------------------------------------------------
// binning: LDS hash aggregation of segments by tile
// ----------------------------------------------------------------------

// Finds (inserting if needed) the LDS slot of `key`; -1 when the bounded
// probe finds neither. A key is either in the table for every caller or
// for none: slots only fill, so a full probe window stays full.
__device__ __forceinline__ int lds_hash_slot(int *keys, int key, bool insert) {
  unsigned h = (static_cast<unsigned>(key) * 2654435761u) >> 20;
  for (int p = 0; p < kHashProbes; ++p, ++h) {
    h &= kHash - 1;
    const int k = keys[h];
    if (k == key) return static_cast<int>(h);
    if (k == -1) {
      if (!insert) return -1;
      const int prev = atomicCAS(keys + h, -1, key);
      if (prev == -1 || prev == key) return static_cast<int>(h);
    }
  }
  return -1;
}

template <bool kPlace>
__global__ void __launch_bounds__(256)
    bin_kernel(const IntegrateArgs g, Workspace w) {
  __shared__ int s_key[kHash];
  __shared__ int s_cnt[kHash];
  __shared__ int s_base[kHash];
  for (int e = threadIdx.x; e < kHash; e += blockDim.x) {
    s_key[e] = -1;
    s_cnt[e] = 0;
  }
  __syncthreads();
  const int r0 = blockIdx.x * kBinRays;
  const int lane_ray = threadIdx.x % kBinRays;
  const int j0 = threadIdx.x / kBinRays;
  constexpr int kJStride = 256 / kBinRays;
  const int ray = r0 + lane_ray;
  const int cnt = ray < g.n ? w.seg_count[ray] : 0;
  const size_t base = static_cast<size_t>(ray) * g.max_seg;

  for (int j = j0; j < cnt; j += kJStride) {
    const int t = w.segs[base + j].tile;
    const int e = lds_hash_slot(s_key, t, true);
    if (e >= 0) {
      atomicAdd(s_cnt + e, 1);
    } else if (!kPlace) {
      atomic_add_i32(w.tile_count + t, 1);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kHash; e += blockDim.x) {
    const int key = s_key[e];
    if (key < 0) continue;
    if (!kPlace) {
      atomic_add_i32(w.tile_count + key, s_cnt[e]);
    } else {
      s_base[e] = atomicAdd(w.tile_cursor + key, s_cnt[e]);
      s_cnt[e] = 0;
    }
  }
  if (!kPlace) return;
  __syncthreads();
  for (int j = j0; j < cnt; j += kJStride) {
    const Segment sg = w.segs[base + j];
    const int e = lds_hash_slot(s_key, sg.tile, false);
    const int slot = e >= 0 ? s_base[e] + atomicAdd(s_cnt + e, 1)
                            : atomicAdd(w.tile_cursor + sg.tile, 1);
    if (slot < w.seg_capacity) w.sorted[slot] = sg;
  }
}

// ----------------------------------------------------------------------
// tile_scan: offsets, work items, hot-tile splits
// ----------------------------------------------------------------------
constexpr int kScanThreads = 1024;

// Inclusive block scan of one int per thread (Hillis-Steele).
__device__ __forceinline__ int block_scan_incl(int *s, int v) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int a = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  const int r = s[t];
  __syncthreads();
  return r;
}

// Cost class of a work item (its record count), larger classes first.
__device__ __forceinline__ int cost_bucket(int records) {
  const int lg = 31 - __clz(max(records, 1));
  return kCostBuckets - 1 - min(lg, kCostBuckets - 1);
}

// Number of chunks a tile with c records is split into (1: not split).
__device__ __forceinline__ int chunks_of(int c) {
  return c > kChunk ? (c + kChunk - 1) / kChunk : 1;
}

__global__ void __launch_bounds__(kScanThreads)
    tile_scan_kernel(const int n_tiles, Workspace w) {
  __shared__ int s_buf[kScanThreads];
  __shared__ int s_bucket[kCostBuckets];
  __shared__ int s_cursor[kCostBuckets];
  const int t = threadIdx.x;
  if (t < kCostBuckets) s_bucket[t] = 0;
  const int per = (n_tiles + kScanThreads - 1) / kScanThreads;
  const int beg = min(t * per, n_tiles);
  const int end = min(beg + per, n_tiles);
  // Pass 1: records and the scratch slabs that splitting would take.
  int seg = 0, slabs = 0;
  for (int k = beg; k < end; ++k) {
    const int c = w.tile_count[k];
    seg += c;
    if (c > kChunk) slabs += chunks_of(c);
  }
  const int seg_incl = block_scan_incl(s_buf, seg);
  const int slab_incl = block_scan_incl(s_buf, slabs);
  // A hot tile is split when its whole slab range fits the scratch area (a
  // prefix of the hot tiles in tile order); otherwise one workgroup takes
  // it in 16-bit passes. Pass 2: work items per cost class.
  int sl = slab_incl - slabs;
  int hot = 0;
  for (int k = beg; k < end; ++k) {
    const int c = w.tile_count[k];
    if (c <= 0) continue;
    const int ch = chunks_of(c);
    if (ch > 1 && sl + ch <= kMaxSlabs) {
      atomicAdd(s_bucket + cost_bucket((c + ch - 1) / ch), ch);
      ++hot;
    } else {
      atomicAdd(s_bucket + cost_bucket(c), 1);
    }
    if (ch > 1) sl += ch;
  }
  const int hot_incl = block_scan_incl(s_buf, hot);
  if (t == 0) {
    int run = 0;
    for (int b = 0; b < kCostBuckets; ++b) {
      s_cursor[b] = run;
      run += s_bucket[b];
    }
    w.misc[0] = min(run, w.max_items);
  }
  __syncthreads();
  // Pass 3: offsets, cursors, items, split tiles; clear the counts.
  int run = seg_incl - seg;
  int sp = hot_incl - hot;
  sl = slab_incl - slabs;
  for (int k = beg; k < end; ++k) {
    const int c = w.tile_count[k];
    w.tile_offset[k] = run;
    w.tile_cursor[k] = run;
    if (c > 0) {
      const int ch = chunks_of(c);
      const bool split = ch > 1 && sl + ch <= kMaxSlabs;
      const int chunks = split ? ch : 1;
      const int per_chunk = (c + chunks - 1) / chunks;
      if (split) w.hot[sp++] = HotTile{k, sl, ch, 0};
      // Every chunk is classed by per_chunk, as counted in pass 2.
      const int bucket = cost_bucket(per_chunk);
      for (int q = 0; q < chunks; ++q) {
        const int ib = run + q * per_chunk;
        const int ie = min(ib + per_chunk, run + c);
        const int pos = atomicAdd(s_cursor + bucket, 1);
        if (pos < w.max_items) {
          w.items[pos] = Item{k, ib, ie, split ? sl + q : -1};
        }
      }
      if (ch > 1) sl += ch;
    }
    run += c;
    w.tile_count[k] = 0;
  }
  if (t == kScanThreads - 1) {
    w.tile_offset[n_tiles] = seg_incl;
    w.misc[1] = seg_incl;
    w.misc[3] = hot_incl;
  }
}

// ----------------------------------------------------------------------
// tile_apply
// ----------------------------------------------------------------------
// Accumulates records [beg, end) of a tile into packed LDS counters
// (op | hit << 16). A pass holds at most kPass records and a record visits
// a voxel at most once, so neither field carries into the other.
__device__ __forceinline__ void tile_accumulate(const IntegrateArgs &g,
                                                const Workspace &w, int beg,
                                                int end, int x0, int y0,
                                                int z0, unsigned *s_cnt) {
  const RayArgs &a = g.a;
  for (int k = beg + static_cast<int>(threadIdx.x); k < end; k += kApplyThreads) {
    const Segment sg = w.sorted[k];
    if (sg.step < 0) {
      const HashCoord e = hash_coord_of(a, sg.ix, sg.iy, sg.iz);
      const int li = ((e.cx - x0) * TY + (e.cy - y0)) * TZ + (e.cz - z0);
      atomicAdd(s_cnt + li, (sg.flags & 64) ? 1u : 0x10001u);
      continue;
    }
    RayState r;
    r.ix = sg.ix;
    r.iy = sg.iy;
    r.iz = sg.iz;
    r.tMx = sg.tMx;
    r.tMy = sg.tMy;
    r.tMz = sg.tMz;
    r.tDx = sg.tDx;
    r.tDy = sg.tDy;
    r.tDz = sg.tDz;
    r.eix = sg.eix;
    r.eiy = sg.eiy;
    r.eiz = sg.eiz;
    r.max_steps = sg.max_steps;
    r.sx = (sg.flags & 3) - 1;
    r.sy = ((sg.flags >> 2) & 3) - 1;
    r.sz = ((sg.flags >> 4) & 3) - 1;
    // Hash coordinates stepped with the same seam wrap as the walk, so
    // tile membership agrees with the segment cut.
    HashCoord hc = hash_coord_of(a, r.ix, r.iy, r.iz);
    for (int step = sg.step; step < r.max_steps; ++step) {
      const int lx = hc.cx - x0, ly = hc.cy - y0, lz = hc.cz - z0;
      const bool go =
          !(r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) &&
          abs(r.ix - a.oix) <= a.hx && abs(r.iy - a.oiy) <= a.hy &&
          abs(r.iz - a.oiz) <= a.hz &&
          static_cast<unsigned>(lx) < static_cast<unsigned>(TX) &&
          static_cast<unsigned>(ly) < static_cast<unsigned>(TY) &&
          static_cast<unsigned>(lz) < static_cast<unsigned>(TZ);
      if (!go) break;
      atomicAdd(s_cnt + (lx * TY + ly) * TZ + lz, 1u);
      const int ax = ray_advance(r);
      hash_coord_step(hc, ax, r, a);
    }
  }
}

__device__ __forceinline__ void lds_zero(uint4 *s4) {
  for (int k = threadIdx.x; k < kTileVox / 4; k += kApplyThreads) {
    s4[k] = make_uint4(0, 0, 0, 0);
  }
}

// Thread t owns tile voxels j * 256 + t (j < 32): a z-run of 32 contiguous
// floats per 32 lanes. Hash offset of its j-th voxel:
__device__ __forceinline__ int owned_hash(const RayArgs &a, int x0, int y0,
                                          int z0, int j) {
  const int t = threadIdx.x;
  const int lx = j >> 1;
  const int ly = ((j & 1) << 3) + (t >> 5);
  const int lz = t & 31;
  return (x0 + lx) * (a.my * a.mz) + (y0 + ly) * a.mz + (z0 + lz);
}

// Fold of a thread's owned counts: every occ load is issued before the
// first update, so a tile costs one memory round trip.
__device__ __forceinline__ void fold_owned(const IntegrateArgs &g, int x0,
                                           int y0, int z0, const int *op,
                                           const int *hit,
                                           float *__restrict__ occ,
                                           int *__restrict__ export_op,
                                           int *__restrict__ export_hit) {
  const RayArgs &a = g.a;
  if (export_op) {
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      if (op[j] == 0) continue;
      const int h = owned_hash(a, x0, y0, z0, j);
      export_op[h] = op[j];
      export_hit[h] = hit[j];
    }
    return;
  }
  float v[kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    v[j] = op[j] != 0 ? occ[owned_hash(a, x0, y0, z0, j)] : 0.0f;
  }
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    if (op[j] == 0) continue;
    occ[owned_hash(a, x0, y0, z0, j)] = apply_logodds(v[j], op[j], hit[j], g.p);
  }
}

// One workgroup per work item, items in decreasing cost order so the
// hardware dispatcher balances them. A plain item accumulates its tile in
// LDS passes of kPass records, sums the passes in registers and folds. A
// chunk of a hot tile stores its packed counts to its scratch slab;
// hot_fold sums a tile's slabs and folds.
__global__ void __launch_bounds__(kApplyThreads)
    tile_apply_kernel(const IntegrateArgs g, Workspace w,
                      float *__restrict__ occ, int *__restrict__ export_op,
                      int *__restrict__ export_hit) {
  __shared__ uint4 s_cnt4[kTileVox / 4];
  unsigned *s_cnt = reinterpret_cast<unsigned *>(s_cnt4);
  if (static_cast<int>(blockIdx.x) >= w.misc[0]) return;
  const Item itm = w.items[blockIdx.x];
  const int tz = itm.tile % g.ntz;
  const int ty = (itm.tile / g.ntz) % g.nty;
  const int tx = itm.tile / (g.ntz * g.nty);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  const int end = static_cast<int>(
      min(static_cast<long long>(itm.end), w.seg_capacity));
  if (itm.slab >= 0) {
    lds_zero(s_cnt4);
    __syncthreads();
    tile_accumulate(g, w, itm.beg, end, x0, y0, z0, s_cnt);
    __syncthreads();
    uint4 *slab = reinterpret_cast<uint4 *>(
        w.scratch + static_cast<size_t>(itm.slab) * kTileVox);
    for (int k = threadIdx.x; k < kTileVox / 4; k += kApplyThreads) {
      slab[k] = s_cnt4[k];
    }
    return;
  }
  int op[kPerThread], hit[kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    op[j] = 0;
    hit[j] = 0;
  }
  for (int b = itm.beg; b < end; b += kPass) {
    lds_zero(s_cnt4);
    __syncthreads();
    tile_accumulate(g, w, b, min(b + kPass, end), x0, y0, z0, s_cnt);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      const unsigned c = s_cnt[j * kApplyThreads + threadIdx.x];
      op[j] += static_cast<int>(c & 0xFFFFu);
      hit[j] += static_cast<int>(c >> 16);
    }
    __syncthreads();
  }
  fold_owned(g, x0, y0, z0, op, hit, occ, export_op, export_hit);
}

// Sums the scratch slabs of each split tile and folds: kHotSlices
// workgroups per tile, four consecutive voxels (one z-run piece) per lane.
__global__ void __launch_bounds__(256)
    hot_fold_kernel(const IntegrateArgs g, Workspace w,
                    float *__restrict__ occ, int *__restrict__ export_op,
                    int *__restrict__ export_hit) {
  const int hot = blockIdx.x / kHotSlices;
  if (hot >= w.misc[3]) return;
  const HotTile ht = w.hot[hot];
  const int k4 = (blockIdx.x % kHotSlices) * 256 + threadIdx.x;  // < 2048
  const uint4 *src = reinterpret_cast<const uint4 *>(
                         w.scratch + static_cast<size_t>(ht.slab0) * kTileVox) +
                     k4;
  unsigned op[4] = {0, 0, 0, 0}, hit[4] = {0, 0, 0, 0};
  int q = 0;
  for (; q + 4 <= ht.chunks; q += 4) {
    uint4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = src[(q + u) * (kTileVox / 4)];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const unsigned vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        op[e] += vv[e] & 0xFFFFu;
        hit[e] += vv[e] >> 16;
      }
    }
  }
  for (; q < ht.chunks; ++q) {
    const uint4 v = src[q * (kTileVox / 4)];
    const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      op[e] += vv[e] & 0xFFFFu;
      hit[e] += vv[e] >> 16;
    }
  }
  const RayArgs &a = g.a;
  const int tz = ht.tile % g.ntz;
  const int ty = (ht.tile / g.ntz) % g.nty;
  const int tx = ht.tile / (g.ntz * g.nty);
  const int k = 4 * k4;
  const int lz = k % TZ, ly = (k / TZ) % TY, lx = k / (TZ * TY);
  const int h0 = (tx * TX + lx) * (a.my * a.mz) + (ty * TY + ly) * a.mz +
                 (tz * TZ + lz);
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = (op[e] != 0 && !export_op) ? occ[h0 + e] : 0.0f;
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (op[e] == 0) continue;
    if (export_op) {
      export_op[h0 + e] = static_cast<int>(op[e]);
      export_hit[h0 + e] = static_cast<int>(hit[e]);
    } else {
      occ[h0 + e] = apply_logodds(v[e], static_cast<int>(op[e]),
                                  static_cast<int>(hit[e]), g.p);
    }
  }
}

// ----------------------------------------------------------------------
// near_apply
// ----------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    near_apply_kernel(const IntegrateArgs g, Workspace w,
                      float *__restrict__ occ, int *__restrict__ export_op,
                      int *__restrict__ export_hit) {
  const RayArgs &a = g.a;
  const bool ctr = a.center != 0;
  const int s0x = pos_to_global(a.ox, a.inv_res, ctr);
  const int s0y = pos_to_global(a.oy, a.inv_res, ctr);
  const int s0z = pos_to_global(a.oz, a.inv_res, ctr);
  const int C = g.cube;
  const int cells = C * C * C;
  const int k1 = g.near_k - 1;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cells;
       c += gridDim.x * blockDim.x) {
    const int o = w.near_op[c];
    if (o == 0) continue;
    const int hit = w.near_hit[c];
    w.near_op[c] = 0;
    w.near_hit[c] = 0;
    const int dz = c % C - k1;
    const int dy = (c / C) % C - k1;
    const int dx = c / (C * C) - k1;
    const int h = global_to_hash(s0x + dx, s0y + dy, s0z + dz, a.mx, a.my,
                                 a.mz, a.hx, a.hy, a.hz);
    if (export_op) {
      export_op[h] = o;
      export_hit[h] = hit;
    } else {
      occ[h] = apply_logodds(occ[h], o, hit, g.p);
    }
  }
}

// ----------------------------------------------------------------------
// Host side: geometry, workspace layout, launch sequence.
// ----------------------------------------------------------------------

struct Layout {
  int ntx, nty, ntz, n_tiles, cube, max_seg, max_items;
  long long seg_capacity;
  size_t off_count, off_segs, off_sorted, off_tcount, off_toff, off_tcur,
      off_items, off_hot, off_scratch, off_misc, off_nop, off_nhit, total;
};

inline size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Per-ray segment bound (tile entries + the endpoint record): inside the
// window the walk is monotone per axis, so it changes tile at most
// ntx + nty + ntz times plus one wrap seam per axis; a walk of S steps also
// changes tile at most S / min(tile dim) + 3 times.
Layout make_layout(const cublox_raycast_cfg &c, int max_points, int near_k) {
  Layout L{};
  L.ntx = (c.map_size_i.x + TX - 1) / TX;
  L.nty = (c.map_size_i.y + TY - 1) / TY;
  L.ntz = (c.map_size_i.z + TZ - 1) / TZ;
  L.n_tiles = L.ntx * L.nty * L.ntz;
  L.cube = 2 * near_k - 1;
  const long long win_bound = L.ntx + L.nty + L.ntz + 8;
  const double reach = static_cast<double>(c.max_range) *
                           static_cast<double>(c.inv_resolution) + 2.0;
  long long range_bound = win_bound;
  if (reach < 1e7) {
    const long long steps = 3 * static_cast<long long>(reach) + 8;
    range_bound = steps / std::min(std::min(TX, TY), TZ) + 10;
  }
  L.max_seg = static_cast<int>(std::min(win_bound, range_bound));
  const size_t n = static_cast<size_t>(std::max(max_points, 1));
  L.seg_capacity = static_cast<long long>(n) * L.max_seg;
  // An unsplit tile is one item; a split tile's chunks each own a slab.
  L.max_items = L.n_tiles + kMaxSlabs;
  size_t o = 0;
  L.off_count = o;
  o = align256(o + n * sizeof(int));
  L.off_segs = o;
  o = align256(o + static_cast<size_t>(L.seg_capacity) * sizeof(Segment));
  L.off_sorted = o;
  o = align256(o + static_cast<size_t>(L.seg_capacity) * sizeof(Segment));
  L.off_tcount = o;
  o = align256(o + static_cast<size_t>(L.n_tiles) * 4);
  L.off_toff = o;
  o = align256(o + static_cast<size_t>(L.n_tiles + 1) * 4);
  L.off_tcur = o;
  o = align256(o + static_cast<size_t>(L.n_tiles) * 4);
  L.off_items = o;
  o = align256(o + static_cast<size_t>(L.max_items) * sizeof(Item));
  L.off_hot = o;
  o = align256(o + kMaxSplit * sizeof(HotTile));
  L.off_scratch = o;
  o = align256(o + static_cast<size_t>(kMaxSlabs) * kTileVox * 4);
  L.off_misc = o;
  o = align256(o + 64 * 4);
  const size_t cells = static_cast<size_t>(L.cube) * L.cube * L.cube;
  L.off_nop = o;
  o = align256(o + cells * 4);
  L.off_nhit = o;
  o = align256(o + cells * 4);
  L.total = o;
  return L;
}

Workspace bind(const Layout &L, void *base) {
  char *b = static_cast<char *>(base);
  Workspace w;
  w.seg_count = reinterpret_cast<int *>(b + L.off_count);
  w.segs = reinterpret_cast<Segment *>(b + L.off_segs);
  w.sorted = reinterpret_cast<Segment *>(b + L.off_sorted);
  w.tile_count = reinterpret_cast<int *>(b + L.off_tcount);
  w.tile_offset = reinterpret_cast<int *>(b + L.off_toff);
  w.tile_cursor = reinterpret_cast<int *>(b + L.off_tcur);
  w.items = reinterpret_cast<Item *>(b + L.off_items);
  w.hot = reinterpret_cast<HotTile *>(b + L.off_hot);
  w.scratch = reinterpret_cast<unsigned *>(b + L.off_scratch);
  w.misc = reinterpret_cast<int *>(b + L.off_misc);
  w.near_op = reinterpret_cast<int *>(b + L.off_nop);
  w.near_hit = reinterpret_cast<int *>(b + L.off_nhit);
  w.seg_capacity = L.seg_capacity;
  w.max_items = L.max_items;
  return w;
}

} // namespace cublox_hip

using namespace cublox_hip;

extern "C" {

size_t cublox_integrate_workspace_bytes(const cublox_raycast_cfg *cfg,
                                        int max_points, int near_steps) {
  if (!cfg || near_steps < 1 || near_steps > kMaxNearK) return 0;
  return make_layout(*cfg, max_points, near_steps).total;
}

int cublox_integrate_workspace_init(void *d_ws, size_t bytes, cublox_stream s) {
  if (!d_ws || bytes == 0) return CUBLOX_OK;
  return cublox_fill_zero_async(d_ws, bytes / 4, s);
}

int cublox_launch_integrate(const cublox_raycast_cfg *cfg, const float *T,
                            const float *d_xs, const float *d_ys,
                            const float *d_zs, int n, float *d_occ,
                            float l_hit, float l_miss, float l_min,
                            float l_max, float l_unknown, void *d_ws,
                            size_t ws_bytes, int max_points, int near_steps,
                            int *d_export_op, int *d_export_hit,
                            cublox_stream s) {
  if (!cfg || n <= 0 || !d_xs || !d_ys || !d_zs || !d_occ) return CUBLOX_OK;
  if (near_steps < 1 || near_steps > kMaxNearK) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: near_steps out of [1, 12]");
  }
  if (n > max_points) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate: n exceeds max_points");
  }
  const Layout L = make_layout(*cfg, max_points, near_steps);
  if (!d_ws || ws_bytes < L.total) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate: workspace too small");
  }
  if ((d_export_op == nullptr) != (d_export_hit == nullptr)) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: export needs both arrays");
  }
  IntegrateArgs g{};
  g.a = make_ray_args(*cfg);
  if (T) {
    for (int k = 0; k < 12; ++k) g.a.T[k] = T[k];
    g.a.ox = T[3];
    g.a.oy = T[7];
    g.a.oz = T[11];
  }
  g.n = n;
  g.ntx = L.ntx;
  g.nty = L.nty;
  g.ntz = L.ntz;
  g.n_tiles = L.n_tiles;
  g.near_k = near_steps;
  g.cube = L.cube;
  g.max_seg = L.max_seg;
  g.p = ApplyArgs{l_hit, l_miss, l_min, l_max, l_unknown};
  const Workspace w = bind(L, d_ws);
  hipStream_t hs = as_stream(s);
  const dim3 wb(kWalkBlock);
  const dim3 wg((n + kWalkBlock - 1) / kWalkBlock);
  if (T) {
    walk_kernel<true><<<wg, wb, 0, hs>>>(g, d_xs, d_ys, d_zs, w);
  } else {
    walk_kernel<false><<<wg, wb, 0, hs>>>(g, d_xs, d_ys, d_zs, w);
  }
  int st = check_launch("walk_kernel");
  if (st != CUBLOX_OK) return st;
  const dim3 bg((n + kBinRays - 1) / kBinRays);
  bin_kernel<false><<<bg, dim3(256), 0, hs>>>(g, w);
  st = check_launch("bin_kernel<count>");
  if (st != CUBLOX_OK) return st;
  tile_scan_kernel<<<dim3(1), dim3(kScanThreads), 0, hs>>>(L.n_tiles, w);
  st = check_launch("tile_scan_kernel");
  if (st != CUBLOX_OK) return st;
  bin_kernel<true><<<bg, dim3(256), 0, hs>>>(g, w);
  st = check_launch("bin_kernel<place>");
  if (st != CUBLOX_OK) return st;
  tile_apply_kernel<<<dim3(L.max_items), dim3(kApplyThreads), 0, hs>>>(
      g, w, d_occ, d_export_op, d_export_hit);
  st = check_launch("tile_apply_kernel");
  if (st != CUBLOX_OK) return st;
  hot_fold_kernel<<<dim3(kMaxSplit * kHotSlices), dim3(256), 0, hs>>>(
      g, w, d_occ, d_export_op, d_export_hit);
  st = check_launch("hot_fold_kernel");
  if (st != CUBLOX_OK) return st;
  const int cells = L.cube * L.cube * L.cube;
  near_apply_kernel<<<dim3(std::max(1, std::min((cells + 255) / 256, 1024))),
                      dim3(256), 0, hs>>>(g, w, d_occ, d_export_op,
                                          d_export_hit);
  return check_launch("near_apply_kernel");
}

} // extern "C"
