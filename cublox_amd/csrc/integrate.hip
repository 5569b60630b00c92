// Tile-binned fused scan update: ray cast + clamped log-odds fold straight
// into the occupancy buffer, without per-voxel global counters.
//
// Same results as the reference's raycast-then-apply pair (kernels.cu:33-210)
// — the same per-voxel op/hit counts folded once per scan — organised for
// CDNA4 instead of one global atomic per traversed voxel:
//
//   walk_count  one ray per lane. Steps < K (the near field: step s visits a
//               voxel at L1 distance s from the sensor voxel) go to a small
//               cube of counters around the sensor voxel with wavefront-
//               deduplicated atomics. Beyond, the walk is cut into
//               "segments" at every entry into a 16x16x32 tile of the
//               hash-ordered grid; the segments are counted per tile.
//   tile_scan   exclusive scan of the per-tile counts (one workgroup) and
//               the list of non-empty tiles.
//   walk_emit   the same walk again, writing each segment — the exact DDA
//               state at the tile entry — into its tile's range. The state
//               is carried, never recomputed: the walk's t_max recurrence is
//               a running FP32 sum, so a segment resumes bit-identically.
//   tile_apply  one workgroup per non-empty tile: resume every segment
//               inside the tile with LDS counters (no global atomics), then
//               fold the tile's counts into occ with coalesced z-runs.
//   near_apply  fold the near-field cube into occ and clear it.
//
// Every voxel's visits happen at one L1 distance from the sensor voxel, so
// the near-field cube and the tiles own disjoint voxel sets and each voxel
// is folded exactly once.
#include "raycast_common.hpp"

#include <algorithm>

namespace cublox_hip {

constexpr int TX = 16, TY = 16, TZ = 32;
constexpr int kTileVox = TX * TY * TZ;  // 8192
constexpr int kWalkBlock = 256;

struct RayInfo {  // 32 B
  float tDx, tDy, tDz;
  int eix, eiy, eiz;
  int max_steps;
  int flags;  // bits 0..5: sx+1, sy+1, sz+1 (2 bits each); bit 6: clipped
};

struct Segment {  // 32 B; step < 0 marks an endpoint record
  int ray;
  int step;
  int ix, iy, iz;
  float tMx, tMy, tMz;
};

struct IntegrateArgs {
  RayArgs a;
  int n;
  int ntx, nty, ntz, n_tiles;
  int near_k;  // K
  int cube;    // 2K - 1
  int max_seg; // per-ray segment bound (capacity check only)
  ApplyArgs p;
};

struct Workspace {
  RayInfo *rays;     // n
  Segment *segs;     // capacity
  int *tile_count;   // n_tiles (zero between launches)
  int *tile_offset;  // n_tiles + 1
  int *tile_cursor;  // n_tiles
  int *tile_list;    // n_tiles
  int *misc;         // [0] non-empty tiles, [1] segments, [2] overflow flag
  int *near_op;      // cube^3 (zero between launches)
  int *near_hit;     // cube^3 (zero between launches)
  long long seg_capacity;
};

__device__ __forceinline__ int wrap_inc(int c, int s, int m) {
  c += s;
  if (c >= m) c -= m;
  if (c < 0) c += m;
  return c;
}

__device__ __forceinline__ int tile_of(const IntegrateArgs &g, int cx, int cy,
                                       int cz) {
  return ((cx / TX) * g.nty + (cy / TY)) * g.ntz + (cz / TZ);
}

__device__ __forceinline__ int cube_index(const IntegrateArgs &g, int dx,
                                          int dy, int dz) {
  const int k1 = g.near_k - 1;
  return ((dx + k1) * g.cube + (dy + k1)) * g.cube + (dz + k1);
}

// The walk shared by walk_count (kEmit = false) and walk_emit (kEmit = true).
template <bool kTransform, bool kEmit>
__global__ void __launch_bounds__(kWalkBlock)
    walk_kernel(const IntegrateArgs g, const float *__restrict__ xs,
                const float *__restrict__ ys, const float *__restrict__ zs,
                Workspace w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n) return;
  const RayArgs &a = g.a;
  RayState r;
  if (!ray_setup<kTransform>(a, xs, ys, zs, i, r)) return;
  if (!kEmit) {
    RayInfo ri;
    ri.tDx = r.tDx;
    ri.tDy = r.tDy;
    ri.tDz = r.tDz;
    ri.eix = r.eix;
    ri.eiy = r.eiy;
    ri.eiz = r.eiz;
    ri.max_steps = r.max_steps;
    ri.flags = (r.sx + 1) | ((r.sy + 1) << 2) | ((r.sz + 1) << 4) |
               (r.clipped ? 64 : 0);
    w.rays[i] = ri;
  }
  const int s0x = r.ix, s0y = r.iy, s0z = r.iz;  // sensor voxel
  int cx = global_to_local(r.ix, a.mx, a.hx) + a.hx;
  int cy = global_to_local(r.iy, a.my, a.hy) + a.hy;
  int cz = global_to_local(r.iz, a.mz, a.hz) + a.hz;
  int cur = -1;
  const int K = g.near_k;

  for (int step = 0; step < r.max_steps; ++step) {
    if (r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) break;
    const bool inwin = abs(r.ix - a.oix) <= a.hx && abs(r.iy - a.oiy) <= a.hy &&
                       abs(r.iz - a.oiz) <= a.hz;
    if (step < K) {
      if (!kEmit) {
        const int key =
            inwin ? cube_index(g, r.ix - s0x, r.iy - s0y, r.iz - s0z) : -1;
        wave_add_unique(w.near_op, key);
      }
    } else if (inwin) {
      const int t = tile_of(g, cx, cy, cz);
      if (t != cur) {
        cur = t;
        if (!kEmit) {
          wave_add_unique(w.tile_count, t);
        } else {
          const int slot = atomicAdd(w.tile_cursor + t, 1);
          if (slot < w.seg_capacity) {
            Segment sg;
            sg.ray = i;
            sg.step = step;
            sg.ix = r.ix;
            sg.iy = r.iy;
            sg.iz = r.iz;
            sg.tMx = r.tMx;
            sg.tMy = r.tMy;
            sg.tMz = r.tMz;
            w.segs[slot] = sg;
          } else {
            w.misc[2] = 1;
          }
        }
      }
    } else {
      cur = -1;
    }
    const int ax = ray_advance(r);
    if (ax == 0) {
      cx = wrap_inc(cx, r.sx, a.mx);
    } else if (ax == 1) {
      cy = wrap_inc(cy, r.sy, a.my);
    } else {
      cz = wrap_inc(cz, r.sz, a.mz);
    }
  }

  // Endpoint (kernels.cu:141-151): op += 1, hit += 1 unless clipped.
  if (inside_window(r.eix, r.eiy, r.eiz, a.oix, a.oiy, a.oiz, a.hx, a.hy, a.hz)) {
    const int ddx = r.eix - s0x, ddy = r.eiy - s0y, ddz = r.eiz - s0z;
    if (abs(ddx) + abs(ddy) + abs(ddz) < K) {
      if (!kEmit) {
        const int c = cube_index(g, ddx, ddy, ddz);
        atomic_add_i32(w.near_op + c, 1);
        if (!r.clipped) atomic_add_i32(w.near_hit + c, 1);
      }
    } else {
      const int ecx = global_to_local(r.eix, a.mx, a.hx) + a.hx;
      const int ecy = global_to_local(r.eiy, a.my, a.hy) + a.hy;
      const int ecz = global_to_local(r.eiz, a.mz, a.hz) + a.hz;
      const int t = tile_of(g, ecx, ecy, ecz);
      if (!kEmit) {
        atomic_add_i32(w.tile_count + t, 1);
      } else {
        const int slot = atomicAdd(w.tile_cursor + t, 1);
        if (slot < w.seg_capacity) {
          Segment sg;
          sg.ray = i;
          sg.step = -1;
          sg.ix = r.eix;
          sg.iy = r.eiy;
          sg.iz = r.eiz;
          sg.tMx = r.clipped ? 1.0f : 0.0f;
          sg.tMy = 0.0f;
          sg.tMz = 0.0f;
          w.segs[slot] = sg;
        } else {
          w.misc[2] = 1;
        }
      }
    }
  }
}

// Exclusive scan of tile counts, non-empty tile list, counts cleared.
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kScanThreads)
    tile_scan_kernel(const int n_tiles, Workspace w) {
  __shared__ int s_sum[kScanThreads];
  __shared__ int s_nz[kScanThreads];
  const int t = threadIdx.x;
  const int per = (n_tiles + kScanThreads - 1) / kScanThreads;
  const int beg = min(t * per, n_tiles);
  const int end = min(beg + per, n_tiles);
  int sum = 0, nz = 0;
  for (int k = beg; k < end; ++k) {
    const int c = w.tile_count[k];
    sum += c;
    nz += c > 0;
  }
  s_sum[t] = sum;
  s_nz[t] = nz;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 partials.
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const int a = t >= off ? s_sum[t - off] : 0;
    const int b = t >= off ? s_nz[t - off] : 0;
    __syncthreads();
    s_sum[t] += a;
    s_nz[t] += b;
    __syncthreads();
  }
  int run = s_sum[t] - sum;
  int pos = s_nz[t] - nz;
  for (int k = beg; k < end; ++k) {
    const int c = w.tile_count[k];
    w.tile_offset[k] = run;
    w.tile_cursor[k] = run;
    if (c > 0) w.tile_list[pos++] = k;
    run += c;
    w.tile_count[k] = 0;
  }
  if (t == kScanThreads - 1) {
    w.tile_offset[n_tiles] = s_sum[t];
    w.misc[0] = s_nz[t];
    w.misc[1] = s_sum[t];
  }
}

// One workgroup per non-empty tile.
__global__ void __launch_bounds__(256)
    tile_apply_kernel(const IntegrateArgs g, Workspace w,
                      float *__restrict__ occ, int *__restrict__ export_op,
                      int *__restrict__ export_hit) {
  __shared__ int s_op[kTileVox];
  __shared__ int s_hit[kTileVox];
  if (static_cast<int>(blockIdx.x) >= w.misc[0]) return;
  const int tile = w.tile_list[blockIdx.x];
  const RayArgs &a = g.a;
  const int tz = tile % g.ntz;
  const int ty = (tile / g.ntz) % g.nty;
  const int tx = tile / (g.ntz * g.nty);
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  for (int k = threadIdx.x; k < kTileVox; k += blockDim.x) {
    s_op[k] = 0;
    s_hit[k] = 0;
  }
  __syncthreads();

  const int beg = w.tile_offset[tile];
  const int end = min(static_cast<long long>(w.tile_offset[tile + 1]),
                      w.seg_capacity);
  for (int k = beg + threadIdx.x; k < end; k += blockDim.x) {
    const Segment sg = w.segs[k];
    const RayInfo ri = w.rays[sg.ray];
    if (sg.step < 0) {
      const int ecx = global_to_local(sg.ix, a.mx, a.hx) + a.hx - x0;
      const int ecy = global_to_local(sg.iy, a.my, a.hy) + a.hy - y0;
      const int ecz = global_to_local(sg.iz, a.mz, a.hz) + a.hz - z0;
      const int li = (ecx * TY + ecy) * TZ + ecz;
      atomicAdd(s_op + li, 1);
      if (sg.tMx == 0.0f) atomicAdd(s_hit + li, 1);
      continue;
    }
    RayState r;
    r.ix = sg.ix;
    r.iy = sg.iy;
    r.iz = sg.iz;
    r.tMx = sg.tMx;
    r.tMy = sg.tMy;
    r.tMz = sg.tMz;
    r.tDx = ri.tDx;
    r.tDy = ri.tDy;
    r.tDz = ri.tDz;
    r.eix = ri.eix;
    r.eiy = ri.eiy;
    r.eiz = ri.eiz;
    r.max_steps = ri.max_steps;
    r.sx = (ri.flags & 3) - 1;
    r.sy = ((ri.flags >> 2) & 3) - 1;
    r.sz = ((ri.flags >> 4) & 3) - 1;
    // Hash coordinates of the segment's first voxel, stepped with the same
    // seam wrap as the emitting walk so that tile membership agrees.
    int cx = global_to_local(r.ix, a.mx, a.hx) + a.hx;
    int cy = global_to_local(r.iy, a.my, a.hy) + a.hy;
    int cz = global_to_local(r.iz, a.mz, a.hz) + a.hz;
    for (int step = sg.step; step < r.max_steps; ++step) {
      if (r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) break;
      if (!(abs(r.ix - a.oix) <= a.hx && abs(r.iy - a.oiy) <= a.hy &&
            abs(r.iz - a.oiz) <= a.hz)) {
        break;
      }
      const int lx = cx - x0, ly = cy - y0, lz = cz - z0;
      if (static_cast<unsigned>(lx) >= static_cast<unsigned>(TX) ||
          static_cast<unsigned>(ly) >= static_cast<unsigned>(TY) ||
          static_cast<unsigned>(lz) >= static_cast<unsigned>(TZ)) {
        break;
      }
      atomicAdd(s_op + (lx * TY + ly) * TZ + lz, 1);
      const int ax = ray_advance(r);
      if (ax == 0) {
        cx = wrap_inc(cx, r.sx, a.mx);
      } else if (ax == 1) {
        cy = wrap_inc(cy, r.sy, a.my);
      } else {
        cz = wrap_inc(cz, r.sz, a.mz);
      }
    }
  }
  __syncthreads();

  // Fold: lanes walk z-runs of TZ voxels (contiguous in occ).
  const int yz = a.my * a.mz;
  for (int k = threadIdx.x; k < kTileVox; k += blockDim.x) {
    const int o = s_op[k];
    if (o == 0) continue;
    const int lz = k % TZ;
    const int ly = (k / TZ) % TY;
    const int lx = k / (TZ * TY);
    const int h = (x0 + lx) * yz + (y0 + ly) * a.mz + (z0 + lz);
    const int hit = s_hit[k];
    if (export_op) {
      export_op[h] = o;
      export_hit[h] = hit;
    } else {
      occ[h] = apply_logodds(occ[h], o, hit, g.p);
    }
  }
}

// Fold and clear the near-field cube.
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
  int ntx, nty, ntz, n_tiles, cube, max_seg;
  long long seg_capacity;
  size_t off_rays, off_segs, off_tcount, off_toff, off_tcur, off_tlist,
      off_misc, off_nop, off_nhit, total;
};

inline size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Per-ray bound on segments (tile entries + the endpoint record): inside
// the window the walk is monotone per axis, so it changes tile at most
// ntx + nty + ntz times plus one wrap seam per axis; a walk of S steps
// also changes tile at most S / min(tile dim) + 3 times.
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
    range_bound = steps / std::min(std::min(TX, TY), TZ) + 8;
  }
  L.max_seg = static_cast<int>(std::min(win_bound, range_bound));
  L.seg_capacity = static_cast<long long>(std::max(max_points, 1)) * L.max_seg;
  size_t o = 0;
  const size_t n = static_cast<size_t>(std::max(max_points, 1));
  L.off_rays = o;
  o = align256(o + n * sizeof(RayInfo));
  L.off_segs = o;
  o = align256(o + static_cast<size_t>(L.seg_capacity) * sizeof(Segment));
  L.off_tcount = o;
  o = align256(o + static_cast<size_t>(L.n_tiles) * 4);
  L.off_toff = o;
  o = align256(o + static_cast<size_t>(L.n_tiles + 1) * 4);
  L.off_tcur = o;
  o = align256(o + static_cast<size_t>(L.n_tiles) * 4);
  L.off_tlist = o;
  o = align256(o + static_cast<size_t>(L.n_tiles) * 4);
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
  w.rays = reinterpret_cast<RayInfo *>(b + L.off_rays);
  w.segs = reinterpret_cast<Segment *>(b + L.off_segs);
  w.tile_count = reinterpret_cast<int *>(b + L.off_tcount);
  w.tile_offset = reinterpret_cast<int *>(b + L.off_toff);
  w.tile_cursor = reinterpret_cast<int *>(b + L.off_tcur);
  w.tile_list = reinterpret_cast<int *>(b + L.off_tlist);
  w.misc = reinterpret_cast<int *>(b + L.off_misc);
  w.near_op = reinterpret_cast<int *>(b + L.off_nop);
  w.near_hit = reinterpret_cast<int *>(b + L.off_nhit);
  w.seg_capacity = L.seg_capacity;
  return w;
}

} // namespace cublox_hip

using namespace cublox_hip;

extern "C" {

size_t cublox_integrate_workspace_bytes(const cublox_raycast_cfg *cfg,
                                        int max_points, int near_steps) {
  if (!cfg || near_steps < 1) return 0;
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
  if (near_steps < 1 || near_steps > 64) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: near_steps out of [1, 64]");
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
    walk_kernel<true, false><<<wg, wb, 0, hs>>>(g, d_xs, d_ys, d_zs, w);
  } else {
    walk_kernel<false, false><<<wg, wb, 0, hs>>>(g, d_xs, d_ys, d_zs, w);
  }
  int st = check_launch("walk_kernel<count>");
  if (st != CUBLOX_OK) return st;
  tile_scan_kernel<<<dim3(1), dim3(kScanThreads), 0, hs>>>(L.n_tiles, w);
  st = check_launch("tile_scan_kernel");
  if (st != CUBLOX_OK) return st;
  if (T) {
    walk_kernel<true, true><<<wg, wb, 0, hs>>>(g, d_xs, d_ys, d_zs, w);
  } else {
    walk_kernel<false, true><<<wg, wb, 0, hs>>>(g, d_xs, d_ys, d_zs, w);
  }
  st = check_launch("walk_kernel<emit>");
  if (st != CUBLOX_OK) return st;
  tile_apply_kernel<<<dim3(L.n_tiles), dim3(256), 0, hs>>>(g, w, d_occ,
                                                            d_export_op,
                                                            d_export_hit);
  st = check_launch("tile_apply_kernel");
  if (st != CUBLOX_OK) return st;
  const int cells = L.cube * L.cube * L.cube;
  near_apply_kernel<<<dim3(std::max(1, std::min((cells + 255) / 256, 1024))),
                      dim3(256), 0, hs>>>(g, w, d_occ, d_export_op,
                                          d_export_hit);
  return check_launch("near_apply_kernel");
}

} // extern "C"
