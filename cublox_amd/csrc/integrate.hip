// Tile-binned fused scan update: ray cast + clamped log-odds fold straight
// into the occupancy buffer, without per-voxel global counters.
//
// Same results as the reference's raycast-then-apply pair (kernels.cu:33-210)
// — the same per-voxel op/hit counts folded once per scan — organised for
// CDNA4, where device-scope atomics are the scarce resource (measured on
// MI355X: ~27 G/s scattered, tens of M/s on one address) and LDS is
// plentiful:
//
//   walk        one ray per lane. Steps < K (the near field: step s visits a
//               voxel at L1 distance s from the sensor voxel) count into an
//               LDS cube around the sensor voxel, flushed once per
//               workgroup. Beyond, the walk is cut into "segments" at every
//               entry into a 16x16x32 tile of the hash-ordered grid: each
//               segment (16 B) carries the exact DDA state at the tile entry
//               and a packed word (the ray's increments are stored once per
//               ray) and goes to the ray's own slot range, its tile key to a
//               parallel array. The four-lane walk (launches of up to
//               262,144 rays) also counts its workgroup's segments per tile
//               in an LDS hash: one global atomic per (workgroup, tile).
//   bin_count   (after the one-lane walk of large launches) the same
//               counts, from the stored segments.
//   tile scan   tile run starts and the work-item plan — one item per
//               tile, or one per chunk of a hot tile — in decreasing cost
//               classes: per-block sums, bases, then per-block cursors
//               and items (for up to 12288 tiles one pass of 256-tile
//               blocks that publish their sums to each other; beyond,
//               sums and bases passes first).
//   place       each record's slot index written in tile order (its tile's
//               run start + its counting group's base there + its rank).
//   tile_apply  one workgroup per work item: replay every segment's voxels
//               into packed 16+16-bit LDS counters (32 KB), then fold the
//               tile's counts into occ with all loads in flight. Chunks of
//               a hot tile store their counts to scratch slabs.
//   hot_fold    sums each split tile's slabs and folds, 8 workgroups a tile;
//               its last workgroups fold the near-field cubes into occ and
//               clear them.
//
// One launch sequence serves one grid or a batch of grids (each with its
// own geometry, sensor pose, points and occupancy buffer): per-grid
// descriptors live in device memory and every kernel maps its rays, tiles
// or cells to their grid, so a batch of any mix of grids drains through
// the same six-to-eight launches.
//
// The walk's t_max recurrence is a running FP32 sum, so a segment resumed
// from its carried state is bit-identical to the uninterrupted walk. Every
// voxel's visits happen at one L1 distance from the sensor voxel, so the
// near-field cube and the tiles own disjoint voxel sets and each voxel is
// folded exactly once.
#include "raycast_common.hpp"
#include "zero_regions.hpp"

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

namespace cublox_hip {

constexpr int TX = 16, TY = 16, TZ = 32;
constexpr int kTileVox = TX * TY * TZ;  // 8192
constexpr int kWalkBlock = 256;
constexpr int kMaxNearK = 12;
constexpr int kRayAlign = 1024;   // grids' ray ranges start at multiples of it
constexpr int kBinThreads = 512;
constexpr int kHash = 4096;
constexpr int kBinCache = 16;  // records per binning thread kept for placing
constexpr int kWalkHash = 1024;  // per-walk-workgroup tile record counts
constexpr int kHashProbes = 64;
// Records per chunk of a hot tile (single-grid default). Every chunk costs a
// 32 KB slab written and read back; larger chunks mean fewer of them and
// longer work items (measured at 701x701x151: 1024 records 218, 2048 210,
// 4096 208, 8192 216 us per update).
constexpr int kChunk = 4096;
constexpr int kApplyThreads = 256;
constexpr int kPerThread = kTileVox / kApplyThreads;  // 32 voxels
constexpr int kHotSlices = 8;     // hot_fold workgroups per split tile
constexpr int kCostBuckets = 4;   // classes of item cost (records per item)

// A segment is the part of one ray inside one tile span: its DDA state at
// the span entry and in `meta` the entry voxel's index in the tile (bits
// 0..12), the number of voxels it visits (13..19), the step signs (20..25,
// s + 1 two bits each), an endpoint flag (26) and the clipped flag (27).
// The increments are the ray's own (ray_td). An endpoint record adds op
// (and hit unless clipped) at its voxel.
struct Segment {  // 16 B
  float tMx, tMy, tMz;
  unsigned meta;
};
static_assert(sizeof(Segment) == 16, "a record moves as one 16 B word");

constexpr unsigned kMetaLi = 0x1FFFu;
constexpr int kMetaN = 13;
constexpr int kMetaSign = 20;
constexpr unsigned kMetaEndpoint = 1u << 26;
constexpr unsigned kMetaClipped = 1u << 27;

struct Item {  // one tile_apply work unit: records [beg, end) of a tile
  int tile, beg, end;
  int slab;  // < 0: accumulate and fold; else a hot-tile chunk's slab
};

// A record's place in its tile, as the counting pass leaves it (the quad
// walk, or the binning count after the one-lane walk; both count records
// per tile in an LDS hash per workgroup, a "counting group" of rays): bit
// 31 set, the place itself, taken from the tile's global count when the
// hash had no room; else the record's hash entry e (bits 19..30) and its
// rank among the group's records of that tile (bits 0..18: a ray's path
// enters one tile once, and once more per wrap seam it crosses, and its
// endpoint adds a record, so a ray has at most five records in a tile and
// a group of at most kRayAlign rays fewer than 2^19). The
// group's first place in the tile is grp_base[group * hash size + e].
constexpr unsigned kRankGlobal = 1u << 31;
constexpr int kRankBits = 19;
__device__ __forceinline__ unsigned rank_code(int e, unsigned rank) {
  return (static_cast<unsigned>(e) << kRankBits) | rank;
}

struct HotTile {  // a tile split into chunks, folded by hot_fold
  int tile, slab0, chunks;
  int atomic;  // the chunks add into one shared op/hit slab pair
};

// Slabs a hot tile of `chunks` chunks takes: one per chunk, or in the
// shared mode one op and one hit slab (32-bit counts) that every chunk adds
// its counts into with global atomics.
constexpr int kSharedSlabs = 2;
__host__ __device__ __forceinline__ int slab_need(int chunks, bool shared) {
  return shared ? kSharedSlabs : chunks;
}

struct IntegrateArgs {
  RayArgs a;
  int n;
  int ntx, nty, ntz, n_tiles;
  int near_k;   // K
  int cube;     // 2K - 1
  int max_seg;  // per-ray segment slots
  int qx, qy, qz;  // window origin modulo the map size (non-negative)
  ApplyArgs p;
};

// Hash coordinate of an in-window global coordinate g (|g - o| <= h, map
// m = 2h + 1): (g + h) mod m, i.e. global_to_local(g) + h, as
// (q + (g - o + h)) mod m with q = o mod m and q + (g - o + h) < 2m.
__device__ __forceinline__ int window_hash(int g, int o, int h, int m, int q) {
  const int v = q + (g - o + h);
  return v >= m ? v - m : v;
}

__device__ __forceinline__ HashCoord window_hash_coord(const IntegrateArgs &g,
                                                       int gx, int gy, int gz) {
  const RayArgs &a = g.a;
  return HashCoord{window_hash(gx, a.oix, a.hx, a.mx, g.qx),
                   window_hash(gy, a.oiy, a.hy, a.my, g.qy),
                   window_hash(gz, a.oiz, a.hz, a.mz, g.qz)};
}

// Tile span origin entered when a walk in direction s leaves the span at
// t0 of an axis with map m and tile width T (the seam joins m - 1 and 0),
// and the width of a span.
__device__ __forceinline__ int next_span(int t0, int s, int m, int T) {
  if (s > 0) return t0 + T >= m ? 0 : t0 + T;
  return t0 == 0 ? ((m - 1) / T) * T : t0 - T;
}

__device__ __forceinline__ int span_width(int t0, int m, int T) {
  return min(T, m - t0);
}

struct Workspace {
  int *seg_count;     // n: segments written by each ray
  int *seg_tile;      // n * max_seg: tile of each slot's segment
  Segment *segs;      // n * max_seg: record j of ray i at j * n + i
  float4 *ray_td;     // per ray: the DDA increments (tDx, tDy, tDz, 0)
  unsigned *order;    // n * max_seg: record slots in tile order
  int *tile_count;    // n_tiles (zero between launches)
  int *tile_cursor;   // n_tiles
  Item *items;        // max_items, costliest first
  HotTile *hot;       // split tiles
  unsigned *scratch;  // one kTileVox slab per hot-tile chunk, op | hit << 16
  int *misc;          // [0] items, [1] segments, [2] overflow flag, [3] split
                      // tiles, [4] unsplit hot tiles
  int *near_op;       // near-field cells (zero between launches)
  int *near_hit;      // near-field cells (zero between launches)
  int *big;           // item slots of unsplit hot tiles (passes)
  int *block_sums;    // multi-block tile scan: per-block sums
  int *scan_flag;     // one-pass small scan: block sums published (zero between launches)
  int *block_bases;   // multi-block tile scan: per-block bases
  unsigned *seg_rank; // n * max_seg: each record's place in its tile (rank_code)
  int *grp_base;      // per counting group and hash entry: the group's first
                      // place in that entry's tile
  long long seg_capacity;
  int max_items;
  int slabs_cap;      // scratch slabs
  int chunk;          // records per hot-tile chunk (single-grid launches)
};

__device__ __forceinline__ int tile_of(const IntegrateArgs &g, int cx, int cy,
                                       int cz) {
  return ((cx / TX) * g.nty + (cy / TY)) * g.ntz + (cz / TZ);
}

// Tile of the spans starting at hash coordinates (t0x, t0y, t0z).
__device__ __forceinline__ int span_tile(const IntegrateArgs &g, int t0x,
                                         int t0y, int t0z) {
  const unsigned tx = static_cast<unsigned>(t0x) / TX;
  const unsigned ty = static_cast<unsigned>(t0y) / TY;
  const unsigned tz = static_cast<unsigned>(t0z) / TZ;
  return static_cast<int>((tx * g.nty + ty) * g.ntz + tz);
}

__device__ __forceinline__ int cube_index(const IntegrateArgs &g, int dx,
                                          int dy, int dz) {
  const int k1 = g.near_k - 1;
  return ((dx + k1) * g.cube + (dy + k1)) * g.cube + (dz + k1);
}

constexpr int kNever = INT_MIN;  // a coordinate no walk reaches

// First global coordinate past the window [o - h, o + h] for travel
// direction s (kNever when the axis does not move).
__device__ __forceinline__ int window_exit(int s, int o, int h) {
  return s > 0 ? o + h + 1 : (s < 0 ? o - h - 1 : kNever);
}

// First global coordinate, travelling in direction s from global g with
// hash coordinate c, outside both the tile span [t0, t0 + tw) of hash
// coordinates and the window (wexit from window_exit). Inside a tile span
// the hash coordinate moves with the global one (the wrap seam is a tile
// boundary).
__device__ __forceinline__ int exit_coord(int g, int s, int c, int t0, int tw,
                                          int wexit) {
  if (s > 0) return min(g + (t0 + tw - c), wexit);
  if (s < 0) return max(g - (c - t0) - 1, wexit);
  return kNever;
}

// ----------------------------------------------------------------------
// walk
// ----------------------------------------------------------------------
constexpr int kBig = 1 << 20;  // steps to the exit of an axis that never moves

// A_k of the running sum A_0 = h, A_{j+1} = A_j + d, formed in order as the
// walk forms its crossing parameters.
__device__ __forceinline__ float seq_at(float h, float d, int k) {
  for (; k >= 4; k -= 4) h = (((h + d) + d) + d) + d;  // each add rounded in turn
  for (; k > 0; --k) h += d;
  return h;
}

// The walk steps the axis with the smallest next crossing parameter, z
// before y before x on ties (reference kernels.cu:120-136), so its steps
// are the merge of the three per-axis running sums in that order. Steps of
// one axis (next parameter h, increment d, at most nb of them) that come
// before an event at parameter e: those below e, or equal to it when the
// axis goes first on ties. h is left at the axis's next parameter.
__device__ __forceinline__ int count_before(float &h, float d, int nb, float e,
                                            bool first_on_tie) {
  // The sums increase (d > 0), so "before e" holds for a prefix of them:
  // four candidates per iteration, counted by how many hold.
  auto before = [&](float x) { return x < e || (first_on_tie && x == e); };
  int c = 0;
  while (c + 4 <= nb) {
    const float h1 = h + d, h2 = h1 + d, h3 = h2 + d;
    if (!before(h3)) {
      if (before(h)) {
        ++c;
        h = h1;
        if (before(h1)) {
          ++c;
          h = h2;
          if (before(h2)) {
            ++c;
            h = h3;
          }
        }
      }
      return c;
    }
    h = h3 + d;
    c += 4;
  }
  while (c < nb && before(h)) {
    h += d;
    ++c;
  }
  return c;
}

// Whether the walk, from its current state, is at the end voxel after
// dx + dy + dz more steps (dx, dy, dz: steps still to go on each axis):
// every step it needs must come before every step of another axis past
// the end voxel.
__device__ __forceinline__ bool end_reached(const RayState &r, int dx, int dy,
                                            int dz) {
  const float inf = __builtin_inff();
  const float inx = dx > 0 ? seq_at(r.tMx, r.tDx, dx - 1) : -inf;
  const float iny = dy > 0 ? seq_at(r.tMy, r.tDy, dy - 1) : -inf;
  const float inz = dz > 0 ? seq_at(r.tMz, r.tDz, dz - 1) : -inf;
  if (!(inx < inf) || !(iny < inf) || !(inz < inf)) return false;
  const float ox = dx > 0 ? inx + r.tDx : r.tMx;
  const float oy = dy > 0 ? iny + r.tDy : r.tMy;
  const float oz = dz > 0 ? inz + r.tDz : r.tMz;
  return inx < oy && inx < oz && iny <= ox && iny < oz && inz <= ox &&
         inz <= oy;
}

// Tile-local index of hash coordinates in their own tile.
__device__ __forceinline__ int tile_local(int cx, int cy, int cz) {
  return ((cx % TX) * TY + cy % TY) * TZ + cz % TZ;
}

// Far field one step at a time from `step` (r: the walk's state there):
// one segment per maximal run of in-window voxels inside one tile span,
// cut at the wrap seam. Used when the sensor voxel is outside the window
// and for the last span of a walk that passes its end voxel by.
template <typename Emit>
__device__ __forceinline__ void far_by_step(const IntegrateArgs &g,
                                            RayState &r, int step,
                                            Emit &emit) {
  const RayArgs &a = g.a;
  HashCoord hc = hash_coord_of(a, r.ix, r.iy, r.iz);
  int cur = -1, open_step = 0, open_li = 0;
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  for (; step < r.max_steps; ++step) {
    if (r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) break;
    const bool inwin = abs(r.ix - a.oix) <= a.hx &&
                       abs(r.iy - a.oiy) <= a.hy && abs(r.iz - a.oiz) <= a.hz;
    const int t = inwin ? tile_of(g, hc.cx, hc.cy, hc.cz) : -1;
    if (t != cur) {
      if (cur >= 0) emit(ox, oy, oz, open_li, step - open_step, cur);
      if (t >= 0) {
        open_step = step;
        open_li = tile_local(hc.cx, hc.cy, hc.cz);
        ox = r.tMx;
        oy = r.tMy;
        oz = r.tMz;
      }
      cur = t;
    }
    const int ax = ray_advance(r);
    const int before = hc.cx + hc.cy + hc.cz;
    hash_coord_step(hc, ax, r, a);
    if (abs(hc.cx + hc.cy + hc.cz - before) != 1 && cur >= 0) {  // seam
      emit(ox, oy, oz, open_li, step + 1 - open_step, cur);
      cur = -1;
    }
  }
  if (cur >= 0) emit(ox, oy, oz, open_li, step - open_step, cur);
}

// Far field from `step` when the sensor voxel is inside the window (r is
// inside too and not at its end voxel), one tile span per iteration. The
// walk is monotone per axis, so each axis reaches the exit of the span
// (tile span or window) after a known number of its own steps; the exit
// comes with the first of those exit steps in merge order, and the other
// axes' steps before it are counted on their running sums. The exit axis
// is guessed from rounded estimates and confirmed exactly. The end voxel
// can only be met at step lend; the span holding that step is decided
// exactly by end_reached, or walked one step at a time when it is passed.
template <typename Emit>
__device__ __forceinline__ void far_by_span(const IntegrateArgs &g,
                                            RayState &r, int step, int lend,
                                            int wx, int wy, int wz,
                                            Emit &emit) {
  const RayArgs &a = g.a;
  const float inf = __builtin_inff();
  const HashCoord hc = window_hash_coord(g, r.ix, r.iy, r.iz);
  int t0x = hc.cx - hc.cx % TX, t0y = hc.cy - hc.cy % TY,
      t0z = hc.cz - hc.cz % TZ;
  int lx = hc.cx - t0x, ly = hc.cy - t0y, lz = hc.cz - t0z;
  int bx = exit_coord(r.ix, r.sx, hc.cx, t0x, span_width(t0x, a.mx, TX), wx);
  int by = exit_coord(r.iy, r.sy, hc.cy, t0y, span_width(t0y, a.my, TY), wy);
  int bz = exit_coord(r.iz, r.sz, hc.cz, t0z, span_width(t0z, a.mz, TZ), wz);
  while (true) {
    const int li = (lx * TY + ly) * TZ + lz;
    const int tile = span_tile(g, t0x, t0y, t0z);
    const int nbx = r.sx != 0 ? abs(bx - r.ix) : kBig;
    const int nby = r.sy != 0 ? abs(by - r.iy) : kBig;
    const int nbz = r.sz != 0 ? abs(bz - r.iz) : kBig;
    const float gxe = nbx == 1 ? r.tMx : r.tMx + static_cast<float>(nbx - 1) * r.tDx;
    const float gye = nby == 1 ? r.tMy : r.tMy + static_cast<float>(nby - 1) * r.tDy;
    const float gze = nbz == 1 ? r.tMz : r.tMz + static_cast<float>(nbz - 1) * r.tDz;
    int ax = (gze <= gxe && gze <= gye) ? 2 : (gye <= gxe ? 1 : 0);
    float e = ax == 0 ? seq_at(r.tMx, r.tDx, nbx - 1)
                      : (ax == 1 ? seq_at(r.tMy, r.tDy, nby - 1)
                                 : seq_at(r.tMz, r.tDz, nbz - 1));
    float hx = r.tMx, hy = r.tMy, hz = r.tMz;
    int cx = count_before(hx, r.tDx, ax == 0 ? 0 : nbx, e, false);
    int cy = count_before(hy, r.tDy, ax == 1 ? 0 : nby, e, ax == 0);
    int cz = count_before(hz, r.tDz, ax == 2 ? 0 : nbz, e, true);
    if ((ax != 0 && cx == nbx) || (ax != 1 && cy == nby) ||
        (ax != 2 && cz == nbz)) {
      // Wrong guess: exact exit parameters of every axis.
      const float ex = r.sx != 0 ? seq_at(r.tMx, r.tDx, nbx - 1) : inf;
      const float ey = r.sy != 0 ? seq_at(r.tMy, r.tDy, nby - 1) : inf;
      const float ez = r.sz != 0 ? seq_at(r.tMz, r.tDz, nbz - 1) : inf;
      ax = (ez <= ex && ez <= ey) ? 2 : (ey <= ex ? 1 : 0);
      e = ax == 0 ? ex : (ax == 1 ? ey : ez);
      hx = r.tMx;
      hy = r.tMy;
      hz = r.tMz;
      cx = count_before(hx, r.tDx, ax == 0 ? 0 : nbx, e, false);
      cy = count_before(hy, r.tDy, ax == 1 ? 0 : nby, e, ax == 0);
      cz = count_before(hz, r.tDz, ax == 2 ? 0 : nbz, e, true);
    }
    if (ax == 0) {
      cx = nbx;
      hx = e + r.tDx;
    } else if (ax == 1) {
      cy = nby;
      hy = e + r.tDy;
    } else {
      cz = nbz;
      hz = e + r.tDz;
    }
    const int n = cx + cy + cz;  // voxels of this span
    if (lend <= step + n) {
      const int tn = lend - step;
      const int dx = abs(r.eix - r.ix), dy = abs(r.eiy - r.iy),
                dz = abs(r.eiz - r.iz);
      if (dx + dy + dz == tn && end_reached(r, dx, dy, dz)) {
        if (tn > 0) emit(r.tMx, r.tMy, r.tMz, li, tn, tile);
      } else {
        far_by_step(g, r, step, emit);
      }
      return;
    }
    emit(r.tMx, r.tMy, r.tMz, li, n, tile);
    if ((ax == 0 && bx == wx) || (ax == 1 && by == wy) || (ax == 2 && bz == wz)) {
      return;  // the next voxel is outside the window
    }
    r.ix += r.sx * cx;
    r.iy += r.sy * cy;
    r.iz += r.sz * cz;
    lx += r.sx * cx;
    ly += r.sy * cy;
    lz += r.sz * cz;
    r.tMx = hx;
    r.tMy = hy;
    r.tMz = hz;
    step += n;
    if (ax == 0) {
      t0x = next_span(t0x, r.sx, a.mx, TX);
      const int tw = span_width(t0x, a.mx, TX);
      bx = r.sx > 0 ? min(r.ix + tw, wx) : max(r.ix - tw, wx);
      lx = r.sx > 0 ? 0 : tw - 1;
    } else if (ax == 1) {
      t0y = next_span(t0y, r.sy, a.my, TY);
      const int tw = span_width(t0y, a.my, TY);
      by = r.sy > 0 ? min(r.iy + tw, wy) : max(r.iy - tw, wy);
      ly = r.sy > 0 ? 0 : tw - 1;
    } else {
      t0z = next_span(t0z, r.sz, a.mz, TZ);
      const int tw = span_width(t0z, a.mz, TZ);
      bz = r.sz > 0 ? min(r.iz + tw, wz) : max(r.iz - tw, wz);
      lz = r.sz > 0 ? 0 : tw - 1;
    }
  }
}

// Value of v in lane k of this lane's quad (lanes 4j..4j+3): a DPP quad
// permute, exact for any 32-bit value. Every lane of the quad is active.
template <int k>
__device__ __forceinline__ int quad_get(int v) {
  return __builtin_amdgcn_update_dpp(0, v, k | (k << 2) | (k << 4) | (k << 6), 0xF, 0xF,
                                     false);
}
template <int k>
__device__ __forceinline__ float quad_get(float v) {
  return __int_as_float(quad_get<k>(__float_as_int(v)));
}

// v of axis q (x, y, z; q = 3 takes z), chosen by value: a conditional over
// members would select their addresses and keep the walk state in memory.
template <typename T>
__device__ __forceinline__ T pick3(int q, T x, T y, T z) {
  T v = z;
  if (q == 0) {
    v = x;
  } else if (q == 1) {
    v = y;
  }
  return v;
}

// far_by_span with the ray's quad of lanes sharing the per-axis work: lane
// q < 3 owns axis q (its exit parameter, its steps before the span exit,
// its end-voxel parameters), lane 3 stands by; the quad exchanges the
// per-axis results and every lane then holds the same walk state and takes
// the same branches. Same spans, same records, same order as far_by_span
// (every axis's exit parameter is formed exactly, so no guess is needed).
template <typename Emit>
__device__ __forceinline__ void far_by_span_quad(const IntegrateArgs &g,
                                                 RayState &r, int step, int lend,
                                                 int wx, int wy, int wz, int q,
                                                 Emit &emit) {
  const RayArgs &a = g.a;
  const float inf = __builtin_inff();
  const HashCoord hc = window_hash_coord(g, r.ix, r.iy, r.iz);
  int t0x = hc.cx - hc.cx % TX, t0y = hc.cy - hc.cy % TY,
      t0z = hc.cz - hc.cz % TZ;
  int lx = hc.cx - t0x, ly = hc.cy - t0y, lz = hc.cz - t0z;
  int bx = exit_coord(r.ix, r.sx, hc.cx, t0x, span_width(t0x, a.mx, TX), wx);
  int by = exit_coord(r.iy, r.sy, hc.cy, t0y, span_width(t0y, a.my, TY), wy);
  int bz = exit_coord(r.iz, r.sz, hc.cz, t0z, span_width(t0z, a.mz, TZ), wz);
  // This lane's axis (lane 3: an axis that never moves).
  const int ms = q == 3 ? 0 : pick3(q, r.sx, r.sy, r.sz);
  const float md = pick3(q, r.tDx, r.tDy, r.tDz);
  while (true) {
    const int li = (lx * TY + ly) * TZ + lz;
    const int tile = span_tile(g, t0x, t0y, t0z);
    const int nbx = r.sx != 0 ? abs(bx - r.ix) : kBig;
    const int nby = r.sy != 0 ? abs(by - r.iy) : kBig;
    const int nbz = r.sz != 0 ? abs(bz - r.iz) : kBig;
    const int mnb = pick3(q, nbx, nby, nbz);
    const float mh0 = pick3(q, r.tMx, r.tMy, r.tMz);
    const float me = ms != 0 ? seq_at(mh0, md, mnb - 1) : inf;
    const float ex = quad_get<0>(me), ey = quad_get<1>(me), ez = quad_get<2>(me);
    const int ax = (ez <= ex && ez <= ey) ? 2 : (ey <= ex ? 1 : 0);
    const float e = ax == 0 ? ex : (ax == 1 ? ey : ez);
    float mh = mh0;
    int mc = 0;
    if (q == ax) {
      mc = mnb;
      mh = e + md;
    } else if (ms != 0) {
      // z before y before x on ties: z always first, y first against x.
      mc = count_before(mh, md, mnb, e, q == 2 || (q == 1 && ax == 0));
    }
    const int cx = quad_get<0>(mc), cy = quad_get<1>(mc), cz = quad_get<2>(mc);
    const float hx = quad_get<0>(mh), hy = quad_get<1>(mh), hz = quad_get<2>(mh);
    const int n = cx + cy + cz;  // voxels of this span
    if (lend <= step + n) {
      const int tn = lend - step;
      const int dx = abs(r.eix - r.ix), dy = abs(r.eiy - r.iy),
                dz = abs(r.eiz - r.iz);
      bool reached = false;
      if (dx + dy + dz == tn) {
        // end_reached, its three running sums one per lane.
        const int md_steps = q == 3 ? 0 : pick3(q, dx, dy, dz);
        const float min_ = md_steps > 0 ? seq_at(mh0, md, md_steps - 1) : -inf;
        const float mout = md_steps > 0 ? min_ + md : mh0;
        const float inx = quad_get<0>(min_), iny = quad_get<1>(min_), inz = quad_get<2>(min_);
        const float ox = quad_get<0>(mout), oy = quad_get<1>(mout), oz = quad_get<2>(mout);
        reached = inx < inf && iny < inf && inz < inf && inx < oy && inx < oz &&
                  iny <= ox && iny < oz && inz <= ox && inz <= oy;
      }
      if (reached) {
        if (tn > 0) emit(r.tMx, r.tMy, r.tMz, li, tn, tile);
      } else {
        far_by_step(g, r, step, emit);
      }
      return;
    }
    emit(r.tMx, r.tMy, r.tMz, li, n, tile);
    if ((ax == 0 && bx == wx) || (ax == 1 && by == wy) || (ax == 2 && bz == wz)) {
      return;  // the next voxel is outside the window
    }
    r.ix += r.sx * cx;
    r.iy += r.sy * cy;
    r.iz += r.sz * cz;
    lx += r.sx * cx;
    ly += r.sy * cy;
    lz += r.sz * cz;
    r.tMx = hx;
    r.tMy = hy;
    r.tMz = hz;
    step += n;
    if (ax == 0) {
      t0x = next_span(t0x, r.sx, a.mx, TX);
      const int tw = span_width(t0x, a.mx, TX);
      bx = r.sx > 0 ? min(r.ix + tw, wx) : max(r.ix - tw, wx);
      lx = r.sx > 0 ? 0 : tw - 1;
    } else if (ax == 1) {
      t0y = next_span(t0y, r.sy, a.my, TY);
      const int tw = span_width(t0y, a.my, TY);
      by = r.sy > 0 ? min(r.iy + tw, wy) : max(r.iy - tw, wy);
      ly = r.sy > 0 ? 0 : tw - 1;
    } else {
      t0z = next_span(t0z, r.sz, a.mz, TZ);
      const int tw = span_width(t0z, a.mz, TZ);
      bz = r.sz > 0 ? min(r.iz + tw, wz) : max(r.iz - tw, wz);
      lz = r.sz > 0 ? 0 : tw - 1;
    }
  }
}

// ----------------------------------------------------------------------
// Grids of one launch
// ----------------------------------------------------------------------
// A launch updates one grid or a batch of grids. Rays, segment slots,
// tiles and near-field cells of all grids live in one index space each;
// every grid owns a contiguous range of them (its *_base fields). A
// workgroup's rays, a work item's tile and a near cell belong to one grid,
// found by binary search over the bases.
struct GridDesc {
  IntegrateArgs g;
  const float *xs, *ys, *zs;
  float *occ;
  int *export_op, *export_hit;
  int transform;       // points are sensor-frame, world = T * p
  int ray_base;        // first ray; a multiple of kRayAlign
  long long slot_base; // first segment slot; record j of ray i at slot_base + j * n + i
  int tile_base;       // first tile
  int near_base;       // first near-field cell
  // Changed-voxel stream (optional, delta_count non-null): the (hash, new
  // value) of every voxel whose stored value changes is appended; the
  // count counts all of them, entries past delta_cap are not written.
  unsigned *delta_count;
  int *delta_idx;
  float *delta_val;
  unsigned delta_cap;
  FastDiv n_div;       // division by g.n (a record slot's ray)
};

static_assert(sizeof(GridDesc) == 304, "descriptor size (documented)");

struct BatchHeader {  // device-resident in batch mode
  int n_grids;
  int n_rays;   // end of the ray range (padded)
  int n_tiles;  // tiles of all grids
  int n_near;   // near-field cells of all grids
  int chunk;    // records per hot-tile work item
  int lanes;    // lanes per ray of the walk: 1 or 4
  int shared;   // hot-tile chunks add into shared slabs (global atomics)
  int cube;     // largest near-field cube edge of the batch's grids (the
                // walk's LDS cube is sized by it)
};

struct Batch {
  GridDesc one;              // the grid of a single-grid launch
  const GridDesc *descs;     // batch: n_grids descriptors (device)
  const BatchHeader *hdr;    // batch: counts (device)
};

// The descriptors and the header stay unchanged while a launch runs, so
// they are read through the constant address space: scalar loads into
// SGPRs, not vector loads into every lane's registers.
template <typename T>
__device__ __forceinline__ int const_int(const T *p, int word) {
  return ((const CUBLOX_CONST int *)(p))[word];
}
__device__ __forceinline__ int batch_word(const Batch &b, int word) {
  return const_int(b.hdr, word);
}
__device__ __forceinline__ int batch_grids(const Batch &b) {
  return b.descs ? batch_word(b, 0) : 1;
}
__device__ __forceinline__ int batch_rays(const Batch &b) {
  return b.descs ? batch_word(b, 1) : b.one.ray_base + b.one.g.n;
}
__device__ __forceinline__ int batch_tiles(const Batch &b) {
  return b.descs ? batch_word(b, 2) : b.one.g.n_tiles;
}
__device__ __forceinline__ int batch_near(const Batch &b) {
  return b.descs ? batch_word(b, 3) : b.one.g.cube * b.one.g.cube * b.one.g.cube;
}
__device__ __forceinline__ int batch_chunk(const Batch &b, const Workspace &w) {
  return b.descs ? batch_word(b, 4) : w.chunk;
}
// Batches carry the walk's lanes per ray and the hot-tile mode in the
// header (so a captured launch sequence follows them); a single-grid launch
// runs only the walk it was launched with and the chunk slabs.
__device__ __forceinline__ bool batch_walk_is(const Batch &b, int lanes) {
  return !b.descs || batch_word(b, 5) == lanes;
}
__device__ __forceinline__ bool batch_shared(const Batch &b) {
  return b.descs && batch_word(b, 6) != 0;
}
static_assert(offsetof(BatchHeader, n_rays) == 4 && offsetof(BatchHeader, chunk) == 16 &&
                  offsetof(BatchHeader, lanes) == 20 && offsetof(BatchHeader, shared) == 24,
              "BatchHeader words");

// Last grid whose base (the int field at byte offset `off`) is <= key.
// Keys are uniform across the calling workgroup.
__device__ __forceinline__ GridDesc grid_at(const Batch &b, int key, size_t off) {
  if (!b.descs) return b.one;
  int lo = 0, hi = batch_grids(b) - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (const_int(b.descs + mid, static_cast<int>(off / 4)) <= key) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return load_const(b.descs + lo);
}

__device__ __forceinline__ GridDesc grid_of_ray(const Batch &b, int ray) {
  return grid_at(b, ray, offsetof(GridDesc, ray_base));
}
__device__ __forceinline__ GridDesc grid_of_tile(const Batch &b, int tile) {
  return grid_at(b, tile, offsetof(GridDesc, tile_base));
}
__device__ __forceinline__ GridDesc grid_of_near(const Batch &b, int cell) {
  return grid_at(b, cell, offsetof(GridDesc, near_base));
}

// Finds (inserting if needed) the LDS slot of `key` in a kSize-entry
// table; -1 when the bounded probe finds neither. A key is either in the
// table for every caller or for none: slots only fill, so a full probe
// window stays full.
template <int kSize>
__device__ __forceinline__ int lds_hash_slot_n(int *keys, int key, bool insert) {
  unsigned h = (static_cast<unsigned>(key) * 2654435761u) >> 20;
  for (int p = 0; p < kHashProbes; ++p, ++h) {
    h &= kSize - 1;
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

__device__ __forceinline__ int lds_hash_slot(int *keys, int key, bool insert) {
  return lds_hash_slot_n<kHash>(keys, key, insert);
}

// ----------------------------------------------------------------------
// walk kernel
// ----------------------------------------------------------------------
constexpr int kLenBuckets = 64;

// A pending recenter's slab clear that a single-grid walk launch carries:
// its last `chunks` workgroups zero the regions (and the counters, when
// the grid has them) while the others walk. Passed by value.
constexpr int kInlineRegions = 16;
struct InlineClear {
  float *occ;
  int *op, *hit;
  int n;
  long long chunks;
  ChunkedRegion r[kInlineRegions];
};

// Workgroup j of an inline clear (every thread of the workgroup calls it).
__device__ __forceinline__ void inline_clear(const InlineClear &z, long long j) {
  int k = 0;  // the region holding chunk j (chunk0 ascending)
  while (k + 1 < z.n && z.r[k + 1].chunk0 <= j) ++k;
  const ZeroRegion part = chunk_part(z.r[k], j - z.r[k].chunk0);
  zero_region(as_global(z.occ), part, 0, 1);
  if (z.op) zero_region(as_global(z.op), part, 0, 1);
  if (z.hit) zero_region(as_global(z.hit), part, 0, 1);
}

// Length class of point i for the in-block ordering: its range in voxels
// (clipped to the max range), 8 voxels per class. Any value is correct;
// similar classes in one wave only shorten the longest lane's walk.
__device__ __forceinline__ int length_class(const GridDesc &d, int i) {
  const RayArgs &a = d.g.a;
  const size_t at = static_cast<size_t>(i) * static_cast<size_t>(a.stride);
  float x = as_global(d.xs)[at], y = as_global(d.ys)[at], z = as_global(d.zs)[at];
  if (!d.transform) {  // world frame: relative to the sensor
    x -= a.ox;
    y -= a.oy;
    z -= a.oz;
  }
  const float r = sqrtf(x * x + y * y + z * z);
  if (!(r < a.max_range)) return kLenBuckets - 1;  // clipped, or not finite
  return min(kLenBuckets - 1, static_cast<int>(r * a.inv_res) >> 3);
}

// kLanes lanes per ray: 1, or 4 (a quad shares each span's per-axis work,
// for launches too small to fill the GPU with one lane per ray). All lanes
// of a ray hold the same walk state; lane 0 writes its records and counts.
// The block's rays are taken in order of length class (a counting sort in
// LDS), so each wave's rays walk similar distances; every ray's records are
// its own, whichever lane walks it.
template <int kLanes>
__global__ void __launch_bounds__(kWalkBlock * kLanes, kLanes == 1 ? 8 : 1)
    walk_kernel(const Batch b, Workspace w, const InlineClear zc) {
  // The last zc.chunks workgroups clear a pending recenter's slabs.
  const long long clear0 = static_cast<long long>(gridDim.x) - zc.chunks;
  if (static_cast<long long>(blockIdx.x) >= clear0) {
    inline_clear(zc, static_cast<long long>(blockIdx.x) - clear0);
    return;
  }
  extern __shared__ int s_near[];  // the grid's near cube, sized at launch
  __shared__ int s_hist[kLenBuckets];
  __shared__ unsigned short s_perm[kWalkBlock];
  // Records per tile of this workgroup (the binning's counts: one global
  // atomic per distinct tile of the workgroup at its end). Only the quad
  // walk counts: its launches are small, and its workgroups' few tiles fit
  // the table; the one-lane walk of large batches leaves the counts to a
  // binning pass.
  constexpr bool kCount = kLanes == 4;
  __shared__ int s_tkey[kCount ? kWalkHash : 1];
  __shared__ int s_tcnt[kCount ? kWalkHash : 1];
  if (!batch_walk_is(b, kLanes)) return;  // a batch tuned to the other walk
  const int ray0 = blockIdx.x * kWalkBlock;
  if (blockIdx.x == 0) {
    // The changed-voxel counts start from zero (the folds run later).
    const int n_grids = batch_grids(b);
    for (int k = threadIdx.x; k < n_grids; k += blockDim.x) {
      unsigned *c = b.descs ? load_const(b.descs + k).delta_count : b.one.delta_count;
      if (c) *c = 0u;
    }
  }
  if (ray0 >= batch_rays(b)) return;
  const GridDesc d = grid_of_ray(b, ray0);
  const IntegrateArgs &g = d.g;
  const int cells = g.cube * g.cube * g.cube;
  for (int c = threadIdx.x; c < cells; c += blockDim.x) s_near[c] = 0;
  if (kCount) {
    for (int e = threadIdx.x; e < kWalkHash; e += blockDim.x) {
      s_tkey[e] = -1;
      s_tcnt[e] = 0;
    }
  }
  if (threadIdx.x < kLenBuckets) s_hist[threadIdx.x] = 0;
  __syncthreads();
  {
    const int t = threadIdx.x;
    const int n_block = min(kWalkBlock, g.n - (ray0 - d.ray_base));
    int cls = -1;
    if (t < kWalkBlock) {
      cls = t < n_block ? length_class(d, ray0 - d.ray_base + t) : kLenBuckets - 1;
      atomicAdd(s_hist + cls, 1);
    }
    __syncthreads();
    if (t < 64) {  // one wavefront: exclusive scan of the classes
      const int v = s_hist[t];
      s_hist[t] = wave_scan_incl(v) - v;
    }
    __syncthreads();
    if (cls >= 0) s_perm[atomicAdd(s_hist + cls, 1)] = static_cast<unsigned short>(t);
    __syncthreads();
  }

  const int q = static_cast<int>(threadIdx.x) % kLanes;
  const bool lead = q == 0;
  const int ray = ray0 + s_perm[static_cast<int>(threadIdx.x) / kLanes];
  const int i = ray - d.ray_base;  // index within the grid's points
  const RayArgs &a = g.a;
  RayState r;
  const bool valid =
      i < g.n &&
      (d.transform
           ? ray_setup<true>(a, as_global(d.xs), as_global(d.ys), as_global(d.zs), i, r)
           : ray_setup<false>(a, as_global(d.xs), as_global(d.ys), as_global(d.zs), i, r));
  if (valid) {
    const unsigned signs =
        static_cast<unsigned>((r.sx + 1) | ((r.sy + 1) << 2) | ((r.sz + 1) << 4))
        << kMetaSign;
    // Record j of ray i sits at slot slot_base + j * n + i: the j-th records
    // of neighbouring rays, which mostly fall in one tile, share lines.
    const long long slot0 = d.slot_base + i;
    const long long sstride = g.n;
    Segment *__restrict__ mine = w.segs + slot0;
    int *__restrict__ mine_tile = w.seg_tile + slot0;
    unsigned *__restrict__ mine_rank = w.seg_rank + slot0;
    int nseg = 0;
    auto put = [&](const Segment &sg, int tile) {
      if (!lead) {
      } else if (nseg < g.max_seg) {
        const int key = d.tile_base + tile;
        mine[nseg * sstride] = sg;
        mine_tile[nseg * sstride] = key;
        if (kCount) {
          const int e = lds_hash_slot_n<kWalkHash>(s_tkey, key, true);
          mine_rank[nseg * sstride] =
              e >= 0 ? rank_code(e, atomicAdd(reinterpret_cast<unsigned *>(s_tcnt) + e, 1u))
                     : kRankGlobal | atomic_add_u32_ret(
                                         reinterpret_cast<unsigned *>(w.tile_count) + key, 1u);
        }
      } else {
        w.misc[2] = 1;
      }
      ++nseg;
    };
    auto emit = [&](float tmx, float tmy, float tmz, int li, int n, int tile) {
      Segment sg;
      sg.tMx = tmx;
      sg.tMy = tmy;
      sg.tMz = tmz;
      sg.meta = static_cast<unsigned>(li) |
                (static_cast<unsigned>(n) << kMetaN) | signs;
      put(sg, tile);
    };
    const int s0x = r.ix, s0y = r.iy, s0z = r.iz;  // sensor voxel
    const int K = g.near_k;
    const int near_end = min(K, r.max_steps);
    int step = 0;
    if (inside_window(s0x, s0y, s0z, a.oix, a.oiy, a.oiz, a.hx, a.hy, a.hz)) {
      // Each step moves one axis one voxel in its travel direction, so the
      // walk is monotone per axis: step s sits at L1 distance s from the
      // sensor voxel, the end voxel can only be met at step lend, and once
      // an axis reaches its window exit coordinate no later voxel is inside.
      const int lend = abs(r.eix - s0x) + abs(r.eiy - s0y) + abs(r.eiz - s0z);
      const int wx = window_exit(r.sx, a.oix, a.hx);
      const int wy = window_exit(r.sy, a.oiy, a.hy);
      const int wz = window_exit(r.sz, a.oiz, a.hz);
      const int C = g.cube;
      const int cdx = r.sx * C * C, cdy = r.sy * C, cdz = r.sz;
      int ci = cube_index(g, 0, 0, 0);
      bool done = false;
      for (; step < near_end; ++step) {
        if ((step == lend && r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) ||
            r.ix == wx || r.iy == wy || r.iz == wz) {
          done = true;
          break;
        }
        if (lead) atomicAdd(s_near + ci, 1);
        const int ax = ray_advance(r);
        ci += ax == 0 ? cdx : (ax == 1 ? cdy : cdz);
      }
      if (!done && step < r.max_steps &&
          !(step == lend && r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) &&
          !(r.ix == wx || r.iy == wy || r.iz == wz)) {
        if (kLanes == 4) {
          far_by_span_quad(g, r, step, lend, wx, wy, wz, q, emit);
        } else {
          far_by_span(g, r, step, lend, wx, wy, wz, emit);
        }
      }
    } else {
      // Sensor outside the window: the ray may enter it later.
      bool ended = false;
      for (; step < near_end; ++step) {
        if (r.ix == r.eix && r.iy == r.eiy && r.iz == r.eiz) {
          ended = true;
          break;
        }
        if (lead && inside_window(r.ix, r.iy, r.iz, a.oix, a.oiy, a.oiz, a.hx, a.hy,
                                  a.hz)) {
          atomicAdd(s_near + cube_index(g, r.ix - s0x, r.iy - s0y, r.iz - s0z), 1);
        }
        ray_advance(r);
      }
      if (!ended) far_by_step(g, r, step, emit);
    }
    // Endpoint (kernels.cu:141-151): op += 1, hit += 1 unless clipped.
    if (inside_window(r.eix, r.eiy, r.eiz, a.oix, a.oiy, a.oiz, a.hx, a.hy,
                      a.hz)) {
      const int ddx = r.eix - s0x, ddy = r.eiy - s0y, ddz = r.eiz - s0z;
      if (abs(ddx) + abs(ddy) + abs(ddz) < K) {
        const int c = cube_index(g, ddx, ddy, ddz);
        if (lead) {
          atomicAdd(s_near + c, 1);
          if (!r.clipped) atomic_add_i32(w.near_hit + d.near_base + c, 1);
        }
      } else {
        const HashCoord e = window_hash_coord(g, r.eix, r.eiy, r.eiz);
        Segment sg;
        sg.tMx = sg.tMy = sg.tMz = 0.0f;
        sg.meta = static_cast<unsigned>(tile_local(e.cx, e.cy, e.cz)) |
                  kMetaEndpoint | (r.clipped ? kMetaClipped : 0u);
        put(sg, tile_of(g, e.cx, e.cy, e.cz));
      }
    }
    if (lead) {
      w.seg_count[ray] = min(nseg, g.max_seg);
      w.ray_td[ray] = make_float4(r.tDx, r.tDy, r.tDz, 0.0f);
    }
  } else if (lead) {
    w.seg_count[ray] = 0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < cells; c += blockDim.x) {
    const int v = s_near[c];
    if (v) atomic_add_i32(w.near_op + d.near_base + c, v);
  }
  if (kCount) {
    // The group's first place in each of its tiles.
    int *base = w.grp_base + static_cast<size_t>(blockIdx.x) * kWalkHash;
    for (int e = threadIdx.x; e < kWalkHash; e += blockDim.x) {
      const int key = s_tkey[e];
      if (key >= 0) {
        base[e] = static_cast<int>(atomic_add_u32_ret(
            reinterpret_cast<unsigned *>(w.tile_count) + key, static_cast<unsigned>(s_tcnt[e])));
      }
    }
  }
}

// ----------------------------------------------------------------------
// binning: counts per tile after the one-lane walk, and the placing pass
// ----------------------------------------------------------------------

// A block counts the records of `rays` consecutive rays (a divisor of
// kRayAlign, so all of one grid; the block is the records' counting group).
// Record j of ray i lives at slot slot_base + j * n + i, so the block walks
// its records j-major: iteration x is record x / rays of ray x % rays
// (skipped when that ray has fewer records), and consecutive lanes read
// consecutive slots. Counts aggregate per tile in an LDS hash: one global
// atomic per distinct tile of the block (records that find no hash slot
// take their place from the global count directly). Each record's rank
// among the block's records of its tile goes to seg_rank, the block's base
// in each of its tiles to grp_base.
__global__ void __launch_bounds__(kBinThreads)
    bin_kernel(const Batch b, Workspace w, int rays) {
  __shared__ int s_key[kHash];
  __shared__ unsigned s_cnt[kHash];
  __shared__ unsigned short s_rc[kRayAlign];  // records per ray
  __shared__ int s_max;
  if (batch_walk_is(b, 4) && b.descs) return;  // the quad walk counted
  const int r0 = blockIdx.x * rays;
  const int n_rays = batch_rays(b);
  if (r0 >= n_rays) return;
  const GridDesc d = grid_of_ray(b, r0);
  const int t = threadIdx.x;
  for (int e = t; e < kHash; e += kBinThreads) {
    s_key[e] = -1;
    s_cnt[e] = 0;
  }
  if (t == 0) s_max = 0;
  __syncthreads();
  int mx = 0;
  for (int r = t; r < rays; r += kBinThreads) {
    const int c = r0 + r < n_rays ? w.seg_count[r0 + r] : 0;
    s_rc[r] = static_cast<unsigned short>(c);
    mx = max(mx, c);
  }
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off, 64));
  if ((t & 63) == 0 && mx > 0) atomicMax(&s_max, mx);
  __syncthreads();
  const int iters = s_max * rays;
  const int i0 = r0 - d.ray_base;  // the block's first ray within its grid
  const int stride = d.g.n;
  const int lg_rays = 31 - __clz(rays);  // rays is a power of two
  // Slot of iteration x relative to the grid's first slot (below the
  // grid's n * max_seg <= INT_MAX slots), or -1.
  auto rel_slot = [&](int x) -> int {
    const int j = x >> lg_rays, r = x & (rays - 1);
    if (j >= static_cast<int>(s_rc[r])) return -1;
    return j * stride + (i0 + r);
  };
  auto count = [&](int rel) {
    const size_t slot = d.slot_base + rel;
    const int tile = w.seg_tile[slot];
    const int e = lds_hash_slot(s_key, tile, true);
    w.seg_rank[slot] =
        e >= 0 ? rank_code(e, atomicAdd(s_cnt + e, 1u))
               : kRankGlobal | atomic_add_u32_ret(
                                   reinterpret_cast<unsigned *>(w.tile_count) + tile, 1u);
  };
  // The first kBinCache iterations of each thread are unrolled so that
  // their loads are in flight together.
  int c_off[kBinCache];
#pragma unroll
  for (int i = 0; i < kBinCache; ++i) {
    const int x = t + i * kBinThreads;
    c_off[i] = x < iters ? rel_slot(x) : -1;
    if (c_off[i] >= 0) count(c_off[i]);
  }
  for (int x = t + kBinCache * kBinThreads; x < iters; x += kBinThreads) {
    const int rel = rel_slot(x);
    if (rel >= 0) count(rel);
  }
  __syncthreads();
  int *base = w.grp_base + static_cast<size_t>(blockIdx.x) * kHash;
  for (int e = t; e < kHash; e += kBinThreads) {
    const int key = s_key[e];
    if (key < 0) continue;
    base[e] = static_cast<int>(
        atomic_add_u32_ret(reinterpret_cast<unsigned *>(w.tile_count) + key, s_cnt[e]));
  }
}

// ----------------------------------------------------------------------
// tile scan: cursors, work items, hot-tile splits
// ----------------------------------------------------------------------
// A tile with more than `chunk` records is split into chunks of about
// `chunk` records, each its own work item with its own scratch slab, while
// the split tiles' slabs (taken in tile order) fit the scratch area; a hot
// tile beyond that is one item, accumulated in 16-bit passes. Items are
// placed in decreasing cost classes (records per item) so the heaviest
// start first. The scan runs over blocks of tiles: per-block sums (pass 1),
// per-block bases (pass 2), then each block's cursors and items (pass 3).
constexpr int kScanThreads = 1024;
constexpr int kScanWaves = kScanThreads / 64;
constexpr int kScanTilesPerBlock = 4096;  // tiles per scan block (large workspaces)
// Workspaces of at most kScanSmall tiles scan in one pass of blocks of
// kScanSmallTiles tiles (one thread each; at most 48 blocks, all resident
// at once): each block publishes its sums and sums every block's itself.
// Measured at 9,680 tiles: 10.0 us, against 4.8 + 7.0 us as a sums launch
// and a cursors launch, and 13.6 us for one 1,024-thread workgroup over
// all tiles, whose instruction issue on one CU bounded it.
constexpr int kScanSmall = 12288;
constexpr int kScanSmallTiles = 256;
constexpr int kScanSums = 3 + kCostBuckets;  // records, slabs, hot tiles, classes
constexpr int kPass = 65535;  // records per LDS pass (16-bit counts)
constexpr int kBigBlocks = 256;  // workgroups over unsplit hot tiles


// Cost class of a work item (its record count), larger classes first:
// above 512 records, above 128, above 32, the rest. Items are placed class
// by class so the heaviest start first; within a class in any order.
__device__ __forceinline__ int cost_bucket(int records) {
  return records > 512 ? 0 : (records > 128 ? 1 : (records > 32 ? 2 : 3));
}

// Number of chunks a tile with c records is split into (1: not split).
__device__ __forceinline__ int chunks_of(int c, int chunk) {
  return c > chunk ? (c + chunk - 1) / chunk : 1;
}

// Inclusive scans over the block's threads of N ints per thread at once
// (one exchange of wave totals for all N); tot receives the block totals.
// s_wave holds N * kScanWaves ints. Every thread of the block takes part.
template <int N>
__device__ __forceinline__ void block_scan_incl_n(int *s_wave, int (&v)[N], int (&tot)[N]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_scan_incl(v[q]);
  if (lane == 63) {
#pragma unroll
    for (int q = 0; q < N; ++q) s_wave[q * kScanWaves + wv] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    int before = 0, all = 0;
    for (int k = 0; k < nw; ++k) {
      const int x = s_wave[q * kScanWaves + k];
      before += k < wv ? x : 0;
      all += x;
    }
    v[q] += before;
    tot[q] = all;
  }
  __syncthreads();  // s_wave is reused by the next scan
}

// Items of tile k (c records from `run`; its slab range would start at sl
// and its split index is sp): ch chunk items when split, else one item.
// The tile's `slots` item positions start at pos0; those an unsplit hot
// tile leaves unused become empty items.
__device__ __forceinline__ void place_tile(const Workspace &w, int k, int c,
                                           long long run, int sl, int sp,
                                           int chunk, bool split, int slots,
                                           int pos0, bool shared) {
  if (c <= 0) return;
  const int ch = split ? chunks_of(c, chunk) : 1;
  const int per_chunk = split ? (c + ch - 1) / ch : c;
  if (split) w.hot[sp] = HotTile{k, sl, ch, shared ? 1 : 0};
  for (int q = 0; q < slots; ++q) {
    const int pos = pos0 + q;
    if (pos >= w.max_items) break;
    if (q < ch) {
      const int ib = static_cast<int>(run) + q * per_chunk;
      const int ie = min(ib + per_chunk, static_cast<int>(run) + c);
      const bool big = !split && c > chunk;  // hot tile without slabs
      w.items[pos] = Item{k, ib, ie, split ? (shared ? sl : sl + q) : (big ? -2 : -1)};
      if (big) w.big[atomicAdd(w.misc + 4, 1)] = pos;
    } else {
      w.items[pos] = Item{-1, 0, 0, -1};
    }
  }
}

// Takes `slots` item positions for an item of class q from the thread's
// class cursors (registers; the loop unrolls to selects).
__device__ __forceinline__ int take_slots(int (&cur)[kCostBuckets], int q, int slots) {
  int pos = 0;
#pragma unroll
  for (int u = 0; u < kCostBuckets; ++u) {
    if (q == u) {
      pos = cur[u];
      cur[u] += slots;
    }
  }
  return pos;
}

// Multi-workgroup scan, pass 1: per block of tpb tiles, the records, the
// slabs its hot tiles would take, its hot tiles, and the items per cost
// class as if every hot tile were split.
__global__ void __launch_bounds__(kScanThreads)
    tile_sums_kernel(const Batch b, Workspace w, int tpb) {
  __shared__ int s_sum[kScanSums];
  const int n_tiles = batch_tiles(b);
  const int chunk = batch_chunk(b, w);
  const bool shared = batch_shared(b);
  const int k0 = blockIdx.x * tpb;
  if (threadIdx.x < kScanSums) s_sum[threadIdx.x] = 0;
  // The split and unsplit hot-tile counters that pass 3 adds to.
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.misc[3] = 0;
    w.misc[4] = 0;
  }
  __syncthreads();
  int v[kScanSums] = {};  // records, slabs, hot tiles, classes
  for (int k = k0 + threadIdx.x; k < min(k0 + tpb, n_tiles);
       k += blockDim.x) {
    const int c = w.tile_count[k];
    v[0] += c;
    const int ch = chunks_of(c, chunk);
    if (c > 0) {
      const int q = cost_bucket((c + ch - 1) / ch);
#pragma unroll
      for (int u = 0; u < kCostBuckets; ++u) v[3 + u] += q == u ? ch : 0;
    }
    if (ch > 1) {
      v[1] += slab_need(ch, shared);
      v[2] += 1;
    }
  }
#pragma unroll
  for (int q = 0; q < kScanSums; ++q) {
    int x = v[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0 && x != 0) atomicAdd(s_sum + q, x);
  }
  __syncthreads();
  if (threadIdx.x < kScanSums) {
    w.block_sums[blockIdx.x * kScanSums + threadIdx.x] = s_sum[threadIdx.x];
  }
}

// Pass 2 (one workgroup): exclusive per-block bases of records and slabs,
// and for every cost class the block's first item slot.
__global__ void __launch_bounds__(kScanThreads)
    tile_bases_kernel(const Batch b, Workspace w, int tpb) {
  __shared__ int s_wave[kScanSums * kScanWaves];
  const int n_blocks = (batch_tiles(b) + tpb - 1) / tpb;
  const int t = threadIdx.x;  // one thread per block (at most kScanThreads)
  int v[kScanSums], tot[kScanSums];
#pragma unroll
  for (int q = 0; q < kScanSums; ++q) v[q] = t < n_blocks ? w.block_sums[t * kScanSums + q] : 0;
  int own[kScanSums];
#pragma unroll
  for (int q = 0; q < kScanSums; ++q) own[q] = v[q];
  block_scan_incl_n<kScanSums>(s_wave, v, tot);
  int start = 0;
#pragma unroll
  for (int q = 0; q < kCostBuckets; ++q) {
    v[3 + q] += start;  // first slot of class q after the earlier classes
    start += tot[3 + q];
  }
  if (t < n_blocks) {
#pragma unroll
    for (int q = 0; q < kScanSums; ++q) w.block_bases[t * kScanSums + q] = v[q] - own[q];
  }
  if (t == 0) {
    w.misc[0] = min(start, w.max_items);
    w.misc[1] = tot[0];
  }
}

// Pass 2 done by a pass-3 block itself (at most 64 blocks): its bases (as
// tile_bases_kernel gives them) from every block's sums, summed by the first
// wave, into s_base; block 0 also writes the item and record totals. Every
// thread of the block calls this.
static_assert((kScanSmall + kScanSmallTiles - 1) / kScanSmallTiles <= 64,
              "one wave sums the block sums of a small scan");
// With `published`, the sums come from the blocks of this same launch:
// lane t waits (a bounded spin) for block t's flag, then reads its sums
// with agent-scope loads.
constexpr int kScanSpinLimit = 1 << 22;
__device__ __forceinline__ void own_block_bases(const Batch &b, const Workspace &w, int tpb,
                                                int *s_base, bool published) {
  const int n_blocks = (batch_tiles(b) + tpb - 1) / tpb;
  const int t = threadIdx.x;
  const int me = static_cast<int>(blockIdx.x);
  if (t < 64) {
    if (published && t < n_blocks) {
      int spins = 0;
      while (__hip_atomic_load(w.scan_flag + t, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
             0) {
        if (++spins > kScanSpinLimit) {  // never expected: flag it and go on
          __hip_atomic_store(w.misc + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    int pre[kScanSums], tot[kScanSums];
#pragma unroll
    for (int q = 0; q < kScanSums; ++q) {
      int x = 0;
      if (t < n_blocks) {
        x = published ? __hip_atomic_load(w.block_sums + t * kScanSums + q, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                      : w.block_sums[t * kScanSums + q];
      }
      pre[q] = t < me ? x : 0;
      tot[q] = x;
    }
#pragma unroll
    for (int q = 0; q < kScanSums; ++q) {
      for (int off = 32; off > 0; off >>= 1) {
        pre[q] += __shfl_xor(pre[q], off, 64);
        tot[q] += __shfl_xor(tot[q], off, 64);
      }
    }
    if (t == 0) {
    int start = 0;
#pragma unroll
    for (int q = 0; q < kCostBuckets; ++q) {
      s_base[3 + q] = start + pre[3 + q];  // first slot of class q in this block
      start += tot[3 + q];
    }
    s_base[0] = pre[0];
    s_base[1] = pre[1];
    s_base[2] = pre[2];
    if (me == 0) {
      w.misc[0] = min(start, w.max_items);
      w.misc[1] = tot[0];
    }
    }
  }
  __syncthreads();
}

// Pass 3: per block, cursors, items and split tiles from the block bases.
// A hot tile whose slab range does not fit keeps one item (in its class
// slot) and turns its other slots into empty items. The slab prefix only
// grows, so the tiles that fit are the first hot tiles in tile order: a
// split tile's index in the hot list is its hot-tile prefix, and the list's
// length the largest such index plus one.
__global__ void __launch_bounds__(kScanThreads)
    tile_place_kernel(const Batch b, Workspace w, int tpb, int own_bases) {
  // own_bases: 0, the bases pass's; 1, summed from the sums pass's; 2, one
  // pass (no sums launch): every block publishes its sums and sums them all.
  __shared__ int s_wave[kScanSums * kScanWaves];
  __shared__ int s_base[kScanSums];
  const int n_tiles = batch_tiles(b);
  const int chunk = batch_chunk(b, w);
  const bool shared = batch_shared(b);
  const int k0 = blockIdx.x * tpb;
  const int k1 = min(k0 + tpb, n_tiles);
  const int t = threadIdx.x;
  const int kPer = tpb / static_cast<int>(blockDim.x);
  const int beg = min(k0 + t * kPer, k1);
  const int end = min(beg + kPer, k1);
  const int c_first = beg < end ? w.tile_count[beg] : 0;
  // Records, slabs, hot tiles and item slots per class (ch each) of the
  // thread's tiles, scanned over the block at once.
  int v[kScanSums] = {}, tv[kScanSums];
  for (int k = beg; k < end; ++k) {
    const int c = k == beg ? c_first : w.tile_count[k];
    const int ch = chunks_of(c, chunk);
    v[0] += c;
    if (c > 0) {
      const int q = cost_bucket((c + ch - 1) / ch);
#pragma unroll
      for (int u = 0; u < kCostBuckets; ++u) v[3 + u] += q == u ? ch : 0;
    }
    if (ch > 1) {
      v[1] += slab_need(ch, shared);
      v[2] += 1;
    }
  }
  int own[kScanSums];
#pragma unroll
  for (int q = 0; q < kScanSums; ++q) own[q] = v[q];
  block_scan_incl_n<kScanSums>(s_wave, v, tv);
  const int *base = w.block_bases + blockIdx.x * kScanSums;
  if (own_bases) {  // uniform
    if (own_bases == 2 && threadIdx.x == 0) {
      // One pass: this block's sums (its scan totals) published for the
      // others; block 0 first clears the hot-list counters they add to.
      if (blockIdx.x == 0) {
        __hip_atomic_store(w.misc + 3, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(w.misc + 4, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int q = 0; q < kScanSums; ++q) {
        __hip_atomic_store(w.block_sums + blockIdx.x * kScanSums + q, tv[q], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
      __hip_atomic_store(w.scan_flag + blockIdx.x, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    own_block_bases(b, w, tpb, s_base, own_bases == 2);
    base = s_base;
  }
  int cur[kCostBuckets];
#pragma unroll
  for (int u = 0; u < kCostBuckets; ++u) cur[u] = base[3 + u] + v[3 + u] - own[3 + u];
  long long run = static_cast<long long>(base[0]) + v[0] - own[0];
  int sl = base[1] + v[1] - own[1];
  int sp = base[2] + v[2] - own[2];
  for (int k = beg; k < end; ++k) {
    const int c = k == beg ? c_first : w.tile_count[k];
    const int ch = chunks_of(c, chunk);
    const bool split = ch > 1 && sl + slab_need(ch, shared) <= w.slabs_cap;
    const int pos0 = c > 0 ? take_slots(cur, cost_bucket((c + ch - 1) / ch), ch) : 0;
    w.tile_cursor[k] = static_cast<int>(run);
    place_tile(w, k, c, run, sl, sp, chunk, split, ch, pos0, shared);
    if (split) atomicMax(w.misc + 3, sp + 1);
    if (ch > 1) {
      sl += slab_need(ch, shared);
      ++sp;
    }
    run += c;
    w.tile_count[k] = 0;
  }
}

// Placing pass (after the tile scan has set each tile's run start in
// tile_cursor): record j of every ray of this block's kWalkBlock rays goes
// to its place in tile order, kPlaceJ values of j at a time: the block of y
// index y takes j = kPlaceJ * y + kPlaceJ * gridDim.y * m (m = 0, 1, ...)
// while the ray has records there. The record's counting group is its ray's
// group of 2^lg_group rays, whose hash has `hash` entries. Every record is
// independent, so the pass is a gather and a scatter with no atomics.
constexpr int kPlaceJ = 4;
constexpr int kPlaceRows = 4;
__global__ void __launch_bounds__(kWalkBlock)
    place_kernel(const Batch b, Workspace w, int lg_group, int hash, int flags) {
  // The one-pass scan's flags, zero again for the next launch (the scan
  // has finished).
  if (blockIdx.x == 0 && blockIdx.y == 0 && static_cast<int>(threadIdx.x) < flags) {
    w.scan_flag[threadIdx.x] = 0;
  }
  const int bx = static_cast<int>(blockIdx.x);
  const int ray = bx * kWalkBlock + static_cast<int>(threadIdx.x);
  const int ray0 = bx * kWalkBlock;
  const int n_rays = batch_rays(b);
  if (ray0 >= n_rays) return;
  const int cnt = ray < n_rays ? w.seg_count[ray] : 0;
  if (static_cast<int>(blockIdx.y) * kPlaceJ >= cnt) return;
  const GridDesc d = grid_of_ray(b, ray0);
  const long long slot0 = d.slot_base + (ray - d.ray_base);
  const long long stride = d.g.n;
  const int *gbase = w.grp_base + static_cast<size_t>(ray >> lg_group) * hash;
  for (int j0 = static_cast<int>(blockIdx.y) * kPlaceJ; j0 < cnt;
       j0 += kPlaceJ * static_cast<int>(gridDim.y)) {
    int tile[kPlaceJ];
    unsigned rk[kPlaceJ];
#pragma unroll
    for (int q = 0; q < kPlaceJ; ++q) {
      if (j0 + q < cnt) {
        tile[q] = w.seg_tile[slot0 + (j0 + q) * stride];
        rk[q] = w.seg_rank[slot0 + (j0 + q) * stride];
      }
    }
    // The records' run starts and group bases, all in flight together.
    int run[kPlaceJ], base[kPlaceJ];
#pragma unroll
    for (int q = 0; q < kPlaceJ; ++q) {
      if (j0 + q < cnt) {
        run[q] = w.tile_cursor[tile[q]];
        base[q] = (rk[q] & kRankGlobal) ? 0 : gbase[rk[q] >> kRankBits];
      }
    }
#pragma unroll
    for (int q = 0; q < kPlaceJ; ++q) {
      if (j0 + q >= cnt) continue;
      const long long pos =
          static_cast<long long>(run[q]) +
          ((rk[q] & kRankGlobal)
               ? static_cast<long long>(rk[q] & ~kRankGlobal)
               : static_cast<long long>(base[q]) +
                     static_cast<long long>(rk[q] & ((1u << kRankBits) - 1u)));
      if (pos < w.seg_capacity) {
        w.order[pos] = static_cast<unsigned>(slot0 + (j0 + q) * stride);
      }
    }
  }
}


// ----------------------------------------------------------------------
// tile_apply
// ----------------------------------------------------------------------

// Replays one record (of a ray with increments td) into the packed LDS
// counters.
__device__ __forceinline__ void replay_record(const Segment &sg, const float4 &td,
                                              unsigned *s_cnt) {
  int li = static_cast<int>(sg.meta & kMetaLi);
  if (sg.meta & kMetaEndpoint) {
    atomicAdd(s_cnt + li, (sg.meta & kMetaClipped) ? 1u : 0x10001u);
    return;
  }
  int n = static_cast<int>((sg.meta >> kMetaN) & 0x7Fu);
  const int sx = static_cast<int>((sg.meta >> kMetaSign) & 3u) - 1;
  const int sy = static_cast<int>((sg.meta >> (kMetaSign + 2)) & 3u) - 1;
  const int sz = static_cast<int>((sg.meta >> (kMetaSign + 4)) & 3u) - 1;
  const int ldx = sx * (TY * TZ), ldy = sy * TZ, ldz = sz;
  float tMx = sg.tMx, tMy = sg.tMy, tMz = sg.tMz;
  for (; n > 0; --n) {
    atomicAdd(s_cnt + li, 1u);
    const bool x_lt_y = tMx < tMy;
    const bool step_x = x_lt_y && (tMx < tMz);
    const bool step_y = !x_lt_y && (tMy < tMz);
    const float nx = tMx + td.x;
    const float ny = tMy + td.y;
    const float nz = tMz + td.z;
    tMx = step_x ? nx : tMx;
    tMy = step_y ? ny : tMy;
    tMz = (!step_x && !step_y) ? nz : tMz;
    li += step_x ? ldx : (step_y ? ldy : ldz);
  }
}

// Accumulates records [beg, end) of a tile into packed LDS counters
// (op | hit << 16): each record replays its n voxels from its entry state
// (the same crossing parameters and tie-break as the walk), one record per
// thread at a time. A call covers at most kPass records and a record visits
// a voxel at most once, so neither field carries into the other.
__device__ __forceinline__ void lds_zero(uint4 *s4);

// The grid fields that map a record slot to its ray: slot = slot_base +
// j * n + i for record j of ray i (global ray ray_base + i).
struct SlotRays {
  long long slot_base;
  int ray_base, n;
  FastDiv n_div;
};
__device__ __forceinline__ SlotRays slot_rays(const GridDesc &d) {
  return SlotRays{d.slot_base, d.ray_base, d.g.n, d.n_div};
}
__device__ __forceinline__ int ray_of_slot(const SlotRays &sr, unsigned slot) {
  const unsigned rel = static_cast<unsigned>(static_cast<long long>(slot) - sr.slot_base);
  const unsigned j = fast_div(rel, sr.n_div);
  return sr.ray_base + static_cast<int>(rel - j * static_cast<unsigned>(sr.n));
}

// Clears the tile's counters and accumulates records [beg, end) into them.
// Each thread's first record (and its ray's increments) is gathered before
// the clear, so those loads' latency overlaps it.
__device__ __forceinline__ void tile_accumulate(const Workspace &w, const SlotRays &sr,
                                                int beg, int end, uint4 *s_cnt4) {
  unsigned *s_cnt = reinterpret_cast<unsigned *>(s_cnt4);
  const int k0 = beg + static_cast<int>(threadIdx.x);
  Segment first{};
  float4 td0{};
  if (k0 < end) {
    const unsigned o = w.order[k0];
    first = w.segs[o];
    td0 = w.ray_td[ray_of_slot(sr, o)];
  }
  lds_zero(s_cnt4);
  __syncthreads();
  if (end - beg <= 2 * kApplyThreads) {  // uniform: at most two records a thread
    if (k0 < end) replay_record(first, td0, s_cnt);
    for (int k = k0 + kApplyThreads; k < end; k += kApplyThreads) {
      const unsigned o = w.order[k];
      replay_record(w.segs[o], w.ray_td[ray_of_slot(sr, o)], s_cnt);
    }
  } else {
    // Longer items (hot-tile chunks) pipeline a thread's records: while
    // record k replays, record k + kApplyThreads is in flight and the slot
    // index after it has arrived, so no replay waits on a dependent pair
    // of loads.
    Segment cur = first;
    float4 tcur = td0;
    unsigned o_next = k0 + kApplyThreads < end ? w.order[k0 + kApplyThreads] : 0u;
    for (int k = k0; k < end; k += kApplyThreads) {
      const int k1 = k + kApplyThreads;
      Segment nxt{};
      float4 tnxt{};
      if (k1 < end) {
        nxt = w.segs[o_next];
        tnxt = w.ray_td[ray_of_slot(sr, o_next)];
      }
      const unsigned o_after =
          k1 + kApplyThreads < end ? w.order[k1 + kApplyThreads] : 0u;
      replay_record(cur, tcur, s_cnt);
      cur = nxt;
      tcur = tnxt;
      o_next = o_after;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void lds_zero(uint4 *s4) {
  for (int k = threadIdx.x; k < kTileVox / 4; k += kApplyThreads) {
    s4[k] = make_uint4(0, 0, 0, 0);
  }
}

// Hash offset of the grid voxel at tile-local index k of the tile whose
// hash corner is (x0, y0, z0).
__device__ __forceinline__ int tile_voxel_hash(const RayArgs &a, int x0,
                                               int y0, int z0, int k) {
  const int lz = k % TZ, ly = (k / TZ) % TY, lx = k / (TZ * TY);
  return (x0 + lx) * (a.my * a.mz) + (y0 + ly) * a.mz + (z0 + lz);
}

// ----------------------------------------------------------------------
// Changed-voxel stream (the reference's modified-voxel compaction,
// kernels.cu:202-209, per fused update): a wave reserves the slots of all
// its changed voxels with one atomic on the grid's count, then writes them
// voxel-index-major, so the lanes of one index write one run.
// ----------------------------------------------------------------------
__device__ __forceinline__ unsigned lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), 0u));
}

__device__ __forceinline__ void delta_put(const GridDesc &d, unsigned slot,
                                          int h, float v) {
  if (slot < d.delta_cap) {
    as_global(d.delta_idx)[slot] = h;
    as_global(d.delta_val)[slot] = v;
  }
}

// Appends the voxels of this thread marked in `changed` (bit j: hash(j),
// new value v[j], j < kN). Every thread of the workgroup (at most four
// waves) takes part, d is the same in all of them, and s_tmp holds 8 ints
// of LDS. The workgroup reserves its slots with one atomic: the grid's
// count is one address that every workgroup of the launch hits.
template <int kN, typename Hash>
__device__ __forceinline__ void delta_append(const GridDesc &d,
                                             unsigned changed, const float *v,
                                             Hash hash, int *s_tmp) {
  int total = __popc(changed);
  for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off, 64);
  const int wv = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  if (__lane_id() == 0) s_tmp[wv] = total;
  __syncthreads();
  int before = 0, all = 0;
  for (int k = 0; k < nw; ++k) {
    const int c = s_tmp[k];
    before += k < wv ? c : 0;
    all += c;
  }
  if (threadIdx.x == 0 && all > 0) {
    s_tmp[4] = static_cast<int>(atomic_add_u32_ret(d.delta_count, static_cast<unsigned>(all)));
  }
  __syncthreads();
  unsigned base = static_cast<unsigned>(s_tmp[4]) + static_cast<unsigned>(before);
  __syncthreads();  // s_tmp is reused by the next call
  if (total == 0) return;  // wave-uniform
#pragma unroll
  for (int j = 0; j < kN; ++j) {
    const bool on = (changed >> j) & 1u;
    const unsigned long long m = __ballot(on);
    if (on) delta_put(d, base + lanes_below(m), hash(j), v[j]);
    base += static_cast<unsigned>(__popcll(m));
  }
}

// Fold of the tile counts a thread owns (voxels j * 256 + t, j < 32: a
// z-run of 32 contiguous floats per 32 lanes): every occ load is issued
// before the first update, so a tile costs one memory round trip. Counts
// come from op(j) and hit(j). kDelta: voxels whose value changes go to the
// grid's changed-voxel stream when it has one.
template <bool kDelta, typename Op, typename Hit>
__device__ __forceinline__ void fold_owned(const GridDesc &d, int x0, int y0,
                                           int z0, Op op, Hit hit, int *s_tmp) {
  const RayArgs &a = d.g.a;
  const int t = threadIdx.x;
  // hash(j) = hbase (this thread) + joff(j) (the same in every lane), so
  // each access is a uniform base plus one per-thread offset.
  const int hbase = (x0 * a.my + y0 + (t >> 5)) * a.mz + z0 + (t & 31);
  const int plane = a.my * a.mz;
  auto joff = [&](int j) { return (j >> 1) * plane + ((j & 1) << 3) * a.mz; };
  auto hash = [&](int j) { return hbase + joff(j); };
  if (d.export_op) {
    CUBLOX_GLOBAL int *eop = as_global(d.export_op);
    CUBLOX_GLOBAL int *ehit = as_global(d.export_hit);
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      if (op(j) == 0) continue;
      (eop + joff(j))[hbase] = static_cast<int>(op(j));
      (ehit + joff(j))[hbase] = static_cast<int>(hit(j));
    }
    return;
  }
  CUBLOX_GLOBAL float *occ = as_global(d.occ);
  float v[kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    v[j] = op(j) != 0 ? (occ + joff(j))[hbase] : 0.0f;
  }
  if (kDelta && d.delta_count) {
    unsigned changed = 0;
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      if (op(j) == 0) continue;
      const float nv = apply_logodds(v[j], static_cast<int>(op(j)),
                                     static_cast<int>(hit(j)), d.g.p);
      (occ + joff(j))[hbase] = nv;
      if (nv != v[j]) changed |= 1u << j;
      v[j] = nv;
    }
    delta_append<kPerThread>(d, changed, v, hash, s_tmp);
    return;
  }
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    if (op(j) == 0) continue;
    (occ + joff(j))[hbase] = apply_logodds(v[j], static_cast<int>(op(j)),
                                           static_cast<int>(hit(j)), d.g.p);
  }
}

// Work items in decreasing cost order, one per workgroup while they last
// (then strided), so the hardware dispatcher balances them. kBig = false:
// the plain items (split chunks and whole tiles); kBig = true: the unsplit
// hot tiles, in passes of kPass records summed in registers (a separate
// kernel, so the common one keeps a register budget for 5 workgroups per
// CU).
// A plain item accumulates its tile in LDS and folds it; a chunk of a hot
// tile stores its packed counts to its scratch slab (hot_fold sums a
// tile's slabs); an unsplit hot tile accumulates in passes of kPass
// records summed in registers. Empty items do nothing.
template <bool kBig, bool kDelta>
__global__ void __launch_bounds__(kApplyThreads)
    tile_apply_kernel(const Batch b, Workspace w) {
  __shared__ uint4 s_cnt4[kTileVox / 4];
  __shared__ int s_tmp[8];
  unsigned *s_cnt = reinterpret_cast<unsigned *>(s_cnt4);
  constexpr bool big = kBig;
  // Items, their count and the hot list are fixed while this kernel runs:
  // read them through the constant address space (scalar loads), so the
  // item's tile, its grid and every address derived from them are uniform.
  const int n_all = const_int(w.misc, kBig ? 4 : 0);
  for (int it = blockIdx.x; it < n_all; it += gridDim.x) {
    const Item itm = load_const(w.items + (kBig ? const_int(w.big, it) : it));
    if (itm.tile < 0 || (!kBig && itm.slab == -2)) continue;  // uniform
    const int end = static_cast<int>(
        min(static_cast<long long>(itm.end), w.seg_capacity));
    const GridDesc d = grid_of_tile(b, itm.tile);
    const SlotRays sr = slot_rays(d);
    if (!big && itm.slab >= 0) {
      tile_accumulate(w, sr, itm.beg, end, s_cnt4);
      if (batch_shared(b)) {
        // Shared mode: the chunk's nonzero counts go into the tile's op and
        // hit slabs (zeroed before this kernel) with global atomics.
        unsigned *op_slab = w.scratch + static_cast<size_t>(itm.slab) * kTileVox;
        unsigned *hit_slab = op_slab + kTileVox;
        for (int k = threadIdx.x; k < kTileVox; k += kApplyThreads) {
          const unsigned c = s_cnt[k];
          if (c == 0) continue;
          atomic_add_u32(op_slab + k, c & 0xFFFFu);
          if (c >> 16) atomic_add_u32(hit_slab + k, c >> 16);
        }
      } else {
        uint4 *slab = reinterpret_cast<uint4 *>(
            w.scratch + static_cast<size_t>(itm.slab) * kTileVox);
        for (int k = threadIdx.x; k < kTileVox / 4; k += kApplyThreads) {
          slab[k] = s_cnt4[k];
        }
      }
      __syncthreads();
      continue;
    }
    const int tile = itm.tile - d.tile_base;
    const int tz = tile % d.g.ntz;
    const int ty = (tile / d.g.ntz) % d.g.nty;
    const int tx = tile / (d.g.ntz * d.g.nty);
    if (!big) {
      // One pass: the packed 16-bit counts fold as they are.
      tile_accumulate(w, sr, itm.beg, end, s_cnt4);
      unsigned c[kPerThread];
#pragma unroll
      for (int j = 0; j < kPerThread; ++j) c[j] = s_cnt[j * kApplyThreads + threadIdx.x];
      __syncthreads();
      fold_owned<kDelta>(d, tx * TX, ty * TY, tz * TZ,
                 [&](int j) { return c[j] & 0xFFFFu; },
                 [&](int j) { return c[j] >> 16; }, s_tmp);
    } else {
      unsigned op[kPerThread], hit[kPerThread];
#pragma unroll
      for (int j = 0; j < kPerThread; ++j) {
        op[j] = 0;
        hit[j] = 0;
      }
      for (int p0 = itm.beg; p0 < end; p0 += kPass) {
        tile_accumulate(w, sr, p0, min(p0 + kPass, end), s_cnt4);
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
          const unsigned c = s_cnt[j * kApplyThreads + threadIdx.x];
          op[j] += c & 0xFFFFu;
          hit[j] += c >> 16;
        }
        __syncthreads();
      }
      fold_owned<kDelta>(d, tx * TX, ty * TY, tz * TZ, [&](int j) { return op[j]; },
                 [&](int j) { return hit[j]; }, s_tmp);
    }
  }
}

// Shared hot-tile mode: zeroes the op and hit slabs of this launch's split
// tiles before the chunks add into them (a batch may have run the chunk-slab
// mode before, which leaves counts in every slab it used).
__global__ void __launch_bounds__(256) hot_zero_kernel(const Batch b, Workspace w) {
  if (!batch_shared(b)) return;
  const int n_hot = w.misc[3];
  constexpr int kWords4 = kSharedSlabs * kTileVox / 4;  // uint4 per hot tile
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < static_cast<long long>(n_hot) * kWords4;
       e += static_cast<long long>(gridDim.x) * 256) {
    const int h = static_cast<int>(e / kWords4);
    const int k = static_cast<int>(e % kWords4);
    const HotTile ht = load_const(w.hot + h);
    reinterpret_cast<uint4 *>(w.scratch + static_cast<size_t>(ht.slab0) * kTileVox)[k] =
        make_uint4(0u, 0u, 0u, 0u);
  }
}

template <bool kDelta>
__device__ __forceinline__ void near_fold(const Batch &b, const Workspace &w, int blk,
                                          int n_blk);

// Sums the scratch slabs of each split tile and folds: kHotSlices
// workgroups per tile, four consecutive voxels (one z-run piece) per lane.
// Workgroups from hot_blocks on fold the near-field cubes instead (both
// touch disjoint voxels and come after tile_apply: one launch for both).
template <bool kDelta>
__global__ void __launch_bounds__(256)
    hot_fold_kernel(const Batch b, Workspace w, int hot_blocks) {
  __shared__ int s_tmp[8];
  if (static_cast<int>(blockIdx.x) >= hot_blocks) {
    near_fold<kDelta>(b, w, static_cast<int>(blockIdx.x) - hot_blocks,
                      static_cast<int>(gridDim.x) - hot_blocks);
    return;
  }
  const int n_hot = w.misc[3];
  for (int blk = blockIdx.x; blk < n_hot * kHotSlices; blk += hot_blocks) {
    const HotTile ht = load_const(w.hot + blk / kHotSlices);
    const int k4 = (blk % kHotSlices) * 256 + threadIdx.x;  // < 2048
    const uint4 *src =
        reinterpret_cast<const uint4 *>(w.scratch +
                                        static_cast<size_t>(ht.slab0) * kTileVox) +
        k4;
    unsigned op[4] = {0, 0, 0, 0}, hit[4] = {0, 0, 0, 0};
    int q = ht.atomic ? ht.chunks : 0;  // shared mode: no chunk slabs to sum
    if (ht.atomic) {
      const uint4 o4 = src[0];
      const uint4 h4 = src[kTileVox / 4];
      op[0] = o4.x; op[1] = o4.y; op[2] = o4.z; op[3] = o4.w;
      hit[0] = h4.x; hit[1] = h4.y; hit[2] = h4.z; hit[3] = h4.w;
    }
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
    const GridDesc d = grid_of_tile(b, ht.tile);
    const RayArgs &a = d.g.a;
    const int tile = ht.tile - d.tile_base;
    const int tz = tile % d.g.ntz;
    const int ty = (tile / d.g.ntz) % d.g.nty;
    const int tx = tile / (d.g.ntz * d.g.nty);
    const int h0 = tile_voxel_hash(a, tx * TX, ty * TY, tz * TZ, 4 * k4);
    CUBLOX_GLOBAL float *occ = as_global(d.occ);
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[e] = (op[e] != 0 && !d.export_op) ? occ[h0 + e] : 0.0f;
    }
    unsigned changed = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (op[e] == 0) continue;
      if (d.export_op) {
        as_global(d.export_op)[h0 + e] = static_cast<int>(op[e]);
        as_global(d.export_hit)[h0 + e] = static_cast<int>(hit[e]);
      } else {
        const float nv = apply_logodds(v[e], static_cast<int>(op[e]),
                                       static_cast<int>(hit[e]), d.g.p);
        occ[h0 + e] = nv;
        if (kDelta && nv != v[e]) changed |= 1u << e;
        v[e] = nv;
      }
    }
    if (kDelta && d.delta_count) {
      delta_append<4>(d, changed, v, [&](int e) { return h0 + e; }, s_tmp);
    }
  }
}

// ----------------------------------------------------------------------
// near fold: the near-field cubes into occ, cleared (workgroup blk of
// n_blk, 256 threads)
// ----------------------------------------------------------------------
template <bool kDelta>
__device__ __forceinline__ void near_fold(const Batch &b, const Workspace &w, int blk,
                                          int n_blk) {
  const int n_near = batch_near(b);
  for (int cell = blk * 256 + static_cast<int>(threadIdx.x); cell < n_near;
       cell += n_blk * 256) {
    const int o = w.near_op[cell];
    if (o == 0) continue;
    const int hit = w.near_hit[cell];
    w.near_op[cell] = 0;
    w.near_hit[cell] = 0;
    const GridDesc d = grid_of_near(b, cell);
    const IntegrateArgs &g = d.g;
    const RayArgs &a = g.a;
    const bool ctr = a.center != 0;
    const int c = cell - d.near_base;
    const int C = g.cube;
    const int k1 = g.near_k - 1;
    const int dz = c % C - k1;
    const int dy = (c / C) % C - k1;
    const int dx = c / (C * C) - k1;
    const int h = global_to_hash(pos_to_global(a.ox, a.inv_res, ctr) + dx,
                                 pos_to_global(a.oy, a.inv_res, ctr) + dy,
                                 pos_to_global(a.oz, a.inv_res, ctr) + dz,
                                 a.mx, a.my, a.mz, a.hx, a.hy, a.hz);
    if (d.export_op) {
      as_global(d.export_op)[h] = o;
      as_global(d.export_hit)[h] = hit;
    } else {
      CUBLOX_GLOBAL float *occ = as_global(d.occ);
      const float v0 = occ[h];
      const float nv = apply_logodds(v0, o, hit, g.p);
      occ[h] = nv;
      // Cells of one wave may belong to different grids: one atomic each.
      if (kDelta && d.delta_count && nv != v0) {
        delta_put(d, atomic_add_u32_ret(d.delta_count, 1u), h, nv);
      }
    }
  }
}

// ----------------------------------------------------------------------
// Host side: geometry, workspace layout, launch sequence.
// ----------------------------------------------------------------------

inline size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct GridGeom {
  int ntx, nty, ntz, n_tiles, cube, max_seg;
};

// Per-ray segment bound (tile entries + the endpoint record): inside the
// window the walk is monotone per axis, so it changes tile at most
// ntx + nty + ntz times plus one wrap seam per axis; a walk of S steps also
// changes tile at most S / min(tile dim) + 3 times.
GridGeom grid_geom(const cublox_raycast_cfg &c, int near_k) {
  GridGeom G{};
  G.ntx = (c.map_size_i.x + TX - 1) / TX;
  G.nty = (c.map_size_i.y + TY - 1) / TY;
  G.ntz = (c.map_size_i.z + TZ - 1) / TZ;
  G.n_tiles = G.ntx * G.nty * G.ntz;
  G.cube = 2 * near_k - 1;
  const long long win_bound = G.ntx + G.nty + G.ntz + 8;
  const double reach = static_cast<double>(c.max_range) *
                           static_cast<double>(c.inv_resolution) + 2.0;
  long long range_bound = win_bound;
  if (reach < 1e7) {
    const long long steps = 3 * static_cast<long long>(reach) + 8;
    range_bound = steps / std::min(std::min(TX, TY), TZ) + 10;
  }
  G.max_seg = static_cast<int>(std::min(win_bound, range_bound));
  return G;
}

// Batch staging: a header and the descriptors, copied to the workspace in
// one transfer at the head of the launch sequence.
struct Staging {
  BatchHeader hdr;
  int pad[60];
  GridDesc desc[1];  // max_grids of them
};

size_t staging_bytes(int max_grids) {
  return offsetof(Staging, desc) + static_cast<size_t>(std::max(max_grids, 1)) * sizeof(GridDesc);
}

// Capacities of a workspace.
struct Caps {
  int max_grids;
  long long max_rays;      // padded ray range
  long long seg_capacity;  // segment slots
  int max_tiles;
  int max_cube;            // largest near cube edge
  int max_near;            // near cells
  int chunk;               // smallest records per hot-tile chunk
  int max_slabs;           // scratch slab cap
  int max_seg;             // record slots per ray
};

struct Layout {
  Caps cap;
  long long slabs;
  bool may_run_out;  // slabs below the chunk bound: hot tiles may stay unsplit
  int max_items, scan_blocks;
  int scan_tpb;  // tiles per multi-block scan block
  size_t off_sflag;
  size_t off_count, off_stile, off_segs, off_sorted, off_tcount, off_tcur,
      off_items, off_big, off_hot, off_scratch, off_misc, off_nop, off_nhit, off_bsum,
      off_bbase, off_srank, off_gbase, off_rtd, off_hdr, off_desc, total;
};

int bin_rays_per_block(const Caps &c);

// Entries of the counting groups' base table: the quad walk's groups
// (kWalkBlock rays, kWalkHash entries) or the binning count's (bin blocks,
// kHash entries), whichever a launch uses.
long long group_table_entries(const Caps &c) {
  const long long walk = (c.max_rays + kWalkBlock - 1) / kWalkBlock * kWalkHash;
  const int br = bin_rays_per_block(c);
  const long long bin = (c.max_rays + br - 1) / br * kHash;
  return std::max(walk, bin);
}

constexpr int kMinChunk = 256;

// Record slots are indexed by 32-bit unsigned integers in tile order.
// Record slots, work-item ranges and tile cursors are 32-bit signed
// indices: the workspace holds at most INT_MAX record slots.
bool caps_fit(const Caps &c) { return c.seg_capacity <= 0x7FFFFFFFll; }

Layout make_layout(const Caps &c) {
  Layout L{};
  L.cap = c;
  // Each hot tile has more than `chunk` records, so the chunks of all hot
  // tiles number at most 2 * segments / chunk: the item bound. Slabs are
  // capped; hot tiles past the cap run unsplit.
  const long long chunks = 2 * c.seg_capacity / std::max(c.chunk, kMinChunk) + 2;
  L.slabs = std::max<long long>(2, std::min<long long>(chunks, c.max_slabs));
  L.may_run_out = chunks > c.max_slabs;
  L.max_items = static_cast<int>(std::min<long long>(c.max_tiles + chunks, 1LL << 30));
  L.scan_tpb = c.max_tiles <= kScanSmall ? kScanSmallTiles : kScanTilesPerBlock;
  L.scan_blocks = (c.max_tiles + L.scan_tpb - 1) / L.scan_tpb;
  size_t o = 0;
  auto take = [&](size_t &off, size_t bytes) {
    off = o;
    o = align256(o + bytes);
  };
  take(L.off_count, static_cast<size_t>(c.max_rays) * 4);
  take(L.off_stile, static_cast<size_t>(c.seg_capacity) * 4);
  take(L.off_segs, static_cast<size_t>(c.seg_capacity) * sizeof(Segment));
  take(L.off_rtd, static_cast<size_t>(c.max_rays) * sizeof(float4));
  take(L.off_sorted, static_cast<size_t>(c.seg_capacity) * sizeof(unsigned));
  take(L.off_tcount, static_cast<size_t>(c.max_tiles) * 4);
  take(L.off_tcur, static_cast<size_t>(c.max_tiles) * 4);
  take(L.off_items, static_cast<size_t>(L.max_items) * sizeof(Item));
  take(L.off_big, static_cast<size_t>(c.max_tiles) * sizeof(int));
  take(L.off_hot, static_cast<size_t>(L.slabs) * sizeof(HotTile));
  take(L.off_scratch, static_cast<size_t>(L.slabs) * kTileVox * 4);
  take(L.off_misc, 64 * 4);
  take(L.off_sflag, static_cast<size_t>(L.scan_blocks) * 4);  // cleared with misc
  take(L.off_nop, static_cast<size_t>(c.max_near) * 4);
  take(L.off_nhit, static_cast<size_t>(c.max_near) * 4);
  take(L.off_bsum, static_cast<size_t>(L.scan_blocks) * kScanSums * 4);
  take(L.off_bbase, static_cast<size_t>(L.scan_blocks) * kScanSums * 4);
  take(L.off_srank, static_cast<size_t>(c.seg_capacity) * 4);
  take(L.off_gbase, static_cast<size_t>(group_table_entries(c)) * 4);
  // Header and descriptors, laid out as the host staging buffer so one
  // copy moves both.
  take(L.off_hdr, staging_bytes(c.max_grids));
  L.off_desc = L.off_hdr + offsetof(Staging, desc);
  L.total = o;
  return L;
}

Workspace bind(const Layout &L, void *base) {
  char *b = static_cast<char *>(base);
  Workspace w{};
  w.seg_count = reinterpret_cast<int *>(b + L.off_count);
  w.seg_tile = reinterpret_cast<int *>(b + L.off_stile);
  w.segs = reinterpret_cast<Segment *>(b + L.off_segs);
  w.ray_td = reinterpret_cast<float4 *>(b + L.off_rtd);
  w.order = reinterpret_cast<unsigned *>(b + L.off_sorted);
  w.tile_count = reinterpret_cast<int *>(b + L.off_tcount);
  w.tile_cursor = reinterpret_cast<int *>(b + L.off_tcur);
  w.items = reinterpret_cast<Item *>(b + L.off_items);
  w.big = reinterpret_cast<int *>(b + L.off_big);
  w.hot = reinterpret_cast<HotTile *>(b + L.off_hot);
  w.scratch = reinterpret_cast<unsigned *>(b + L.off_scratch);
  w.misc = reinterpret_cast<int *>(b + L.off_misc);
  w.near_op = reinterpret_cast<int *>(b + L.off_nop);
  w.near_hit = reinterpret_cast<int *>(b + L.off_nhit);
  w.block_sums = reinterpret_cast<int *>(b + L.off_bsum);
  w.scan_flag = reinterpret_cast<int *>(b + L.off_sflag);
  w.block_bases = reinterpret_cast<int *>(b + L.off_bbase);
  w.seg_rank = reinterpret_cast<unsigned *>(b + L.off_srank);
  w.grp_base = reinterpret_cast<int *>(b + L.off_gbase);
  w.seg_capacity = L.cap.seg_capacity;
  w.max_items = L.max_items;
  w.slabs_cap = static_cast<int>(L.slabs);
  w.chunk = L.cap.chunk;
  return w;
}

// Single-grid capacities.
Caps single_caps(const cublox_raycast_cfg &cfg, int max_points, int near_k) {
  const GridGeom G = grid_geom(cfg, near_k);
  Caps c{};
  c.max_grids = 1;
  c.max_rays = static_cast<long long>(std::max(max_points, 1) + kWalkBlock);
  c.seg_capacity = static_cast<long long>(std::max(max_points, 1)) * G.max_seg;
  c.max_tiles = G.n_tiles;
  c.max_cube = G.cube;
  c.max_near = G.cube * G.cube * G.cube;
  c.chunk = kChunk;
  c.max_slabs = 4096;
  c.max_seg = G.max_seg;
  return c;
}

// A changed-voxel stream: a count (zeroed by the caller) and, for a
// nonzero capacity, both entry arrays; not with a count export (no value
// is stored then).
int check_delta(const unsigned *count, const int *idx, const float *val,
                unsigned cap, bool exporting, const char *who) {
  if (!count) return CUBLOX_OK;
  if (exporting) {
    return set_error(CUBLOX_ERR_INVALID,
                     std::string(who) + ": a changed-voxel stream with a count export");
  }
  if (cap > 0 && (!idx || !val)) {
    return set_error(CUBLOX_ERR_INVALID,
                     std::string(who) + ": changed-voxel stream needs both arrays");
  }
  return CUBLOX_OK;
}

// Point layout: point_stride floats between consecutive points (0: 1,
// separate columns); a range drop applies to sensor-frame points only.
int check_input(const cublox_integrate_grid &q, const char *who) {
  if (q.point_stride < 0) {
    return set_error(CUBLOX_ERR_INVALID, std::string(who) + ": negative point stride");
  }
  if (q.range_drop_sq > 0.0f && !q.use_T) {
    return set_error(CUBLOX_ERR_INVALID,
                     std::string(who) + ": a range drop needs sensor-frame points");
  }
  return CUBLOX_OK;
}

void set_input(GridDesc &d, const cublox_integrate_grid &q) {
  d.g.a.stride = q.point_stride > 0 ? q.point_stride : 1;
  d.g.a.drop_sq = q.range_drop_sq > 0.0f ? q.range_drop_sq : 0.0f;
}

void set_delta(GridDesc &d, unsigned *count, int *idx, float *val, unsigned cap) {
  d.delta_count = count;
  d.delta_idx = count ? idx : nullptr;
  d.delta_val = count ? val : nullptr;
  d.delta_cap = count ? cap : 0u;
}

int check_cfg(const cublox_raycast_cfg &cfg, int near_k) {
  if (near_k < 1 || near_k > kMaxNearK) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: near_steps out of [1, 12]");
  }
  if (cfg.map_size_i.x != 2 * cfg.half_map_size_i.x + 1 ||
      cfg.map_size_i.y != 2 * cfg.half_map_size_i.y + 1 ||
      cfg.map_size_i.z != 2 * cfg.half_map_size_i.z + 1) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: map size must be 2 * half + 1");
  }
  return CUBLOX_OK;
}

// Fills the descriptor of one grid (bases set by the caller).
GridDesc make_desc(const cublox_raycast_cfg &cfg, const float *T,
                   const float *xs, const float *ys, const float *zs, int n,
                   float *occ, const ApplyArgs &p, int near_k, int *eop,
                   int *ehit) {
  const GridGeom G = grid_geom(cfg, near_k);
  GridDesc d{};
  IntegrateArgs &g = d.g;
  g.a = make_ray_args(cfg);
  if (T) {
    for (int k = 0; k < 12; ++k) g.a.T[k] = T[k];
    g.a.ox = T[3];
    g.a.oy = T[7];
    g.a.oz = T[11];
  }
  g.n = n;
  d.n_div = make_fastdiv(static_cast<unsigned>(std::max(n, 1)));
  g.ntx = G.ntx;
  g.nty = G.nty;
  g.ntz = G.ntz;
  g.n_tiles = G.n_tiles;
  g.near_k = near_k;
  g.cube = G.cube;
  g.max_seg = G.max_seg;
  auto pmod = [](int v, int m) {
    const int r = v % m;
    return r < 0 ? r + m : r;
  };
  g.qx = pmod(g.a.oix, g.a.mx);
  g.qy = pmod(g.a.oiy, g.a.my);
  g.qz = pmod(g.a.oiz, g.a.mz);
  g.p = p;
  d.xs = xs;
  d.ys = ys;
  d.zs = zs;
  d.occ = occ;
  d.export_op = eop;
  d.export_hit = ehit;
  d.transform = T ? 1 : 0;
  return d;
}

// Launches of at most this many rays walk with a quad of lanes per ray.
constexpr long long kQuadWalkRays = 1 << 18;

// Rays per bin workgroup: larger blocks share more tiles (fewer global
// atomics), smaller ones fill the GPU on small launches. A power of two
// dividing kRayAlign, from the capacities only.
int bin_rays_per_block(const Caps &c) {
  int r = 256;
  while (r < kRayAlign && c.max_rays / (2 * r) >= 1024) r *= 2;
  return r;
}

// The fold launches; kDelta: the variants that feed changed-voxel streams.
template <bool kDelta>
int launch_folds(const Batch &b, const Layout &L, const Workspace &w,
                 hipStream_t hs, bool shared) {
  const Caps &c = L.cap;
  if (shared) {  // a batch in the shared hot-tile mode
    hot_zero_kernel<<<dim3(2048), dim3(256), 0, hs>>>(b, w);
    const int st0 = check_launch("hot_zero_kernel");
    if (st0 != CUBLOX_OK) return st0;
  }
  const int apply_blocks = std::min(L.max_items, 1 << 16);
  tile_apply_kernel<false, kDelta><<<dim3(apply_blocks), dim3(kApplyThreads), 0, hs>>>(b, w);
  int st = check_launch("tile_apply_kernel");
  if (st != CUBLOX_OK) return st;
  // Hot tiles stay unsplit only when the scratch slabs run out; with
  // slabs for every possible chunk the pass is not launched.
  if (L.may_run_out) {
    tile_apply_kernel<true, kDelta><<<dim3(kBigBlocks), dim3(kApplyThreads), 0, hs>>>(b, w);
    st = check_launch("tile_apply_kernel<big>");
    if (st != CUBLOX_OK) return st;
  }
  // A grid-stride loop over the split tiles' slices: a few workgroups per
  // CU (empty workgroups of a larger grid only cost dispatch time); then
  // the near-field cubes' workgroups.
  const int hot_blocks = static_cast<int>(std::min<long long>((L.slabs / 2 + 1) * kHotSlices, 4096));
  const int near_blocks = std::max(1, std::min((c.max_near + 255) / 256, 1024));
  hot_fold_kernel<kDelta><<<dim3(static_cast<unsigned>(hot_blocks + near_blocks)), dim3(256), 0,
                            hs>>>(b, w, hot_blocks);
  return check_launch("hot_fold_kernel");
}

// Parts of a launch sequence: the kernels that read no occupancy (walk,
// record counting, tile scan, placing) and the folds.
constexpr int kPartFront = 1, kPartBack = 2, kPartBoth = 3;

// The launch sequence for a workspace of layout L. Launch shapes depend on
// the capacities only (so the sequence can be captured once and replayed);
// the kernels read the actual counts from the batch. delta selects the fold
// variants that feed changed-voxel streams (grids without one skip it).
// `parts` selects the front, the folds or both.
int launch_sequence(const Batch &b, const Layout &L, const Workspace &w,
                    bool delta, hipStream_t hs, hipEvent_t occ_ready = nullptr,
                    int batch_lanes = 0, bool shared = false,
                    int parts = kPartBoth, int cube = 0,
                    const InlineClear *clear = nullptr) {
  if (!(parts & kPartFront)) {
    return delta ? launch_folds<true>(b, L, w, hs, shared)
                 : launch_folds<false>(b, L, w, hs, shared);
  }
  const Caps &c = L.cap;
  // The walk's LDS near-field cube: the batch's largest (cube > 0), else the
  // capacity's. It bounds the walk's occupancy, so it is not sized for the
  // limits when the tick needs less.
  const int edge = cube > 0 ? std::min(cube, c.max_cube) : c.max_cube;
  const size_t near_lds = static_cast<size_t>(edge) * edge * edge * 4;
  const unsigned walk_blocks = static_cast<unsigned>((c.max_rays + kWalkBlock - 1) / kWalkBlock);
  // A single grid takes the walk its size calls for; a batch the one its
  // staged header names (the autotuner's lanes per ray; a captured sequence
  // holds one of them, so a graph is per lanes value).
  const bool quad = batch_lanes ? batch_lanes == 4 : c.max_rays <= kQuadWalkRays;
  const InlineClear zc = clear ? *clear : InlineClear{};
  const unsigned grid = walk_blocks + static_cast<unsigned>(zc.chunks);
  if (quad) {
    walk_kernel<4><<<dim3(grid), dim3(4 * kWalkBlock), near_lds, hs>>>(b, w, zc);
  } else {
    walk_kernel<1><<<dim3(grid), dim3(kWalkBlock), near_lds, hs>>>(b, w, zc);
  }
  int st = check_launch("walk_kernel");
  if (st != CUBLOX_OK) return st;
  // The quad walk counted each tile's records; after the one-lane walk a
  // binning pass counts them.
  const int bin_rays = bin_rays_per_block(c);
  const unsigned bin_blocks = static_cast<unsigned>((c.max_rays + bin_rays - 1) / bin_rays);
  if (!quad) {
    bin_kernel<<<dim3(bin_blocks), dim3(kBinThreads), 0, hs>>>(b, w, bin_rays);
    st = check_launch("bin_kernel");
    if (st != CUBLOX_OK) return st;
  }
  const bool small = L.scan_tpb == kScanSmallTiles;
  if (small) {
    // Few blocks, all resident at once: one pass, each block publishing its
    // sums and waiting for the others'.
    tile_place_kernel<<<dim3(L.scan_blocks), dim3(kScanSmallTiles), 0, hs>>>(b, w, L.scan_tpb,
                                                                             2);
    st = check_launch("tile_place_kernel");
  } else {
    tile_sums_kernel<<<dim3(L.scan_blocks), dim3(kScanThreads), 0, hs>>>(b, w, L.scan_tpb);
    st = check_launch("tile_sums_kernel");
    if (st != CUBLOX_OK) return st;
    tile_bases_kernel<<<dim3(1), dim3(kScanThreads), 0, hs>>>(b, w, L.scan_tpb);
    st = check_launch("tile_bases_kernel");
    if (st != CUBLOX_OK) return st;
    tile_place_kernel<<<dim3(L.scan_blocks), dim3(kScanThreads), 0, hs>>>(b, w, L.scan_tpb, 0);
    st = check_launch("tile_place_kernel");
  }
  if (st != CUBLOX_OK) return st;
  {
    // Counting groups: the quad walk's workgroups, else the bin blocks.
    const int lg_group = 31 - __builtin_clz(static_cast<unsigned>(quad ? kWalkBlock : bin_rays));
    // At most kPlaceRows block rows (j chunks of kPlaceJ; a row loops over
    // its further chunks): rows past a ray's records are empty workgroups,
    // whose dispatch costs time on large launches.
    const dim3 grid(walk_blocks,
                    static_cast<unsigned>(std::min((c.max_seg + kPlaceJ - 1) / kPlaceJ, kPlaceRows)));
    place_kernel<<<grid, dim3(kWalkBlock), 0, hs>>>(b, w, lg_group, quad ? kWalkHash : kHash,
                                                     small ? L.scan_blocks : 0);
    st = check_launch("place_kernel");
  }
  if (st != CUBLOX_OK) return st;
  if (!(parts & kPartBack)) return CUBLOX_OK;
  // Nothing above reads or writes occ: the folds wait for it here.
  if (occ_ready) CUBLOX_TRY(hipStreamWaitEvent(hs, occ_ready, 0));
  return delta ? launch_folds<true>(b, L, w, hs, shared)
               : launch_folds<false>(b, L, w, hs, shared);
}

Caps batch_caps(const cublox_batch_limits &lim) {
  Caps c{};
  c.max_grids = std::max(lim.max_grids, 1);
  c.max_rays = static_cast<long long>(std::max(lim.max_points, 1)) +
               static_cast<long long>(c.max_grids) * kRayAlign;
  c.seg_capacity = static_cast<long long>(std::max(lim.max_points, 1)) *
                   std::max(lim.max_segments_per_ray, 1);
  c.max_tiles = std::max(lim.max_tiles, 1);
  c.max_cube = 2 * std::max(1, std::min(lim.max_near_steps, kMaxNearK)) - 1;
  c.max_near = c.max_grids * c.max_cube * c.max_cube * c.max_cube;
  c.chunk = std::max(kMinChunk, std::min(lim.chunk > 0 ? lim.chunk : kChunk, 8192));
  c.max_slabs = lim.max_slabs > 0 ? lim.max_slabs : 65536;
  c.max_seg = std::max(lim.max_segments_per_ray, 1);
  return c;
}

} // namespace cublox_hip

namespace cublox_hip {

int read_counters(const Layout &L, const void *d_ws, size_t ws_bytes, int *out,
                  hipStream_t s) {
  if (!d_ws || ws_bytes < L.total) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate_counters: workspace too small");
  }
  CUBLOX_TRY(hipMemcpyAsync(out, static_cast<const char *>(d_ws) + L.off_misc,
                            5 * sizeof(int), hipMemcpyDeviceToHost, s));
  CUBLOX_TRY(hipStreamSynchronize(s));
  return CUBLOX_OK;
}

// ---------------------------------------------------------------------
// The update without allocation: its workspace is this file-scope device
// array. A host-side reference to `scratch` is only a stand-in for the
// symbol, so host code uses the address hipGetSymbolAddress returns for the
// current device (each device holds its own copy), looked up on a device's
// first call and kept per device.
// ---------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = CUBLOX_NOALLOC_SCRATCH_BYTES;
constexpr int kNoAllocSlabs = 64;  // hot tiles past them run unsplit
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int noalloc_scratch(void **out) {
  static std::mutex mu;
  static std::vector<void *> ptr;  // per device
  int dev = 0;
  CUBLOX_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if (static_cast<size_t>(dev) >= ptr.size()) ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
  if (!ptr[dev]) {
    void *p = nullptr;
    CUBLOX_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(CUBLOX_ERR_RUNTIME, "integrate_noalloc: no scratch address");
    ptr[dev] = p;
  }
  *out = ptr[dev];
  return CUBLOX_OK;
}

// Checks a single-grid update and builds its descriptor; Caps from the
// grid's geometry for `max_points` points.
int prepare_grid(const cublox_integrate_grid *q, int max_points, int near_steps,
                 Caps &caps, Batch &b) {
  int st = check_cfg(q->cfg, near_steps);
  if (st != CUBLOX_OK) return st;
  if (q->n > max_points) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate: n exceeds max_points");
  }
  caps = single_caps(q->cfg, max_points, near_steps);
  if (!caps_fit(caps)) {
    return set_error(CUBLOX_ERR_CAPACITY,
                     "integrate: max_points x segment bound above 2^31 slots");
  }
  if ((q->d_export_op == nullptr) != (q->d_export_hit == nullptr)) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: export needs both arrays");
  }
  st = check_delta(q->d_delta_count, q->d_delta_idx, q->d_delta_val,
                   q->delta_capacity, q->d_export_op != nullptr, "integrate");
  if (st != CUBLOX_OK) return st;
  st = check_input(*q, "integrate");
  if (st != CUBLOX_OK) return st;
  b = Batch{};
  b.one = make_desc(q->cfg, q->use_T ? q->T : nullptr, q->d_xs, q->d_ys, q->d_zs,
                    q->n, q->d_occ,
                    ApplyArgs{q->l_hit, q->l_miss, q->l_min, q->l_max, q->l_unknown},
                    near_steps, q->d_export_op, q->d_export_hit);
  set_input(b.one, *q);
  set_delta(b.one, q->d_delta_count, q->d_delta_idx, q->d_delta_val,
            q->delta_capacity);
  return CUBLOX_OK;
}

} // namespace cublox_hip

using namespace cublox_hip;

extern "C" {

size_t cublox_integrate_workspace_bytes(const cublox_raycast_cfg *cfg,
                                        int max_points, int near_steps) {
  if (!cfg || near_steps < 1 || near_steps > kMaxNearK) return 0;
  const Caps c = single_caps(*cfg, max_points, near_steps);
  return caps_fit(c) ? make_layout(c).total : 0;
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
  return cublox_launch_integrate_delta(cfg, T, d_xs, d_ys, d_zs, n, d_occ, l_hit,
                                       l_miss, l_min, l_max, l_unknown, d_ws,
                                       ws_bytes, max_points, near_steps,
                                       d_export_op, d_export_hit, nullptr,
                                       nullptr, nullptr, 0u, s);
}

int cublox_launch_integrate_delta(
    const cublox_raycast_cfg *cfg, const float *T, const float *d_xs,
    const float *d_ys, const float *d_zs, int n, float *d_occ, float l_hit,
    float l_miss, float l_min, float l_max, float l_unknown, void *d_ws,
    size_t ws_bytes, int max_points, int near_steps, int *d_export_op,
    int *d_export_hit, unsigned int *d_delta_count, int *d_delta_idx,
    float *d_delta_val, unsigned int delta_capacity, cublox_stream s) {
  if (!cfg) {
    return d_delta_count ? cublox_fill_zero_async(d_delta_count, 1, s) : CUBLOX_OK;
  }
  if (near_steps < 1 || near_steps > kMaxNearK) {
    return set_error(CUBLOX_ERR_INVALID, "integrate: near_steps out of [1, 12]");
  }
  cublox_integrate_grid q{};
  q.cfg = *cfg;
  if (T) {
    for (int k = 0; k < 12; ++k) q.T[k] = T[k];
    q.use_T = 1;
  }
  q.d_xs = d_xs;
  q.d_ys = d_ys;
  q.d_zs = d_zs;
  q.n = n;
  q.d_occ = d_occ;
  q.l_hit = l_hit;
  q.l_miss = l_miss;
  q.l_min = l_min;
  q.l_max = l_max;
  q.l_unknown = l_unknown;
  q.near_steps = near_steps;
  q.d_export_op = d_export_op;
  q.d_export_hit = d_export_hit;
  q.d_delta_count = d_delta_count;
  q.d_delta_idx = d_delta_idx;
  q.d_delta_val = d_delta_val;
  q.delta_capacity = delta_capacity;
  return cublox_launch_integrate_grid(&q, d_ws, ws_bytes, max_points, s);
}

// Target workgroups of an inline recenter clear: about one per CU, beside
// the walk's.
constexpr long long kInlineClearBlocks = 256;

int cublox_launch_integrate_grid(const cublox_integrate_grid *q, void *d_ws,
                                 size_t ws_bytes, int max_points,
                                 cublox_stream s) {
  if (!q) return set_error(CUBLOX_ERR_INVALID, "integrate: null grid");
  // A staged recenter clear goes into the walk launch when it is a slab
  // clear of few regions, else it is launched on its own first.
  InlineClear zc{};
  const StagedClear *plan = static_cast<const StagedClear *>(q->zero_plan);
  bool inline_ok = plan && q->d_occ && plan->action == 1 && plan->n > 0 &&
                   plan->n <= kInlineRegions;
  if (q->n <= 0 || !q->d_xs || !q->d_ys || !q->d_zs || !q->d_occ) {
    // Nothing to integrate: nothing changed (but a staged clear still runs).
    if (plan) {
      const int st = cublox_launch_recenter_plan(plan, q->d_occ, q->d_zero_op,
                                                 q->d_zero_hit, s);
      if (st != CUBLOX_OK) return st;
    }
    return q->d_delta_count ? cublox_fill_zero_async(q->d_delta_count, 1, s) : CUBLOX_OK;
  }
  const int near_steps = q->near_steps > 0 ? q->near_steps : 8;
  Caps caps{};
  Batch b{};
  int st = prepare_grid(q, max_points, near_steps, caps, b);
  if (st != CUBLOX_OK) return st;
  const Layout L = make_layout(caps);
  if (!d_ws || ws_bytes < L.total) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate: workspace too small");
  }
  if (plan && !inline_ok) {
    st = cublox_launch_recenter_plan(plan, q->d_occ, q->d_zero_op, q->d_zero_hit, s);
    if (st != CUBLOX_OK) return st;
  }
  if (inline_ok) {
    long long elems = 0;
    for (int k = 0; k < plan->n; ++k) elems += plan->r[k].n_outer * plan->r[k].run;
    const long long piece = chunk_piece(elems, kInlineClearBlocks);
    zc.occ = q->d_occ;
    zc.op = q->d_zero_op;
    zc.hit = q->d_zero_hit;
    zc.n = plan->n;
    for (int k = 0; k < plan->n; ++k) {
      zc.chunks += chunk_region(plan->r[k], piece, zc.chunks, &zc.r[k]);
    }
  }
  return launch_sequence(b, L, bind(L, d_ws), q->d_delta_count != nullptr,
                         as_stream(s), reinterpret_cast<hipEvent_t>(q->occ_ready), 0, false,
                         kPartBoth, 0, inline_ok ? &zc : nullptr);
}

int cublox_launch_integrate_noalloc(const cublox_raycast_cfg *cfg, const float *T,
                                    const float *d_xs, const float *d_ys,
                                    const float *d_zs, int n, float *d_occ,
                                    float l_hit, float l_miss, float l_min,
                                    float l_max, float l_unknown, int near_steps,
                                    int *d_export_op, int *d_export_hit,
                                    cublox_stream s) {
  if (!cfg || n <= 0 || !d_xs || !d_ys || !d_zs || !d_occ) return CUBLOX_OK;
  if (near_steps < 1 || near_steps > kMaxNearK) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_noalloc: near_steps out of [1, 12]");
  }
  cublox_integrate_grid q{};
  q.cfg = *cfg;
  if (T) {
    for (int k = 0; k < 12; ++k) q.T[k] = T[k];
    q.use_T = 1;
  }
  q.d_xs = d_xs;
  q.d_ys = d_ys;
  q.d_zs = d_zs;
  q.n = n;
  q.d_occ = d_occ;
  q.l_hit = l_hit;
  q.l_miss = l_miss;
  q.l_min = l_min;
  q.l_max = l_max;
  q.l_unknown = l_unknown;
  q.near_steps = near_steps;
  q.d_export_op = d_export_op;
  q.d_export_hit = d_export_hit;
  Caps caps{};
  Batch b{};
  int st = prepare_grid(&q, n, near_steps, caps, b);
  if (st != CUBLOX_OK) return st;
  caps.max_slabs = kNoAllocSlabs;
  const Layout L = make_layout(caps);
  if (L.total > kNoAllocScratchBytes) {
    return set_error(CUBLOX_ERR_CAPACITY,
                     "integrate_noalloc: " + std::to_string(n) + " points on this grid need " +
                         std::to_string(L.total) + " bytes of workspace, above the " +
                         std::to_string(kNoAllocScratchBytes) + "-byte scratch");
  }
  void *ws = nullptr;
  st = noalloc_scratch(&ws);
  if (st != CUBLOX_OK) return st;
  // The array keeps the last call's contents: the counters the kernels
  // expect at zero are cleared first, with fill kernels (no memset node
  // when the caller captures the stream). They lie at off_misc, off_nop,
  // off_nhit (consecutive) and off_tcount.
  char *base = static_cast<char *>(ws);
  st = cublox_fill_zero_async(base + L.off_tcount, static_cast<size_t>(caps.max_tiles), s);
  if (st != CUBLOX_OK) return st;
  st = cublox_fill_zero_async(base + L.off_misc,
                              (L.off_nhit + static_cast<size_t>(caps.max_near) * 4 -
                               L.off_misc) / 4,
                              s);
  if (st != CUBLOX_OK) return st;
  return launch_sequence(b, L, bind(L, ws), false, as_stream(s));
}

int cublox_integrate_segment_bound(const cublox_raycast_cfg *cfg) {
  if (!cfg) return 0;
  return grid_geom(*cfg, 1).max_seg;
}

int cublox_integrate_tile_count(const cublox_raycast_cfg *cfg) {
  if (!cfg) return 0;
  return grid_geom(*cfg, 1).n_tiles;
}

// The multi-block tile scan's base pass runs one thread per block.
constexpr int kMaxBatchTiles = kScanThreads * kScanTilesPerBlock;

size_t cublox_integrate_batch_workspace_bytes(const cublox_batch_limits *lim) {
  if (!lim || lim->max_tiles > kMaxBatchTiles) return 0;
  const Caps c = batch_caps(*lim);
  return caps_fit(c) ? make_layout(c).total : 0;
}

size_t cublox_integrate_batch_staging_bytes(const cublox_batch_limits *lim) {
  if (!lim) return 0;
  return staging_bytes(lim->max_grids);
}

int cublox_integrate_batch_stage(const cublox_integrate_grid *grids,
                                 int n_grids, const cublox_batch_limits *lim,
                                 int chunk_override, void *h_staging) {
  cublox_batch_tuning t{};
  t.chunk = chunk_override;
  return cublox_integrate_batch_stage_tuned(grids, n_grids, lim, &t, h_staging);
}

int cublox_integrate_batch_stage_tuned(const cublox_integrate_grid *grids, int n_grids,
                                       const cublox_batch_limits *lim,
                                       cublox_batch_tuning *tuning, void *h_staging) {
  const int chunk_override = tuning ? tuning->chunk : 0;
  if (!lim || !h_staging || n_grids < 0 || (n_grids > 0 && !grids)) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_batch: bad arguments");
  }
  if (lim->max_tiles > kMaxBatchTiles) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate_batch: max_tiles above 4194304");
  }
  const Caps c = batch_caps(*lim);
  if (!caps_fit(c)) {
    return set_error(CUBLOX_ERR_CAPACITY,
                     "integrate_batch: max_points x max_segments_per_ray above 2^32 slots");
  }
  if (n_grids > c.max_grids) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate_batch: too many grids");
  }
  Staging *sg = static_cast<Staging *>(h_staging);
  long long ray = 0, slot = 0, points = 0;
  int tile = 0, near = 0, max_cube = 1;
  for (int k = 0; k < n_grids; ++k) {
    const cublox_integrate_grid &q = grids[k];
    const int near_k = q.near_steps > 0 ? q.near_steps : 8;
    int st = check_cfg(q.cfg, near_k);
    if (st != CUBLOX_OK) return st;
    if (2 * near_k - 1 > c.max_cube) {
      return set_error(CUBLOX_ERR_CAPACITY, "integrate_batch: near_steps above the limit");
    }
    if ((q.d_export_op == nullptr) != (q.d_export_hit == nullptr)) {
      return set_error(CUBLOX_ERR_INVALID, "integrate_batch: export needs both arrays");
    }
    if (q.d_delta_count && !lim->delta) {
      return set_error(CUBLOX_ERR_INVALID,
                       "integrate_batch: a changed-voxel stream needs limits.delta");
    }
    st = check_delta(q.d_delta_count, q.d_delta_idx, q.d_delta_val,
                     q.delta_capacity, q.d_export_op != nullptr, "integrate_batch");
    if (st != CUBLOX_OK) return st;
    const bool live = q.n > 0 && q.d_xs && q.d_ys && q.d_zs && q.d_occ;
    GridDesc d = make_desc(q.cfg, q.use_T ? q.T : nullptr, q.d_xs, q.d_ys, q.d_zs,
                           live ? q.n : 0, q.d_occ,
                           ApplyArgs{q.l_hit, q.l_miss, q.l_min, q.l_max, q.l_unknown},
                           near_k, q.d_export_op, q.d_export_hit);
    set_delta(d, q.d_delta_count, q.d_delta_idx, q.d_delta_val, q.delta_capacity);
    st = check_input(q, "integrate_batch");
    if (st != CUBLOX_OK) return st;
    set_input(d, q);
    if (d.g.max_seg > lim->max_segments_per_ray) {
      return set_error(CUBLOX_ERR_CAPACITY,
                       "integrate_batch: a grid needs more segments per ray than the limit");
    }
    d.ray_base = static_cast<int>(ray);
    d.slot_base = slot;
    d.tile_base = tile;
    d.near_base = near;
    sg->desc[k] = d;
    const long long n = d.g.n;
    points += n;
    ray += (n + kRayAlign - 1) / kRayAlign * kRayAlign;
    slot += n * d.g.max_seg;
    tile += d.g.n_tiles;
    near += d.g.cube * d.g.cube * d.g.cube;
    max_cube = std::max(max_cube, d.g.cube);
  }
  if (points > lim->max_points || ray > c.max_rays || slot > c.seg_capacity ||
      tile > c.max_tiles || near > c.max_near) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate_batch: batch exceeds the limits");
  }
  const int chunk = chunk_override > 0 ? chunk_override : c.chunk;
  if (chunk < c.chunk || chunk > 65535) {
    return set_error(CUBLOX_ERR_INVALID,
                     "integrate_batch: chunk below the limits' chunk or above 65535");
  }
  int lanes = tuning ? tuning->lanes_per_ray : 0;
  if (lanes == 0) lanes = c.max_rays <= kQuadWalkRays ? 4 : 1;
  if (lanes != 1 && lanes != 4) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_batch: lanes per ray must be 1 or 4");
  }
  const int shared = tuning && tuning->hot_mode == 1 ? 1 : 0;
  if (tuning && (tuning->hot_mode < 0 || tuning->hot_mode > 1)) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_batch: hot-tile mode must be 0 or 1");
  }
  sg->hdr = BatchHeader{n_grids, static_cast<int>(ray), tile, near, chunk, lanes, shared,
                        max_cube};
  if (tuning) tuning->lanes_per_ray = lanes;  // the walk the launch will take
  return CUBLOX_OK;
}

int cublox_launch_integrate_batch_staged_part(const void *h_staging, void *d_ws,
                                              size_t ws_bytes,
                                              const cublox_batch_limits *lim,
                                              int part, cublox_stream s) {
  if (!lim || !h_staging || lim->max_tiles > kMaxBatchTiles || !caps_fit(batch_caps(*lim)) ||
      part < kPartFront || part > kPartBoth) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_batch: bad arguments");
  }
  const Layout L = make_layout(batch_caps(*lim));
  if (!d_ws || ws_bytes < L.total) {
    return set_error(CUBLOX_ERR_CAPACITY, "integrate_batch: workspace too small");
  }
  hipStream_t hs = as_stream(s);
  char *base = static_cast<char *>(d_ws);
  if (part & kPartFront) {
    const int st = check(hipMemcpyAsync(base + L.off_hdr, h_staging,
                                        staging_bytes(L.cap.max_grids),
                                        hipMemcpyHostToDevice, hs),
                         "hipMemcpyAsync(batch descriptors)");
    if (st != CUBLOX_OK) return st;
  }
  Batch b{};
  b.descs = reinterpret_cast<const GridDesc *>(base + L.off_desc);
  b.hdr = reinterpret_cast<const BatchHeader *>(base + L.off_hdr);
  const BatchHeader &h = static_cast<const Staging *>(h_staging)->hdr;
  return launch_sequence(b, L, bind(L, d_ws), lim->delta != 0, hs, nullptr,
                         h.lanes == 4 ? 4 : 1, h.shared != 0, part, h.cube);
}

int cublox_launch_integrate_batch_staged(const void *h_staging, void *d_ws,
                                         size_t ws_bytes,
                                         const cublox_batch_limits *lim,
                                         cublox_stream s) {
  return cublox_launch_integrate_batch_staged_part(h_staging, d_ws, ws_bytes, lim,
                                                   kPartBoth, s);
}

int cublox_launch_integrate_batch(const cublox_integrate_grid *grids,
                                  int n_grids, const cublox_batch_limits *lim,
                                  void *h_staging, void *d_ws, size_t ws_bytes,
                                  cublox_stream s) {
  int st = cublox_integrate_batch_stage(grids, n_grids, lim, 0, h_staging);
  if (st != CUBLOX_OK) return st;
  return cublox_launch_integrate_batch_staged(h_staging, d_ws, ws_bytes, lim, s);
}

int cublox_integrate_batch_counters(const void *d_ws, size_t ws_bytes,
                                    const cublox_batch_limits *lim, int *out,
                                    cublox_stream s) {
  if (!lim || !out || lim->max_tiles > kMaxBatchTiles) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_batch_counters: bad arguments");
  }
  return read_counters(make_layout(batch_caps(*lim)), d_ws, ws_bytes, out,
                       as_stream(s));
}

int cublox_integrate_counters(const void *d_ws, size_t ws_bytes,
                              const cublox_raycast_cfg *cfg, int max_points,
                              int near_steps, int *out, cublox_stream s) {
  if (!cfg || !out || near_steps < 1 || near_steps > kMaxNearK) {
    return set_error(CUBLOX_ERR_INVALID, "integrate_counters: bad arguments");
  }
  return read_counters(make_layout(single_caps(*cfg, max_points, near_steps)),
                       d_ws, ws_bytes, out, as_stream(s));
}

} // extern "C"
