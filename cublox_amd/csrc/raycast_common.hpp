// Ray-cast pieces shared by the per-voxel-atomic kernel (kernels.hip) and
// the tile-binned fused update (integrate.hip): launch arguments, the
// Amanatides-Woo set-up and wave-level helpers.
#pragma once

#include "hip_common.hpp"

namespace cublox_hip {

__device__ __forceinline__ void atomic_add_i32(int *p, int v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void atomic_add_u32(unsigned *p, unsigned v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned atomic_add_u32_ret(unsigned *p,
                                                       unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}

// Adds 1 at arr[key] for every active lane whose key is >= 0, issuing one
// atomic per distinct key among the active lanes of the wavefront (a
// match-any built from readfirstlane + ballot). Lanes with key < 0 take part
// in the matching but add nothing.
__device__ __forceinline__ void wave_add_unique(int *arr, int key) {
  const int lane = __lane_id();
  while (true) {
    const int lead = __builtin_amdgcn_readfirstlane(key);
    const bool mine = key == lead;
    const unsigned long long m = __ballot(mine);
    if (mine) {
      if (lead >= 0 && lane == __ffsll(static_cast<long long>(m)) - 1) {
        atomic_add_i32(arr + lead, __popcll(m));
      }
      break;
    }
  }
}

struct RayArgs {
  float T[12];  // sensor->world, row-major [R | t]; used when transforming
  float ox, oy, oz;
  int oix, oiy, oiz;  // window origin (global voxel)
  int mx, my, mz;
  int hx, hy, hz;
  float res, inv_res, max_range;
  int vox_num;
  int center;
  int stride;      // floats from one point's x (y, z) to the next one's
  float drop_sq;   // > 0: sensor-frame points with |p|^2 above it are dropped
};

inline RayArgs make_ray_args(const cublox_raycast_cfg &c) {
  RayArgs a{};
  a.ox = c.origin.x;
  a.oy = c.origin.y;
  a.oz = c.origin.z;
  a.oix = c.origin_i.x;
  a.oiy = c.origin_i.y;
  a.oiz = c.origin_i.z;
  a.mx = c.map_size_i.x;
  a.my = c.map_size_i.y;
  a.mz = c.map_size_i.z;
  a.hx = c.half_map_size_i.x;
  a.hy = c.half_map_size_i.y;
  a.hz = c.half_map_size_i.z;
  a.res = c.resolution;
  a.inv_res = c.inv_resolution;
  a.max_range = c.max_range;
  a.vox_num = c.map_vox_num;
  a.center = c.origin_at_center;
  a.stride = 1;
  a.drop_sq = 0.0f;
  return a;
}

// Initial Amanatides-Woo state of one ray (reference kernels.cu:45-105).
struct RayState {
  int ix, iy, iz;     // current voxel (global)
  int eix, eiy, eiz;  // endpoint voxel (global, after the range clip)
  int sx, sy, sz;     // step signs
  float tDx, tDy, tDz;
  float tMx, tMy, tMz;
  int max_steps;
  bool clipped;
};

// Loads point i (transforming it when kTransform) and sets up its walk.
// Returns false when the ray is skipped (non-finite, zero length).
template <bool kTransform, typename P>
__device__ __forceinline__ bool ray_setup(const RayArgs &a, P xs, P ys, P zs,
                                          int i, RayState &r) {
  const size_t at = static_cast<size_t>(i) * static_cast<size_t>(a.stride);
  float ex = xs[at], ey = ys[at], ez = zs[at];
  if (kTransform) {
    // Scan transform (driver main.cpp:672-682): drop non-finite sensor-frame
    // points and, when asked, points beyond the range (|p|^2 > r^2,
    // main.cpp:676-679), then p_w = R p_s + t in a fixed operand order.
    if (!isfinite(ex) || !isfinite(ey) || !isfinite(ez)) return false;
    const float px = ex, py = ey, pz = ez;
    if (a.drop_sq > 0.0f && (px * px + py * py) + pz * pz > a.drop_sq) return false;
    ex = ((a.T[0] * px + a.T[1] * py) + a.T[2] * pz) + a.T[3];
    ey = ((a.T[4] * px + a.T[5] * py) + a.T[6] * pz) + a.T[7];
    ez = ((a.T[8] * px + a.T[9] * py) + a.T[10] * pz) + a.T[11];
  }
  const float ox = a.ox, oy = a.oy, oz = a.oz;
  // kernels.cu:49-54
  float dx = ex - ox, dy = ey - oy, dz = ez - oz;
  float L = sqrtf((dx * dx + dy * dy) + dz * dz);
  if (L < 1e-6f || !isfinite(L) || !isfinite(ex) || !isfinite(ey) ||
      !isfinite(ez)) {
    return false;
  }
  // kernels.cu:57-67: range clip marks free space only.
  r.clipped = L > a.max_range;
  if (r.clipped) {
    const float s = a.max_range / L;
    ex = ox + dx * s;
    ey = oy + dy * s;
    ez = oz + dz * s;
    dx = dx * s;
    dy = dy * s;
    dz = dz * s;
    L = a.max_range;
  }
  const float invL = 1.0f / L;
  const float ux = dx * invL, uy = dy * invL, uz = dz * invL;

  // kernels.cu:73-100
  const bool ctr = a.center != 0;
  r.ix = pos_to_global(ox, a.inv_res, ctr);
  r.iy = pos_to_global(oy, a.inv_res, ctr);
  r.iz = pos_to_global(oz, a.inv_res, ctr);
  r.eix = pos_to_global(ex, a.inv_res, ctr);
  r.eiy = pos_to_global(ey, a.inv_res, ctr);
  r.eiz = pos_to_global(ez, a.inv_res, ctr);
  r.sx = (ux > 0.0f) - (ux < 0.0f);
  r.sy = (uy > 0.0f) - (uy < 0.0f);
  r.sz = (uz > 0.0f) - (uz < 0.0f);
  const float inf = __builtin_inff();
  r.tDx = r.sx == 0 ? inf : fabsf(a.res / ux);
  r.tDy = r.sy == 0 ? inf : fabsf(a.res / uy);
  r.tDz = r.sz == 0 ? inf : fabsf(a.res / uz);
  const float bx =
      ctr ? (static_cast<float>(r.ix) + 0.5f * static_cast<float>(r.sx)) * a.res
          : (static_cast<float>(r.ix) + static_cast<float>(r.sx > 0)) * a.res;
  const float by =
      ctr ? (static_cast<float>(r.iy) + 0.5f * static_cast<float>(r.sy)) * a.res
          : (static_cast<float>(r.iy) + static_cast<float>(r.sy > 0)) * a.res;
  const float bz =
      ctr ? (static_cast<float>(r.iz) + 0.5f * static_cast<float>(r.sz)) * a.res
          : (static_cast<float>(r.iz) + static_cast<float>(r.sz > 0)) * a.res;
  r.tMx = r.sx == 0 ? inf : (bx - ox) / ux;
  r.tMy = r.sy == 0 ? inf : (by - oy) / uy;
  r.tMz = r.sz == 0 ? inf : (bz - oz) / uz;
  // kernels.cu:105: Manhattan distance + 4 bounds the walk.
  r.max_steps =
      abs(r.eix - r.ix) + abs(r.eiy - r.iy) + abs(r.eiz - r.iz) + 4;
  return true;
}

// One Amanatides-Woo step with the reference's z-biased tie-break
// (kernels.cu:120-136):
//   x if tMx < tMy and tMx < tMz; else y if !(tMx < tMy) and tMy < tMz;
//   else z.
// Written as selects so a wavefront whose lanes step different axes runs
// one straight-line path; each candidate sum is formed and then selected,
// so the kept values are exactly the reference's.
__device__ __forceinline__ int ray_advance(RayState &r) {
  const bool x_lt_y = r.tMx < r.tMy;
  const bool step_x = x_lt_y && (r.tMx < r.tMz);
  const bool step_y = !x_lt_y && (r.tMy < r.tMz);
  const bool step_z = !step_x && !step_y;
  const float nx = r.tMx + r.tDx;
  const float ny = r.tMy + r.tDy;
  const float nz = r.tMz + r.tDz;
  r.ix += step_x ? r.sx : 0;
  r.iy += step_y ? r.sy : 0;
  r.iz += step_z ? r.sz : 0;
  r.tMx = step_x ? nx : r.tMx;
  r.tMy = step_y ? ny : r.tMy;
  r.tMz = step_z ? nz : r.tMz;
  return step_x ? 0 : (step_y ? 1 : 2);
}

// Hash coordinates (local index + half, in [0, map)) tracked along a walk:
// a +-1 step wraps at the map edge. Branch-free.
__device__ __forceinline__ int wrap_step(int c, int s, int m) {
  c += s;
  c = c >= m ? c - m : c;
  return c < 0 ? c + m : c;
}

struct HashCoord {
  int cx, cy, cz;
};

__device__ __forceinline__ HashCoord hash_coord_of(const RayArgs &a, int gx,
                                                   int gy, int gz) {
  return HashCoord{global_to_local(gx, a.mx, a.hx) + a.hx,
                   global_to_local(gy, a.my, a.hy) + a.hy,
                   global_to_local(gz, a.mz, a.hz) + a.hz};
}

// Applies the axis step returned by ray_advance to tracked hash coordinates.
__device__ __forceinline__ void hash_coord_step(HashCoord &c, int ax,
                                                const RayState &r,
                                                const RayArgs &a) {
  c.cx = ax == 0 ? wrap_step(c.cx, r.sx, a.mx) : c.cx;
  c.cy = ax == 1 ? wrap_step(c.cy, r.sy, a.my) : c.cy;
  c.cz = ax == 2 ? wrap_step(c.cz, r.sz, a.mz) : c.cz;
}

// Log-odds fold of one voxel with op != 0 (reference kernels.cu:184-197):
// decode, add misses and hits, clamp once, re-encode.
struct ApplyArgs {
  float l_hit, l_miss, l_min, l_max, l_unk;
};

__device__ __forceinline__ float apply_logodds(float stored0, int op, int hit,
                                               const ApplyArgs &p) {
  const int miss = op > hit ? op - hit : 0;
  float l = stored0 + p.l_unk;
  l = l + p.l_miss * static_cast<float>(miss);
  l = l + p.l_hit * static_cast<float>(hit);
  if (l < p.l_min) l = p.l_min;
  if (l > p.l_max) l = p.l_max;
  return l - p.l_unk;
}

} // namespace cublox_hip
