/*
 * cublox oracle — CPU restatement of the reference occupancy-update path.
 * TEST INFRASTRUCTURE ONLY (see cublox_oracle.h for the contract and pins).
 *
 * Build: g++ -O2 -ffp-contract=off -fopenmp (no -ffast-math). All FP32
 * expressions are written in the reference's left-to-right order so that,
 * with contraction off, each rounding step is the same IEEE operation.
 */
#include "cublox_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct I3 {
  int x, y, z;
};
struct F3 {
  float x, y, z;
};

/* utils.cuh:24-32 — corner binning floors; centre binning rounds half away. */
inline int pos_to_global(float p, float inv_res, bool center) {
  if (center) {
    const float s = p < 0.0f ? -1.0f : 1.0f;
    const float q = p * inv_res;
    const float r = q + s * 0.5f;
    return static_cast<int>(r);
  }
  const float q = p * inv_res;
  return static_cast<int>(std::floor(q));
}

/* utils.cuh:44-48 */
inline float global_to_pos(int g, float res, bool center) {
  if (center) {
    return static_cast<float>(g) * res;
  }
  const float c = static_cast<float>(g) + 0.5f;
  return c * res;
}

/* utils.cuh:60-70 — C remainder, then fold into [-half, half]. */
inline int global_to_local(int g, int map, int half) {
  int l = g % map;
  if (l > half) {
    l -= map;
  } else if (l < -half) {
    l += map;
  }
  return l;
}

/* utils.cuh:82-92 */
inline int local_to_global(int l, int map, int half, int origin_g) {
  const int min_g = -half + origin_g;
  int min_l = min_g % map;
  if (min_l > half) min_l -= map;
  if (min_l < -half) min_l += map;
  int diff = l - min_l;
  if (diff < 0) diff += map;
  return min_g + diff;
}

/* utils.cuh:131-136 — x-major, z fastest. */
inline int local_to_hash(I3 l, I3 map, I3 half) {
  const int ix = l.x + half.x, iy = l.y + half.y, iz = l.z + half.z;
  return ix * map.y * map.z + iy * map.z + iz;
}

/* utils.cuh:120-129 */
inline I3 hash_to_local(int h, I3 map, I3 half) {
  const int yz = map.y * map.z;
  I3 l;
  l.x = h / yz;
  l.y = (h - l.x * yz) / map.z;
  l.z = h - l.x * yz - l.y * map.z;
  l.x -= half.x;
  l.y -= half.y;
  l.z -= half.z;
  return l;
}

/* utils.cuh:160-166 */
inline int global_to_hash(I3 g, I3 map, I3 half) {
  I3 l{global_to_local(g.x, map.x, half.x), global_to_local(g.y, map.y, half.y),
       global_to_local(g.z, map.z, half.z)};
  return local_to_hash(l, map, half);
}

/* kernels.cu:23-28 */
inline bool inside_window(I3 g, I3 o, I3 half) {
  return std::abs(g.x - o.x) <= half.x && std::abs(g.y - o.y) <= half.y &&
         std::abs(g.z - o.z) <= half.z;
}

struct RayCfg {
  F3 origin;
  I3 origin_i, map, half;
  float res, inv_res, max_range;
  int vox_num;
  bool center;
};

/* One ray of rayCastUpdateKernel (kernels.cu:39-151). `visit(h, endpoint,
 * clipped)` is called for each counter increment in reference order.
 * Returns false when the ray is skipped (kernels.cu:51-54). */
template <typename Visit>
inline bool trace_ray(const RayCfg &c, float ex, float ey, float ez,
                      Visit &&visit) {
  const float ox = c.origin.x, oy = c.origin.y, oz = c.origin.z;
  float dx = ex - ox, dy = ey - oy, dz = ez - oz;
  /* kernels.cu:50 — ((dx*dx + dy*dy) + dz*dz), then IEEE sqrt. */
  const float sxx = dx * dx;
  const float syy = dy * dy;
  const float szz = dz * dz;
  float L = std::sqrt((sxx + syy) + szz);
  if (L < 1e-6f || !std::isfinite(L) || !std::isfinite(ex) ||
      !std::isfinite(ey) || !std::isfinite(ez)) {
    return false;
  }
  /* kernels.cu:57-67 — clip to max_range; the clip point is not a hit. */
  const bool clipped = L > c.max_range;
  if (clipped) {
    const float s = c.max_range / L;
    ex = ox + dx * s;
    ey = oy + dy * s;
    ez = oz + dz * s;
    dx = dx * s;
    dy = dy * s;
    dz = dz * s;
    L = c.max_range;
  }
  /* kernels.cu:69-70 */
  const float invL = 1.0f / L;
  const float ux = dx * invL, uy = dy * invL, uz = dz * invL;

  /* kernels.cu:73-100 */
  int ix = pos_to_global(ox, c.inv_res, c.center);
  int iy = pos_to_global(oy, c.inv_res, c.center);
  int iz = pos_to_global(oz, c.inv_res, c.center);
  const int eix = pos_to_global(ex, c.inv_res, c.center);
  const int eiy = pos_to_global(ey, c.inv_res, c.center);
  const int eiz = pos_to_global(ez, c.inv_res, c.center);
  const int sx = (ux > 0.0f) - (ux < 0.0f);
  const int sy = (uy > 0.0f) - (uy < 0.0f);
  const int sz = (uz > 0.0f) - (uz < 0.0f);
  const float inf = std::numeric_limits<float>::infinity();
  const float tDx = sx == 0 ? inf : std::fabs(c.res / ux);
  const float tDy = sy == 0 ? inf : std::fabs(c.res / uy);
  const float tDz = sz == 0 ? inf : std::fabs(c.res / uz);
  float bx, by, bz;
  if (c.center) {
    bx = (static_cast<float>(ix) + 0.5f * static_cast<float>(sx)) * c.res;
    by = (static_cast<float>(iy) + 0.5f * static_cast<float>(sy)) * c.res;
    bz = (static_cast<float>(iz) + 0.5f * static_cast<float>(sz)) * c.res;
  } else {
    bx = (static_cast<float>(ix) + static_cast<float>(sx > 0)) * c.res;
    by = (static_cast<float>(iy) + static_cast<float>(sy > 0)) * c.res;
    bz = (static_cast<float>(iz) + static_cast<float>(sz > 0)) * c.res;
  }
  float tMx = sx == 0 ? inf : (bx - ox) / ux;
  float tMy = sy == 0 ? inf : (by - oy) / uy;
  float tMz = sz == 0 ? inf : (bz - oz) / uz;

  /* kernels.cu:105-137 — Manhattan bound + 4, z-biased tie-break. */
  const int max_steps =
      std::abs(eix - ix) + std::abs(eiy - iy) + std::abs(eiz - iz) + 4;
  for (int step = 0; step < max_steps; ++step) {
    if (ix == eix && iy == eiy && iz == eiz) break;
    const I3 g{ix, iy, iz};
    if (inside_window(g, c.origin_i, c.half)) {
      const int h = global_to_hash(g, c.map, c.half);
      if (h >= 0 && h < c.vox_num) visit(h, false, clipped);
    }
    if (tMx < tMy) {
      if (tMx < tMz) {
        ix += sx;
        tMx += tDx;
      } else {
        iz += sz;
        tMz += tDz;
      }
    } else {
      if (tMy < tMz) {
        iy += sy;
        tMy += tDy;
      } else {
        iz += sz;
        tMz += tDz;
      }
    }
  }
  /* kernels.cu:141-151 — endpoint counted even if the walk never reached
   * it; hit only for real (unclipped) returns. */
  const I3 e{eix, eiy, eiz};
  if (inside_window(e, c.origin_i, c.half)) {
    const int h = global_to_hash(e, c.map, c.half);
    if (h >= 0 && h < c.vox_num) visit(h, true, clipped);
  }
  return true;
}

/* kernels.cu:303-311 */
inline int normalize_slice(int x, int a, int b) {
  const int range = b - a + 1;
  int y = (x - a) % range;
  if (y < 0) y += range;
  return y + a;
}

struct Plan {
  int action = 0;
  I3 new_origin{0, 0, 0};
  std::vector<int> slices[3];
};

/* planRecenterKernel (kernels.cu:313-398), including the origin-independent
 * "entering" slices (kernels.cu:381-384, 392-395). */
inline Plan plan_recenter(F3 pos, F3 origin_f, I3 origin_i, I3 map, I3 half,
                          float inv_res, float thr_sq, bool center) {
  Plan p;
  const float dx = pos.x - origin_f.x;
  const float dy = pos.y - origin_f.y;
  const float dz = pos.z - origin_f.z;
  const float d2 = (dx * dx + dy * dy) + dz * dz;
  if (d2 < thr_sq) return p;
  const I3 no{pos_to_global(pos.x, inv_res, center),
              pos_to_global(pos.y, inv_res, center),
              pos_to_global(pos.z, inv_res, center)};
  const I3 shift{no.x - origin_i.x, no.y - origin_i.y, no.z - origin_i.z};
  const int max_shift =
      std::max(std::abs(shift.x), std::max(std::abs(shift.y), std::abs(shift.z)));
  const int max_map = std::max(map.x, std::max(map.y, map.z));
  p.new_origin = no;
  if (max_shift >= max_map) {
    p.action = 2;
    return p;
  }
  p.action = 1;
  const int sh[3] = {shift.x, shift.y, shift.z};
  const int hs[3] = {half.x, half.y, half.z};
  const int ms[3] = {map.x, map.y, map.z};
  const int og[3] = {origin_i.x, origin_i.y, origin_i.z};
  for (int a = 0; a < 3; ++a) {
    const int s = sh[a];
    if (s == 0) continue;
    const int h = hs[a];
    const int min_l = (-h + og[a]) % ms[a];
    std::vector<int> &out = p.slices[a];
    if (s > 0) {
      for (int k = 0; k < s; ++k) out.push_back(normalize_slice(min_l + k, -h, h));
      for (int k = 0; k < s; ++k) out.push_back(normalize_slice(h - k, -h, h));
    } else {
      for (int k = -1; k >= s; --k) out.push_back(normalize_slice(min_l + k, -h, h));
      for (int k = 0; k < -s; ++k) out.push_back(normalize_slice(-h + k, -h, h));
    }
  }
  return p;
}

int default_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

} // namespace

/* =================== C ABI: free functions =================== */

extern "C" int orc_pos_to_global(float p, float inv_res, int c) {
  return pos_to_global(p, inv_res, c != 0);
}
extern "C" float orc_global_to_pos(int g, float res, int c) {
  return global_to_pos(g, res, c != 0);
}
extern "C" int orc_global_to_local(int g, int map, int half) {
  return global_to_local(g, map, half);
}
extern "C" int orc_local_to_global(int l, int map, int half, int origin_g) {
  return local_to_global(l, map, half, origin_g);
}
extern "C" int orc_local_to_hash(const int l[3], const int map[3],
                                 const int half[3]) {
  return local_to_hash(I3{l[0], l[1], l[2]}, I3{map[0], map[1], map[2]},
                       I3{half[0], half[1], half[2]});
}
extern "C" void orc_hash_to_local(int h, const int map[3], const int half[3],
                                  int l[3]) {
  const I3 r = hash_to_local(h, I3{map[0], map[1], map[2]},
                             I3{half[0], half[1], half[2]});
  l[0] = r.x;
  l[1] = r.y;
  l[2] = r.z;
}
extern "C" int orc_global_to_hash(const int g[3], const int map[3],
                                  const int half[3]) {
  return global_to_hash(I3{g[0], g[1], g[2]}, I3{map[0], map[1], map[2]},
                        I3{half[0], half[1], half[2]});
}
/* utils.hpp:185-188 */
extern "C" float orc_logit(float p) {
  const float pc = std::clamp(p, 1.0e-6f, 1.0f - 1.0e-6f);
  return std::log(pc / (1.0f - pc));
}

extern "C" int orc_trace_ray(const float origin[3], const float endpoint[3],
                             const int origin_i[3], const int map[3],
                             const int half[3], float res, float inv_res,
                             float max_range, int vox_num, int center,
                             int *out_h, int cap, int *clipped,
                             int *endpoint_counted) {
  RayCfg c;
  c.origin = F3{origin[0], origin[1], origin[2]};
  c.origin_i = I3{origin_i[0], origin_i[1], origin_i[2]};
  c.map = I3{map[0], map[1], map[2]};
  c.half = I3{half[0], half[1], half[2]};
  c.res = res;
  c.inv_res = inv_res;
  c.max_range = max_range;
  c.vox_num = vox_num;
  c.center = center != 0;
  int n = 0;
  int clip = 0, ep = 0;
  const bool ok = trace_ray(c, endpoint[0], endpoint[1], endpoint[2],
                            [&](int h, bool is_end, bool cl) {
                              if (n < cap) out_h[n] = h;
                              ++n;
                              clip = cl ? 1 : 0;
                              if (is_end) ep = 1;
                            });
  if (clipped) *clipped = clip;
  if (endpoint_counted) *endpoint_counted = ep;
  return ok ? n : -1;
}

extern "C" int orc_plan_recenter(const float pos[3], const float origin_f[3],
                                 const int origin_i[3], const int map[3],
                                 const int half[3], float inv_res,
                                 float thr_sq, int center, int new_origin[3],
                                 int *sx, int *sy, int *sz, int n_slices[3]) {
  const Plan p = plan_recenter(
      F3{pos[0], pos[1], pos[2]}, F3{origin_f[0], origin_f[1], origin_f[2]},
      I3{origin_i[0], origin_i[1], origin_i[2]}, I3{map[0], map[1], map[2]},
      I3{half[0], half[1], half[2]}, inv_res, thr_sq, center != 0);
  new_origin[0] = p.new_origin.x;
  new_origin[1] = p.new_origin.y;
  new_origin[2] = p.new_origin.z;
  int *outs[3] = {sx, sy, sz};
  for (int a = 0; a < 3; ++a) {
    n_slices[a] = static_cast<int>(p.slices[a].size());
    if (outs[a]) {
      std::copy(p.slices[a].begin(), p.slices[a].end(), outs[a]);
    }
  }
  return p.action;
}

/* main.cpp:54, 640-693. Eigen's squaredNorm of a Vector3f and its
 * Isometry3f * Vector3f are taken as (x*x + y*y) + z*z and the row sums in
 * column order (Eigen is not available here to pin either). */
extern "C" int orc_build_cloud(const uint8_t *data, long long n_records,
                               int point_step, int off_x, int off_y, int off_z,
                               const float T[12], float max_range_sq,
                               float *xw, float *yw, float *zw) {
  if (n_records <= 0) return 0;
  const long long n = std::min<long long>(n_records, 1000000);
  int row = 0;
  for (long long i = 0; i < n; ++i) {
    const uint8_t *rec = data + i * point_step;
    float lx, ly, lz;
    std::memcpy(&lx, rec + off_x, 4);
    std::memcpy(&ly, rec + off_y, 4);
    std::memcpy(&lz, rec + off_z, 4);
    if (!std::isfinite(lx) || !std::isfinite(ly) || !std::isfinite(lz)) continue;
    if ((lx * lx + ly * ly) + lz * lz > max_range_sq) continue;
    xw[row] = ((T[0] * lx + T[1] * ly) + T[2] * lz) + T[3];
    yw[row] = ((T[4] * lx + T[5] * ly) + T[6] * lz) + T[7];
    zw[row] = ((T[8] * lx + T[9] * ly) + T[10] * lz) + T[11];
    ++row;
  }
  return row;
}

/* main.cpp:558-638. */
extern "C" void orc_depth_to_cloud(const uint8_t *img, int width, int height,
                                   long long step, int encoding,
                                   const float K[4], float *xs, float *ys,
                                   float *zs) {
  const float fx = K[0], fy = K[1], cx = K[2], cy = K[3];
  const float nan = std::numeric_limits<float>::quiet_NaN();
  long long k = 0;
  for (int v = 0; v < height; ++v) {
    const uint8_t *row = img + static_cast<long long>(v) * step;
    for (int u = 0; u < width; ++u, ++k) {
      float z;
      if (encoding == 0) {
        uint16_t raw;
        std::memcpy(&raw, row + 2 * u, 2);
        if (raw == 0) {
          xs[k] = ys[k] = zs[k] = nan;
          continue;
        }
        z = static_cast<float>(raw) * 0.001f;
      } else {
        std::memcpy(&z, row + 4 * u, 4);
        if (!std::isfinite(z) || z <= 0.f) {
          xs[k] = ys[k] = zs[k] = nan;
          continue;
        }
      }
      xs[k] = (static_cast<float>(u) - cx) * z / fx;
      ys[k] = (static_cast<float>(v) - cy) * z / fy;
      zs[k] = z;
    }
  }
}

/* main.cpp:275-291 then 293-304. */
static float viz_color_packed(float d, float d_max) {
  const float d_max_clamped = std::max(d_max, 1.0f);
  const float t = std::clamp(d / d_max_clamped, 0.0f, 1.0f);
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
  const uint32_t ur = static_cast<uint32_t>(std::clamp(r, 0.f, 1.f) * 255.f);
  const uint32_t ug = static_cast<uint32_t>(std::clamp(g, 0.f, 1.f) * 255.f);
  const uint32_t ub = static_cast<uint32_t>(std::clamp(b, 0.f, 1.f) * 255.f);
  const uint32_t packed = (ur << 16) | (ug << 8) | ub;
  float rgb = 0.f;
  std::memcpy(&rgb, &packed, sizeof(float));
  return rgb;
}

/* main.cpp:71-84 and 763-790. */
extern "C" int orc_viz_points(const float *pos, int n, const float robot[3],
                              float d_max, float inv_res, int origin_at_center,
                              int step, int max_pts, float *out4) {
  int written = 0;
  for (int i = 0; i < n && written < max_pts; ++i) {
    const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    if (step > 1) {
      const unsigned mix =
          static_cast<unsigned>(orc_pos_to_global(x, inv_res, origin_at_center)) ^
          static_cast<unsigned>(orc_pos_to_global(y, inv_res, origin_at_center)) ^
          static_cast<unsigned>(orc_pos_to_global(z, inv_res, origin_at_center));
      if (mix % static_cast<unsigned>(step) != 0u) continue;
    }
    const float dx = x - robot[0], dy = y - robot[1], dz = z - robot[2];
    const float d = std::sqrt((dx * dx + dy * dy) + dz * dz);
    out4[4 * written + 0] = x;
    out4[4 * written + 1] = y;
    out4[4 * written + 2] = z;
    out4[4 * written + 3] = viz_color_packed(d, d_max);
    ++written;
  }
  return written;
}

extern "C" void orc_transform_points(const float T[12], const float *xs,
                                     const float *ys, const float *zs, int n,
                                     float *xw, float *yw, float *zw) {
  for (int i = 0; i < n; ++i) {
    const float px = xs[i], py = ys[i], pz = zs[i];
    xw[i] = ((T[0] * px + T[1] * py) + T[2] * pz) + T[3];
    yw[i] = ((T[4] * px + T[5] * py) + T[6] * pz) + T[7];
    zw[i] = ((T[8] * px + T[9] * py) + T[10] * pz) + T[11];
  }
}

/* =================== stateful grid =================== */

struct orc_grid {
  /* Grid::Config (Grid.hpp:44-53, Grid.cpp:27-54) */
  float res = 0.f, inv_res = 0.f;
  I3 map{0, 0, 0}, half{0, 0, 0};
  int vox_num = 0;
  bool center = false;
  bool has_thr = false;
  float thr_sq = 0.f;
  /* Grid public state (Grid.hpp:55-56) */
  F3 origin_f{0, 0, 0};
  I3 origin_i{0, 0, 0};
  /* OccupancyGrid private state (OccupancyGrid.hpp:158-168 defaults) */
  bool first_run = true;
  float max_range = 25.0f;
  float l_hit = 0.847f, l_miss = -0.405f, l_min = -1.992f, l_max = 3.476f;
  float l_free = -0.0004f, l_occ = 1.7346f, l_unk = 0.0f;
  std::vector<float> occ;
  std::vector<int> op, hit;
  int threads = 1;
};

namespace {

/* Grid::updateOriginAndBound (Grid.cpp:80-92) — bounds are derived data. */
void set_origin(orc_grid *g, I3 o) {
  g->origin_i = o;
  g->origin_f = F3{global_to_pos(o.x, g->res, g->center),
                   global_to_pos(o.y, g->res, g->center),
                   global_to_pos(o.z, g->res, g->center)};
}

RayCfg ray_cfg(const orc_grid *g, F3 sensor) {
  RayCfg c;
  c.origin = sensor;
  c.origin_i = g->origin_i;
  c.map = g->map;
  c.half = g->half;
  c.res = g->res;
  c.inv_res = g->inv_res;
  c.max_range = g->max_range;
  c.vox_num = g->vox_num;
  c.center = g->center;
  return c;
}

/* rayCastUpdateKernel over all points, OpenMP over rays with atomic int
 * increments (order-independent, hence deterministic). */
void raycast(orc_grid *g, const float *x, const float *y, const float *z,
             int n, F3 sensor) {
  const RayCfg c = ray_cfg(g, sensor);
  int *op = g->op.data();
  int *hit = g->hit.data();
#pragma omp parallel for schedule(dynamic, 256) num_threads(g->threads)
  for (int i = 0; i < n; ++i) {
    trace_ray(c, x[i], y[i], z[i], [&](int h, bool is_end, bool clipped) {
#pragma omp atomic
      op[h] += 1;
      if (is_end && !clipped) {
#pragma omp atomic
        hit[h] += 1;
      }
    });
  }
}

/* applyUpdateKernel (kernels.cu:167-200). */
void apply(orc_grid *g) {
  float *occ = g->occ.data();
  int *op = g->op.data();
  int *hit = g->hit.data();
  const int V = g->vox_num;
  const float l_hit = g->l_hit, l_miss = g->l_miss, l_min = g->l_min,
              l_max = g->l_max, l_unk = g->l_unk;
#pragma omp parallel for schedule(static) num_threads(g->threads)
  for (int h = 0; h < V; ++h) {
    const int o = op[h];
    if (o == 0) continue;
    const int ht = hit[h];
    const int miss = o > ht ? o - ht : 0;
    float l = occ[h] + l_unk;
    l = l + l_miss * static_cast<float>(miss);
    l = l + l_hit * static_cast<float>(ht);
    if (l < l_min) l = l_min;
    if (l > l_max) l = l_max;
    occ[h] = l - l_unk;
    op[h] = 0;
    hit[h] = 0;
  }
}

/* clearRecenterSlabAxis{0,1,2}Kernel (kernels.cu:404-499): zero every voxel
 * whose local index on `axis` is in `slices` (occ and both counters). */
void clear_slabs(orc_grid *g, int axis, const std::vector<int> &slices) {
  const I3 m = g->map, hf = g->half;
  const int ns = static_cast<int>(slices.size());
  const int outer = axis == 0 ? m.y : m.x;
#pragma omp parallel for schedule(static) num_threads(g->threads)
  for (int t = 0; t < ns * outer; ++t) {
    const int si = t / outer;
    const int r = t % outer;
    I3 l0;
    int step, count;
    if (axis == 0) {
      l0 = I3{slices[si], r - hf.y, -hf.z};
      step = 1;
      count = m.z;
    } else if (axis == 1) {
      l0 = I3{r - hf.x, slices[si], -hf.z};
      step = 1;
      count = m.z;
    } else {
      l0 = I3{r - hf.x, -hf.y, slices[si]};
      step = m.z;
      count = m.y;
    }
    int h = local_to_hash(l0, m, hf);
    for (int k = 0; k < count; ++k, h += step) {
      g->occ[h] = 0.0f;
      g->op[h] = 0;
      g->hit[h] = 0;
    }
  }
}

} // namespace

extern "C" orc_grid *orc_grid_create(int hx, int hy, int hz, float res,
                                     int center, int has_thr, double thr,
                                     float ox, float oy, float oz) {
  orc_grid *g = new orc_grid();
  g->res = res;
  g->inv_res = 1.0f / res;
  g->half = I3{hx, hy, hz};
  g->map = I3{2 * hx + 1, 2 * hy + 1, 2 * hz + 1};
  g->has_thr = has_thr != 0;
  if (g->has_thr) {
    /* optional<double> -> optional<float> (OccupancyGrid.cpp:59), then
     * squared in float (Grid.cpp:33-38). */
    const float t = static_cast<float>(thr);
    g->thr_sq = t * t;
  }
  g->center = center != 0;
  const int64_t total =
      static_cast<int64_t>(g->map.x) * g->map.y * static_cast<int64_t>(g->map.z);
  if (total > INT_MAX) {
    delete g;
    return nullptr;
  }
  g->vox_num = static_cast<int>(total);
  set_origin(g, I3{pos_to_global(ox, g->inv_res, g->center),
                   pos_to_global(oy, g->inv_res, g->center),
                   pos_to_global(oz, g->inv_res, g->center)});
  g->occ.assign(static_cast<size_t>(g->vox_num), 0.0f);
  g->op.assign(static_cast<size_t>(g->vox_num), 0);
  g->hit.assign(static_cast<size_t>(g->vox_num), 0);
  g->threads = default_threads();
  return g;
}

extern "C" void orc_grid_destroy(orc_grid *g) { delete g; }
extern "C" void orc_grid_set_threads(orc_grid *g, int n) {
  g->threads = n > 0 ? n : default_threads();
}
extern "C" void orc_grid_set_max_range(orc_grid *g, float r) { g->max_range = r; }
extern "C" void orc_grid_set_log_odds(orc_grid *g, float l_hit, float l_miss,
                                      float l_min, float l_max, float l_free,
                                      float l_occupied) {
  g->l_hit = l_hit;
  g->l_miss = l_miss;
  g->l_min = l_min;
  g->l_max = l_max;
  g->l_free = l_free;
  g->l_occ = l_occupied;
  g->l_unk = 0.0f;
}

extern "C" void orc_grid_raycast_world(orc_grid *g, const float *x,
                                       const float *y, const float *z, int n,
                                       const float s[3]) {
  if (g->vox_num <= 0 || n <= 0) return;
  const F3 sensor{s[0], s[1], s[2]};
  /* OccupancyGrid.cpp:192-198 — first update pins the window origin to the
   * sensor voxel. */
  if (g->first_run) {
    set_origin(g, I3{pos_to_global(sensor.x, g->inv_res, g->center),
                     pos_to_global(sensor.y, g->inv_res, g->center),
                     pos_to_global(sensor.z, g->inv_res, g->center)});
    g->first_run = false;
  }
  raycast(g, x, y, z, n, sensor);
}

extern "C" void orc_grid_apply(orc_grid *g) { apply(g); }

extern "C" void orc_grid_update(orc_grid *g, const float *x, const float *y,
                                const float *z, int n, const float s[3]) {
  if (g->vox_num <= 0 || n <= 0) return;
  orc_grid_raycast_world(g, x, y, z, n, s);
  apply(g);
}

extern "C" void orc_grid_update_sensor(orc_grid *g, const float *xs,
                                       const float *ys, const float *zs, int n,
                                       const float T[12]) {
  std::vector<float> xw, yw, zw;
  xw.reserve(n);
  yw.reserve(n);
  zw.reserve(n);
  for (int i = 0; i < n; ++i) {
    const float px = xs[i], py = ys[i], pz = zs[i];
    if (!std::isfinite(px) || !std::isfinite(py) || !std::isfinite(pz)) continue;
    xw.push_back(((T[0] * px + T[1] * py) + T[2] * pz) + T[3]);
    yw.push_back(((T[4] * px + T[5] * py) + T[6] * pz) + T[7]);
    zw.push_back(((T[8] * px + T[9] * py) + T[10] * pz) + T[11]);
  }
  const float s[3] = {T[3], T[7], T[11]};
  orc_grid_update(g, xw.data(), yw.data(), zw.data(),
                  static_cast<int>(xw.size()), s);
}

extern "C" int orc_grid_recenter(orc_grid *g, const float pos[3]) {
  if (!g->has_thr || g->vox_num <= 0) return 0;
  const F3 p{pos[0], pos[1], pos[2]};
  /* OccupancyGrid.cpp:147-150 host pre-check, (a0 + a1) + a2 order. */
  const float dx = p.x - g->origin_f.x;
  const float dy = p.y - g->origin_f.y;
  const float dz = p.z - g->origin_f.z;
  if ((dx * dx + dy * dy) + dz * dz < g->thr_sq) return 0;
  const Plan plan = plan_recenter(p, g->origin_f, g->origin_i, g->map, g->half,
                                  g->inv_res, g->thr_sq, g->center);
  if (plan.action == 0) return 0;
  if (plan.action == 2) {
    std::fill(g->occ.begin(), g->occ.end(), 0.0f);
    std::fill(g->op.begin(), g->op.end(), 0);
    std::fill(g->hit.begin(), g->hit.end(), 0);
  } else {
    for (int a = 0; a < 3; ++a) {
      if (!plan.slices[a].empty()) clear_slabs(g, a, plan.slices[a]);
    }
  }
  set_origin(g, plan.new_origin);
  return plan.action;
}

extern "C" int orc_grid_fetch(const orc_grid *g, const float c[3],
                              float radius, int filter, int max_results,
                              float *out_pos, uint8_t *out_state, int cap) {
  if (g->vox_num <= 0 || !(radius > 0.0f)) return 0;
  const F3 center{c[0], c[1], c[2]};
  /* Grid::maxHorizontalInWindowRadius (Grid.cpp:67-78) */
  const I3 cg{pos_to_global(center.x, g->inv_res, g->center),
              pos_to_global(center.y, g->inv_res, g->center),
              pos_to_global(center.z, g->inv_res, g->center)};
  const int mx = g->half.x - std::abs(cg.x - g->origin_i.x);
  const int my = g->half.y - std::abs(cg.y - g->origin_i.y);
  float margin = 0.0f;
  if (mx >= 0 && my >= 0) {
    margin = static_cast<float>(std::min(mx, my)) * g->res;
  }
  /* OccupancyGrid.cpp:248-290 */
  const float r = std::min(radius, margin);
  if (!(r > 0.0f)) return 0;
  const F3 wmin{center.x - r, center.y - r, center.z - r};
  const F3 wmax{center.x + r, center.y + r, center.z + r};
  I3 gmin{pos_to_global(wmin.x, g->inv_res, g->center),
          pos_to_global(wmin.y, g->inv_res, g->center),
          pos_to_global(wmin.z, g->inv_res, g->center)};
  I3 gmax{pos_to_global(wmax.x, g->inv_res, g->center),
          pos_to_global(wmax.y, g->inv_res, g->center),
          pos_to_global(wmax.z, g->inv_res, g->center)};
  const I3 lo{g->origin_i.x - g->half.x, g->origin_i.y - g->half.y,
              g->origin_i.z - g->half.z};
  const I3 hi{g->origin_i.x + g->half.x, g->origin_i.y + g->half.y,
              g->origin_i.z + g->half.z};
  gmin = I3{std::max(gmin.x, lo.x), std::max(gmin.y, lo.y),
            std::max(gmin.z, lo.z)};
  gmax = I3{std::min(gmax.x, hi.x), std::min(gmax.y, hi.y),
            std::min(gmax.z, hi.z)};
  if (gmax.x < gmin.x || gmax.y < gmin.y || gmax.z < gmin.z) return 0;
  const int disk_r =
      static_cast<int>(std::ceil(static_cast<double>(r * g->inv_res)));
  const int nz = gmax.z - gmin.z + 1;
  const int diam = 2 * disk_r + 1;
  const int64_t est = static_cast<int64_t>(diam) * diam * nz;
  if (est <= 0 || est > static_cast<int64_t>(INT_MAX)) return 0;
  const unsigned out_cap =
      max_results > 0 ? static_cast<unsigned>(max_results) : 500000u;
  const float r2 = r * r;
  /* fetchOccupancySphereKernel (kernels.cu:506-562) per voxel. */
  unsigned count = 0;
  for (int tx = 0; tx < diam; ++tx) {
    for (int ty = 0; ty < diam; ++ty) {
      for (int gz = gmin.z; gz <= gmax.z; ++gz) {
        const I3 id{cg.x + tx - disk_r, cg.y + ty - disk_r, gz};
        if (!inside_window(id, g->origin_i, g->half)) continue;
        const F3 p{global_to_pos(id.x, g->res, g->center),
                   global_to_pos(id.y, g->res, g->center),
                   global_to_pos(id.z, g->res, g->center)};
        const float dx = p.x - center.x;
        const float dy = p.y - center.y;
        const float dz = p.z - center.z;
        if ((dx * dx + dy * dy) + dz * dz > r2) continue;
        const int h = global_to_hash(id, g->map, g->half);
        const float l = g->occ[h] + g->l_unk;
        int st = 0;
        if (l >= g->l_occ) {
          st = 2;
        } else if (l < g->l_free) {
          st = 1;
        }
        if (filter >= 0 && st != filter) continue;
        const unsigned slot = count++;
        if (slot >= out_cap) continue;
        if (static_cast<int>(slot) < cap) {
          out_pos[3 * slot + 0] = p.x;
          out_pos[3 * slot + 1] = p.y;
          out_pos[3 * slot + 2] = p.z;
          out_state[slot] = static_cast<uint8_t>(st);
        }
      }
    }
  }
  /* OccupancyGrid.cpp:329-335 */
  unsigned n = std::min(count, out_cap);
  if (max_results > 0) n = std::min(n, static_cast<unsigned>(max_results));
  return static_cast<int>(n);
}

extern "C" void orc_grid_reset(orc_grid *g) {
  std::fill(g->occ.begin(), g->occ.end(), 0.0f);
  std::fill(g->op.begin(), g->op.end(), 0);
  std::fill(g->hit.begin(), g->hit.end(), 0);
  g->first_run = true;
}
extern "C" void orc_grid_reset_voxel(orc_grid *g, int h) {
  if (h < 0 || h >= g->vox_num) return;
  g->occ[h] = 0.0f;
  g->op[h] = 0;
  g->hit[h] = 0;
}
extern "C" float *orc_grid_occ(orc_grid *g) { return g->occ.data(); }
extern "C" int *orc_grid_op(orc_grid *g) { return g->op.data(); }
extern "C" int *orc_grid_hit(orc_grid *g) { return g->hit.data(); }
extern "C" int orc_grid_voxel_num(const orc_grid *g) { return g->vox_num; }
extern "C" void orc_grid_origin(const orc_grid *g, int oi[3], float of[3]) {
  oi[0] = g->origin_i.x;
  oi[1] = g->origin_i.y;
  oi[2] = g->origin_i.z;
  of[0] = g->origin_f.x;
  of[1] = g->origin_f.y;
  of[2] = g->origin_f.z;
}
extern "C" void orc_grid_load_occ(orc_grid *g, const float *occ) {
  std::memcpy(g->occ.data(), occ, sizeof(float) * g->occ.size());
}
