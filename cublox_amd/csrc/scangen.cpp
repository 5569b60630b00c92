// Synthetic LiDAR workload: a seeded scene of axis-aligned boxes on a ground
// plane, a rotating multi-beam sensor on a moving base, and fetch-query
// centres. Deterministic across machines: own SplitMix64-based counter RNG
// (no std::*_distribution) and a fixed evaluation order, so committed
// fixtures stay valid.
//
// Geometry of one scan k:
//   base pose    t_wb = (speed*k, 0, base_z), yaw = yaw_rate*k about +z
//   extrinsic    T_bs from the reference driver config
//                (config/cublox_driver.yaml:26-32)
//   sensor pose  T_ws = T_wb * T_bs, composed in float, fixed operand order
//   beams        elevations evenly spaced in [el_lo, el_hi], azimuths evenly
//                spaced over 360 deg, emitted azimuth-major (all beams of one
//                azimuth, then the next), like a spinning sensor
//   range        exact first hit (ground plane z=0 or a box) in double, plus
//                Gaussian noise; no hit within max_range -> a miss emitted at
//                max_range + miss_extra so the ray is range-clipped
//                (free space only) by the ray cast
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Uniform in (0, 1), 53-bit, never exactly 0.
inline double u01(uint64_t bits) {
  return (static_cast<double>(bits >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Counter-based stream: value j of stream (seed, tag).
inline double uniform_at(uint64_t seed, uint64_t tag, uint64_t j) {
  return u01(splitmix64(splitmix64(seed ^ (tag * 0xD1B54A32D192ED03ull)) + j));
}

// Box-Muller standard normal from two counter-based uniforms.
inline double normal_at(uint64_t seed, uint64_t tag, uint64_t j) {
  const double a = uniform_at(seed, tag, 2 * j);
  const double b = uniform_at(seed, tag, 2 * j + 1);
  return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
}

struct Box {
  double lo[3], hi[3];
};

} // namespace

extern "C" {

typedef struct {
  int n_boxes;
  float box_extent;    // box xy centres uniform in [-extent, extent]
  float box_min_side;  // sides uniform in [min, max] per axis
  float box_max_side;
  float corridor_half; // boxes whose footprint meets |y| < this are redrawn
  float speed;         // base advance along world +x per scan [m]
  float yaw_rate_deg;  // base yaw per scan [deg]
  float sensor_height; // sensor above the ground plane [m]
  float noise_sigma;   // range noise [m]
  float miss_extra;    // misses are emitted at max_range + miss_extra
  unsigned long long seed;
} cublox_scene_cfg;

typedef struct cublox_scene cublox_scene;

} // extern "C"

struct cublox_scene {
  cublox_scene_cfg cfg;
  std::vector<Box> boxes;
  float R_bs[9];
  float t_bs[3];
  float base_z;
};

namespace {

void quat_to_rot(float qx, float qy, float qz, float qw, float R[9]) {
  const float n = std::sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
  qx /= n;
  qy /= n;
  qz /= n;
  qw /= n;
  R[0] = 1.0f - 2.0f * (qy * qy + qz * qz);
  R[1] = 2.0f * (qx * qy - qz * qw);
  R[2] = 2.0f * (qx * qz + qy * qw);
  R[3] = 2.0f * (qx * qy + qz * qw);
  R[4] = 1.0f - 2.0f * (qx * qx + qz * qz);
  R[5] = 2.0f * (qy * qz - qx * qw);
  R[6] = 2.0f * (qx * qz - qy * qw);
  R[7] = 2.0f * (qy * qz + qx * qw);
  R[8] = 1.0f - 2.0f * (qx * qx + qy * qy);
}

// Sensor pose of scan k as row-major [R | t] (float), plus the base position.
void scan_pose(const cublox_scene &s, int k, float T[12], float base[3]) {
  const double yaw = static_cast<double>(s.cfg.yaw_rate_deg) * k * kPi / 180.0;
  const float c = static_cast<float>(std::cos(yaw));
  const float sn = static_cast<float>(std::sin(yaw));
  const float Rwb[9] = {c, -sn, 0.0f, sn, c, 0.0f, 0.0f, 0.0f, 1.0f};
  const float twb[3] = {s.cfg.speed * static_cast<float>(k), 0.0f, s.base_z};
  for (int r = 0; r < 3; ++r) {
    for (int col = 0; col < 3; ++col) {
      T[4 * r + col] = (Rwb[3 * r + 0] * s.R_bs[0 + col] +
                        Rwb[3 * r + 1] * s.R_bs[3 + col]) +
                       Rwb[3 * r + 2] * s.R_bs[6 + col];
    }
    T[4 * r + 3] = ((Rwb[3 * r + 0] * s.t_bs[0] + Rwb[3 * r + 1] * s.t_bs[1]) +
                    Rwb[3 * r + 2] * s.t_bs[2]) +
                   twb[r];
  }
  base[0] = twb[0];
  base[1] = twb[1];
  base[2] = twb[2];
}

// First intersection distance of the ray o + t d (|d| = 1) with the scene.
double cast(const std::vector<Box> &boxes, const std::vector<int> &cands,
            const double o[3], const double d[3]) {
  double best = INFINITY;
  if (d[2] < 0.0) best = -o[2] / d[2];
  for (int bi : cands) {
    const Box &b = boxes[static_cast<size_t>(bi)];
    double t0 = -INFINITY, t1 = INFINITY;
    bool miss = false;
    for (int a = 0; a < 3; ++a) {
      if (d[a] == 0.0) {
        if (o[a] < b.lo[a] || o[a] > b.hi[a]) {
          miss = true;
          break;
        }
        continue;
      }
      double ta = (b.lo[a] - o[a]) / d[a];
      double tb = (b.hi[a] - o[a]) / d[a];
      if (ta > tb) std::swap(ta, tb);
      t0 = std::max(t0, ta);
      t1 = std::min(t1, tb);
    }
    if (miss || t0 > t1 || t0 <= 0.0) continue;
    best = std::min(best, t0);
  }
  return best;
}

constexpr int kBins = 720;

int az_bin(double az) {
  double u = az / (2.0 * kPi);
  u -= std::floor(u);
  int b = static_cast<int>(u * kBins);
  return std::min(std::max(b, 0), kBins - 1);
}

// Buckets boxes by the azimuth interval they subtend from (ox, oy).
void bin_boxes(const std::vector<Box> &boxes, double ox, double oy,
               std::vector<std::vector<int>> &bins) {
  bins.assign(kBins, {});
  for (size_t i = 0; i < boxes.size(); ++i) {
    const Box &b = boxes[i];
    if (ox >= b.lo[0] && ox <= b.hi[0] && oy >= b.lo[1] && oy <= b.hi[1]) {
      for (auto &bin : bins) bin.push_back(static_cast<int>(i));
      continue;
    }
    const double cx = 0.5 * (b.lo[0] + b.hi[0]) - ox;
    const double cy = 0.5 * (b.lo[1] + b.hi[1]) - oy;
    const double ac = std::atan2(cy, cx);
    double lo = 0.0, hi = 0.0;
    for (int k = 0; k < 4; ++k) {
      const double x = (k & 1 ? b.hi[0] : b.lo[0]) - ox;
      const double y = (k & 2 ? b.hi[1] : b.lo[1]) - oy;
      double da = std::atan2(y, x) - ac;
      while (da > kPi) da -= 2.0 * kPi;
      while (da < -kPi) da += 2.0 * kPi;
      lo = std::min(lo, da);
      hi = std::max(hi, da);
    }
    const double step = 2.0 * kPi / kBins;
    const int nb = static_cast<int>(std::ceil((hi - lo) / step)) + 2;
    const int b0 = az_bin(ac + lo) - 1 + kBins;
    for (int j = 0; j < nb && j < kBins; ++j) {
      bins[static_cast<size_t>((b0 + j) % kBins)].push_back(static_cast<int>(i));
    }
  }
}

} // namespace

extern "C" {

cublox_scene *cublox_scene_create(const cublox_scene_cfg *cfg) {
  if (!cfg) return nullptr;
  cublox_scene *s = new cublox_scene();
  s->cfg = *cfg;
  // Extrinsic of the reference driver config (config/cublox_driver.yaml:26-32).
  quat_to_rot(0.0f, 0.0f, 0.99987200f, 0.01599934f, s->R_bs);
  s->t_bs[0] = 0.98837f;
  s->t_bs[1] = -0.06666f;
  s->t_bs[2] = -1.37611f;
  // Base height that puts the sensor at sensor_height (yaw keeps z).
  s->base_z = cfg->sensor_height - s->t_bs[2];
  const uint64_t seed = cfg->seed;
  uint64_t j = 0;
  for (int i = 0; i < cfg->n_boxes; ++i) {
    for (int tries = 0; tries < 1000; ++tries, j += 5) {
      const double W = cfg->box_extent;
      const double lo = cfg->box_min_side, hi = cfg->box_max_side;
      const double sx = lo + (hi - lo) * uniform_at(seed, 1, j + 0);
      const double sy = lo + (hi - lo) * uniform_at(seed, 1, j + 1);
      const double sz = lo + (hi - lo) * uniform_at(seed, 1, j + 2);
      const double cx = -W + 2.0 * W * uniform_at(seed, 1, j + 3);
      const double cy = -W + 2.0 * W * uniform_at(seed, 1, j + 4);
      Box b;
      b.lo[0] = cx - 0.5 * sx;
      b.hi[0] = cx + 0.5 * sx;
      b.lo[1] = cy - 0.5 * sy;
      b.hi[1] = cy + 0.5 * sy;
      b.lo[2] = 0.0;  // boxes rest on the ground
      b.hi[2] = sz;
      if (b.hi[1] > -cfg->corridor_half && b.lo[1] < cfg->corridor_half) continue;
      s->boxes.push_back(b);
      j += 5;
      break;
    }
  }
  return s;
}

void cublox_scene_destroy(cublox_scene *s) { delete s; }

int cublox_scene_box_count(const cublox_scene *s) {
  return s ? static_cast<int>(s->boxes.size()) : 0;
}

// lo/hi corners, 6 doubles per box.
void cublox_scene_boxes(const cublox_scene *s, double *out) {
  for (size_t i = 0; i < s->boxes.size(); ++i) {
    std::memcpy(out + 6 * i, s->boxes[i].lo, 3 * sizeof(double));
    std::memcpy(out + 6 * i + 3, s->boxes[i].hi, 3 * sizeof(double));
  }
}

void cublox_scene_pose(const cublox_scene *s, int k, float T[12],
                       float base[3]) {
  scan_pose(*s, k, T, base);
}

// Scan k: beams*azimuths sensor-frame points (azimuth-major), the sensor
// pose T (row-major 3x4) and the base position. Returns the point count, and
// writes the number of misses (clipped rays) to *n_miss when given.
int cublox_scene_scan(const cublox_scene *s, int k, int beams, int azimuths,
                      float el_lo_deg, float el_hi_deg, float max_range,
                      float *xs, float *ys, float *zs, float T[12],
                      float base[3], int *n_miss, int n_threads) {
  if (!s || beams <= 0 || azimuths <= 0) return 0;
  scan_pose(*s, k, T, base);
  const double o[3] = {T[3], T[7], T[11]};
  const double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
  std::vector<std::vector<int>> bins;
  bin_boxes(s->boxes, o[0], o[1], bins);
  const uint64_t tag = 0x5CA7ull + static_cast<uint64_t>(k) * 7919ull;
  const int n = beams * azimuths;
  int misses = 0;
#ifdef _OPENMP
  const int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt) reduction(+ : misses)
#endif
  for (int a = 0; a < azimuths; ++a) {
    const double az = 2.0 * kPi * a / azimuths;
    for (int b = 0; b < beams; ++b) {
      const double el_deg =
          beams == 1 ? 0.5 * (el_lo_deg + el_hi_deg)
                     : el_lo_deg + (el_hi_deg - el_lo_deg) * b / (beams - 1);
      const double el = el_deg * kPi / 180.0;
      const double ds[3] = {std::cos(el) * std::cos(az),
                            std::cos(el) * std::sin(az), std::sin(el)};
      const double dw[3] = {R[0] * ds[0] + R[1] * ds[1] + R[2] * ds[2],
                            R[3] * ds[0] + R[4] * ds[1] + R[5] * ds[2],
                            R[6] * ds[0] + R[7] * ds[1] + R[8] * ds[2]};
      const int bin = az_bin(std::atan2(dw[1], dw[0]));
      const double t = cast(s->boxes, bins[static_cast<size_t>(bin)], o, dw);
      const int idx = a * beams + b;
      double r;
      if (t <= static_cast<double>(max_range)) {
        r = t + static_cast<double>(s->cfg.noise_sigma) *
                    normal_at(s->cfg.seed, tag, static_cast<uint64_t>(idx));
      } else {
        r = static_cast<double>(max_range) + s->cfg.miss_extra;
        misses += 1;
      }
      xs[idx] = static_cast<float>(r * ds[0]);
      ys[idx] = static_cast<float>(r * ds[1]);
      zs[idx] = static_cast<float>(r * ds[2]);
    }
  }
  if (n_miss) *n_miss = misses;
  return n;
}

// n query centres for scan k: uniform in [-spread, spread] in x and y around
// the sensor, at the sensor height.
void cublox_scene_queries(const cublox_scene *s, int k, int n, float spread,
                          float *centres) {
  float T[12], base[3];
  scan_pose(*s, k, T, base);
  const uint64_t tag = 0x0E7Cull + static_cast<uint64_t>(k) * 104729ull;
  for (int i = 0; i < n; ++i) {
    const double ux = uniform_at(s->cfg.seed, tag, 2 * static_cast<uint64_t>(i));
    const double uy =
        uniform_at(s->cfg.seed, tag, 2 * static_cast<uint64_t>(i) + 1);
    centres[3 * i + 0] = T[3] + static_cast<float>((2.0 * ux - 1.0) * spread);
    centres[3 * i + 1] = T[7] + static_cast<float>((2.0 * uy - 1.0) * spread);
    centres[3 * i + 2] = T[11];
  }
}

// Generic counter-based uniform in [0, 1) for host-side workload draws
// (per-tick point counts, arrival times).
double cublox_rng_uniform(unsigned long long seed, unsigned long long tag,
                          unsigned long long j) {
  return uniform_at(seed, tag, j);
}

} // extern "C"
