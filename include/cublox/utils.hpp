// Host index math of the toroidal voxel window (Eigen-typed public API).
//
// Drop-in for the reference's include/cublox/utils.hpp:24-188: same names,
// signatures and results. Positions bin to global voxel indices; global
// indices fold into local window indices in [-half, half] by C remainder
// (the window is a ring buffer, so recentering moves no data); local indices
// flatten x-major with z fastest into hash ids, which are public API
// (OccupancyGrid::isOccupied(int), resetVoxel(int)).
#pragma once

#include <Eigen/Dense>
#include <algorithm>
#include <cmath>

namespace cublox {

// utils.hpp:24-31. Corner binning floors p/res; centre binning rounds half
// away from zero (bins are centred on multiples of res).
inline int posToGlobalIndex(const float pos, const float resolution_inv,
                            const bool origin_at_center) {
  const float q = pos * resolution_inv;
  if (origin_at_center) {
    const float half = pos < 0.0f ? -0.5f : 0.5f;
    return static_cast<int>(q + half);
  }
  return static_cast<int>(std::floor(q));
}

inline void posToGlobalIndex(const Eigen::Vector3f &pos,
                             const float resolution_inv,
                             const bool origin_at_center,
                             Eigen::Vector3i &id_g) {
  for (int a = 0; a < 3; ++a) {
    id_g(a) = posToGlobalIndex(pos(a), resolution_inv, origin_at_center);
  }
}

// utils.hpp:44-47: voxel centre (centre binning) or voxel centre of the
// [i, i+1) cell (corner binning).
inline float globalIndexToPos(const int id_g, const float resolution,
                              const bool origin_at_center) {
  const float g = static_cast<float>(id_g);
  return origin_at_center ? g * resolution : (g + 0.5f) * resolution;
}

inline void globalIndexToPos(const Eigen::Vector3i &id_g,
                             const float resolution,
                             const bool origin_at_center,
                             Eigen::Vector3f &pos) {
  for (int a = 0; a < 3; ++a) {
    pos(a) = globalIndexToPos(id_g(a), resolution, origin_at_center);
  }
}

// utils.hpp:60-69
inline int globalIndexToLocalIndex(const int id_g, const int map_size_i,
                                   const int half_map_size_i) {
  const int r = id_g % map_size_i;
  if (r > half_map_size_i) return r - map_size_i;
  if (r < -half_map_size_i) return r + map_size_i;
  return r;
}

inline void globalIndexToLocalIndex(const Eigen::Vector3i &id_g,
                                    const Eigen::Vector3i &map_size_i,
                                    const Eigen::Vector3i &half_map_size_i,
                                    Eigen::Vector3i &id_l) {
  for (int a = 0; a < 3; ++a) {
    id_l(a) = globalIndexToLocalIndex(id_g(a), map_size_i(a),
                                      half_map_size_i(a));
  }
}

// utils.hpp:82-92: the global index in the window around origin_g whose
// local index is id_l.
inline int localIndexToGlobalIndex(const int id_l, const int map_size_i,
                                   const int half_map_size_i,
                                   const int origin_g) {
  const int lowest_g = origin_g - half_map_size_i;
  const int lowest_l =
      globalIndexToLocalIndex(lowest_g, map_size_i, half_map_size_i);
  int offset = id_l - lowest_l;
  if (offset < 0) offset += map_size_i;
  return lowest_g + offset;
}

inline void localIndexToGlobalIndex(const Eigen::Vector3i &id_l,
                                    const Eigen::Vector3i &map_size_i,
                                    const Eigen::Vector3i &half_map_size_i,
                                    const Eigen::Vector3i &local_map_origin_i,
                                    Eigen::Vector3i &id_g) {
  for (int a = 0; a < 3; ++a) {
    id_g(a) = localIndexToGlobalIndex(id_l(a), map_size_i(a),
                                      half_map_size_i(a),
                                      local_map_origin_i(a));
  }
}

// utils.hpp:109-119
inline void localIndexToPos(const Eigen::Vector3i &id_l,
                            const Eigen::Vector3i &map_size_i,
                            const Eigen::Vector3i &half_map_size_i,
                            const Eigen::Vector3i &local_map_origin_i,
                            const float resolution, const bool origin_at_center,
                            Eigen::Vector3f &pos) {
  Eigen::Vector3i id_g;
  localIndexToGlobalIndex(id_l, map_size_i, half_map_size_i,
                          local_map_origin_i, id_g);
  globalIndexToPos(id_g, resolution, origin_at_center, pos);
}

// utils.hpp:122-131
inline void hashIdToLocalIndex(const int hash_id,
                               const Eigen::Vector3i &map_size_i,
                               const Eigen::Vector3i &half_map_size_i,
                               Eigen::Vector3i &id_l) {
  const int plane = map_size_i.y() * map_size_i.z();
  const int ix = hash_id / plane;
  const int rem = hash_id - ix * plane;
  const int iy = rem / map_size_i.z();
  const int iz = rem - iy * map_size_i.z();
  id_l = Eigen::Vector3i(ix - half_map_size_i.x(), iy - half_map_size_i.y(),
                         iz - half_map_size_i.z());
}

// utils.hpp:133-139
inline int localIndexToHashId(const Eigen::Vector3i &id_l,
                              const Eigen::Vector3i &map_size_i,
                              const Eigen::Vector3i &half_map_size_i) {
  const int ix = id_l.x() + half_map_size_i.x();
  const int iy = id_l.y() + half_map_size_i.y();
  const int iz = id_l.z() + half_map_size_i.z();
  return ix * map_size_i.y() * map_size_i.z() + iy * map_size_i.z() + iz;
}

// utils.hpp:142-151
inline void hashIdToGlobalIndex(const int hash_id,
                                const Eigen::Vector3i &map_size_i,
                                const Eigen::Vector3i &half_map_size_i,
                                const Eigen::Vector3i &local_map_origin_i,
                                Eigen::Vector3i &id_g) {
  Eigen::Vector3i id_l;
  hashIdToLocalIndex(hash_id, map_size_i, half_map_size_i, id_l);
  localIndexToGlobalIndex(id_l, map_size_i, half_map_size_i,
                          local_map_origin_i, id_g);
}

// utils.hpp:153-162
inline void hashIdToPos(const int hash_id, const Eigen::Vector3i &map_size_i,
                        const Eigen::Vector3i &half_map_size_i,
                        const Eigen::Vector3i &local_map_origin_i,
                        const float resolution, const bool origin_at_center,
                        Eigen::Vector3f &pos) {
  Eigen::Vector3i id_l;
  hashIdToLocalIndex(hash_id, map_size_i, half_map_size_i, id_l);
  localIndexToPos(id_l, map_size_i, half_map_size_i, local_map_origin_i,
                  resolution, origin_at_center, pos);
}

// utils.hpp:165-171
inline int globalIndexToHashId(const Eigen::Vector3i &id_g,
                               const Eigen::Vector3i &map_size_i,
                               const Eigen::Vector3i &half_map_size_i) {
  Eigen::Vector3i id_l;
  globalIndexToLocalIndex(id_g, map_size_i, half_map_size_i, id_l);
  return localIndexToHashId(id_l, map_size_i, half_map_size_i);
}

// utils.hpp:173-181
inline int posToHashIndex(const Eigen::Vector3f &pos,
                          const Eigen::Vector3i &map_size_i,
                          const Eigen::Vector3i &half_map_size_i,
                          const float resolution_inv,
                          const bool origin_at_center) {
  Eigen::Vector3i id_g;
  posToGlobalIndex(pos, resolution_inv, origin_at_center, id_g);
  return globalIndexToHashId(id_g, map_size_i, half_map_size_i);
}

// utils.hpp:185-188: probability -> log-odds, p clamped to [1e-6, 1-1e-6].
inline float logit(const float p) {
  const float pc = std::min(std::max(p, 1.0e-6f), 1.0f - 1.0e-6f);
  return std::log(pc / (1.0f - pc));
}

} // namespace cublox
