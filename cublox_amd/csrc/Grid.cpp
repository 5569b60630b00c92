// Sliding-window geometry (reference src/Grid.cpp:27-92).
#include <cublox/Grid.hpp>
#include <cublox/utils.hpp>

#include <cstdint>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

namespace cublox {

Grid::Grid(const Eigen::Vector3i &half_map_size_i, const float resolution,
           const std::optional<float> recenter_threshold,
           const bool origin_at_center) {
  config_.resolution = resolution;
  config_.resolution_inv = 1.0f / resolution;
  config_.half_map_size_i = half_map_size_i;
  config_.map_size_i = 2 * half_map_size_i + Eigen::Vector3i::Ones();
  config_.recenter_threshold = recenter_threshold;
  if (recenter_threshold) {
    config_.recenter_threshold_squared = *recenter_threshold * *recenter_threshold;
  }
  config_.origin_at_center = origin_at_center;

  std::int64_t total = 1;
  for (int a = 0; a < 3; ++a) total *= config_.map_size_i(a);
  if (total > std::numeric_limits<int>::max()) {
    throw std::runtime_error(
        "cublox::Grid: requested voxel_num = " + std::to_string(total) +
        " exceeds int32 max (" +
        std::to_string(std::numeric_limits<int>::max()) +
        "). Pick a smaller half_map_size_i.");
  }
  config_.voxel_num = static_cast<int>(total);
  std::cout << "cublox::Grid initialized with " << config_.voxel_num
            << " voxels." << std::endl;
}

bool Grid::inside(const Eigen::Vector3f &pos) const {
  Eigen::Vector3i id_g;
  posToGlobalIndex(pos, config_.resolution_inv, config_.origin_at_center, id_g);
  return inside(id_g);
}

bool Grid::inside(const Eigen::Vector3i &id_g) const {
  return ((id_g - origin_i_).cwiseAbs().array() <=
          config_.half_map_size_i.array())
      .all();
}

float Grid::maxHorizontalInWindowRadius(const Eigen::Vector3f &center) const {
  Eigen::Vector3i center_g;
  posToGlobalIndex(center, config_.resolution_inv, config_.origin_at_center,
                   center_g);
  const Eigen::Vector3i off = (center_g - origin_i_).cwiseAbs();
  const int mx = config_.half_map_size_i.x() - off.x();
  const int my = config_.half_map_size_i.y() - off.y();
  if (mx < 0 || my < 0) return 0.0f;
  return static_cast<float>(mx < my ? mx : my) * config_.resolution;
}

void Grid::updateOriginAndBound(const Eigen::Vector3i &new_origin_i) {
  origin_i_ = new_origin_i;
  bound_min_i_ = origin_i_ - config_.half_map_size_i;
  bound_max_i_ = origin_i_ + config_.half_map_size_i;
  globalIndexToPos(origin_i_, config_.resolution, config_.origin_at_center,
                   origin_f_);
  globalIndexToPos(bound_min_i_, config_.resolution, config_.origin_at_center,
                   bound_min_f_);
  globalIndexToPos(bound_max_i_, config_.resolution, config_.origin_at_center,
                   bound_max_f_);
}

} // namespace cublox
