// Sliding-window voxel geometry (drop-in for the reference's
// include/cublox/Grid.hpp:22-57: same constructor, methods and public data).
#pragma once

#include <Eigen/Dense>
#include <optional>

namespace cublox {

class Grid {
public:
  // map_size = 2*half + 1 per axis; throws std::runtime_error when the voxel
  // count exceeds INT_MAX (hash ids are int).
  Grid(const Eigen::Vector3i &half_map_size_i, const float resolution,
       const std::optional<float> recenter_threshold = std::nullopt,
       const bool origin_at_center = false);
  Grid() = default;
  // Virtual so that deleting through a Grid* is defined (the reference's
  // destructor is not virtual, Grid.hpp:28).
  virtual ~Grid() = default;

  bool inside(const Eigen::Vector3f &pos) const;
  bool inside(const Eigen::Vector3i &id_g) const;
  void updateOriginAndBound(const Eigen::Vector3i &new_origin_i);

  virtual void reset() = 0;
  virtual void resetVoxel(const int hash_id) = 0;

  Eigen::Vector3i getOriginIndex() const { return origin_i_; }
  Eigen::Vector3f getOriginPosition() const { return origin_f_; }

  // Horizontal (xy) distance from `center` to the window edge [m], 0 when
  // the centre is outside the window in x or y.
  float maxHorizontalInWindowRadius(const Eigen::Vector3f &center) const;

  struct Config {
    float resolution{0.0};
    float resolution_inv{0.0};
    Eigen::Vector3i map_size_i{Eigen::Vector3i::Zero()};
    Eigen::Vector3i half_map_size_i{Eigen::Vector3i::Zero()};
    int voxel_num{0};
    bool origin_at_center{false};
    std::optional<float> recenter_threshold{std::nullopt};
    float recenter_threshold_squared{0.0};
  } config_;

  Eigen::Vector3f origin_f_, bound_min_f_, bound_max_f_;
  Eigen::Vector3i origin_i_, bound_min_i_, bound_max_i_;
};

} // namespace cublox
