// GPU occupancy grid on MI355X (drop-in for the reference's
// include/cublox/OccupancyGrid.hpp:29-169).
//
// The public interface is the reference's: update / recenter /
// fetchOccupancyAround / point queries / log-odds setters, with the same
// defaults and blocking semantics. Private members differ: buffers are
// plain device pointers behind the C ABI in cublox/capi.h, every grid owns
// its own HIP stream, and this header includes no HIP or CUDA header.
//
// Additions with no reference counterpart are marked "extension".
#pragma once

#include <Eigen/Dense>
#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include <cublox/Grid.hpp>

namespace cublox {

using PointCloud = Eigen::Matrix<float, Eigen::Dynamic, 3>;

class GridBatch;

class OccupancyGrid : public Grid {
public:
  enum class VoxelState {
    UNKNOWN = 0,
    FREE = 1,
    OCCUPIED = 2,
  };

  struct OccupancySample {
    Eigen::Vector3f position;
    VoxelState state;
  };

  OccupancyGrid(const Eigen::Vector3i &half_map_size_i, const float resolution,
                const bool origin_at_center,
                const std::optional<double> recenter_threshold,
                const Eigen::Vector3f &origin = Eigen::Vector3f::Zero());
  OccupancyGrid() = default;
  ~OccupancyGrid() override;

  // Owning device buffers: not copyable, movable.
  OccupancyGrid(const OccupancyGrid &) = delete;
  OccupancyGrid &operator=(const OccupancyGrid &) = delete;
  OccupancyGrid(OccupancyGrid &&other) noexcept;
  OccupancyGrid &operator=(OccupancyGrid &&other) noexcept;

  void reset() override;
  void resetVoxel(const int hash_id) override;

  // One ray-cast pass of `input_cloud` (world frame, N x 3) from
  // `sensor_origin`, then the clamped log-odds fold. Blocks until done.
  void update(const PointCloud &input_cloud,
              const Eigen::Vector3f &sensor_origin);

  // Sliding-window recenter: host slab plan, one device launch. The window
  // origin and bounds change at once; the slab clear may still run when the
  // call returns, and every later call on this grid is ordered after it
  // (the next update's ray walk overlaps it).
  void recenter(const Eigen::Vector3f &pos);

  // Voxels within `radius` [m] of `center`, optionally of one state.
  // `max_results` caps the count (0 = the 500000-entry default cap).
  std::vector<OccupancySample> fetchOccupancyAround(
      const Eigen::Vector3f &center, const float radius,
      const std::optional<VoxelState> state_filter = std::nullopt,
      const int max_results = 0) const;

  inline void setMaxRaycastRange(const float range) {
    max_raycast_range_ = range;
  }
  inline float getMaxRaycastRange() const { return max_raycast_range_; }

  void setLogOddsParams(const float l_hit, const float l_miss,
                        const float l_min, const float l_max,
                        const float l_free, const float l_occupied) {
    l_hit_ = l_hit;
    l_miss_ = l_miss;
    l_min_ = l_min;
    l_max_ = l_max;
    l_free_ = l_free;
    l_occupied_ = l_occupied;
    l_unknown_ = 0.0f;
  }

  inline float getResolutionInv() const { return config_.resolution_inv; }
  inline bool getOriginAtCenter() const { return config_.origin_at_center; }

  bool isOccupied(const Eigen::Vector3f &pos) const;
  bool isUnknown(const Eigen::Vector3f &pos) const;
  bool isFree(const Eigen::Vector3f &pos) const;
  bool isOccupied(const Eigen::Vector3i &id_g) const;
  bool isUnknown(const Eigen::Vector3i &id_g) const;
  bool isFree(const Eigen::Vector3i &id_g) const;
  bool isOccupied(const int hash_id) const;
  bool isUnknown(const int hash_id) const;
  bool isFree(const int hash_id) const;

  // ---- extensions -------------------------------------------------------
  // update() with the three host columns given directly (no PointCloud).
  void updateColumns(const float *x, const float *y, const float *z, int n,
                     const Eigen::Vector3f &sensor_origin);
  // updateSensorFrame() with host columns and T row-major 3x4 [R | t].
  void updateSensorFrameColumns(const float *xs, const float *ys,
                                const float *zs, int n, const float T[12]);
  // Scan transform fused into the ray cast: `sensor_cloud` in the sensor
  // frame, world pose [R | t]; non-finite points are skipped. Equivalent to
  // update(R*p + t for finite p, t).
  void updateSensorFrame(const PointCloud &sensor_cloud,
                         const Eigen::Matrix3f &R, const Eigen::Vector3f &t);
  // Same, with the sensor-frame SoA columns already in device memory and T
  // row-major 3x4 [R | t]. Blocks.
  void updateSensorFrameDevice(const float *d_xs, const float *d_ys,
                               const float *d_zs, int n, const float T[12]);
  // World-frame SoA columns already in device memory. Blocks.
  void updateDevice(const float *d_x, const float *d_y, const float *d_z,
                    int n, const Eigen::Vector3f &sensor_origin);
  // The two halves of update, for stage-wise checks: castRays leaves the
  // per-voxel counters populated, applyCounters folds and zeroes them.
  void castRaysDevice(const float *d_x, const float *d_y, const float *d_z,
                      int n, const Eigen::Vector3f &sensor_origin);
  void castRaysSensorFrameDevice(const float *d_xs, const float *d_ys,
                                 const float *d_zs, int n, const float T[12]);
  void applyCounters();
  // Work census of the pending counters: out = {voxels with op != 0,
  // sum of op (voxel visits), sum of hit, voxels with hit != 0}.
  void counterCensus(unsigned long long out[4]) const;

  // Batched sphere fetch: one launch for all centres. counts[q] is the
  // reference's returned count for query q (min(logical count, cap));
  // positions (x, y, z per sample) and states hold the samples of query 0,
  // then query 1, and so on. Order within a query is unspecified. Blocks.
  void fetchOccupancyBatch(const std::vector<Eigen::Vector3f> &centers,
                           const float radius,
                           const std::optional<VoxelState> state_filter,
                           const int per_query_cap,
                           std::vector<unsigned int> &counts,
                           std::vector<float> &positions,
                           std::vector<unsigned char> &states) const;

  // How update() integrates a scan. FUSED_TILES (default) bins ray segments
  // by tile and folds counts straight into occ; PER_VOXEL_ATOMICS is the
  // reference's shape (global op/hit counters, then a sweep over all
  // voxels). Both give the same occupancy.
  enum class UpdatePath { FUSED_TILES, PER_VOXEL_ATOMICS };
  void setUpdatePath(UpdatePath p);
  UpdatePath getUpdatePath() const { return update_path_; }
  // Steps of each ray handled in the near-sensor cube by the fused path.
  void setNearSteps(int k);

  // Fused path, counts only: the per-voxel op/hit counts the fused kernels
  // produce are written to the counter buffers (read them with
  // copyCountersToHost, fold them with applyCounters).
  void castRaysFusedSensorFrameDevice(const float *d_xs, const float *d_ys,
                                      const float *d_zs, int n,
                                      const float T[12]);

  // ---- driver ingest on the device (reference src/main.cpp) -------------
  // A PointCloud2 xyz payload: width * height records of point_step bytes,
  // float32 x/y/z at byte offsets off_x/off_y/off_z (all four multiples of
  // 4), little-endian, sensor frame.
  struct PointCloud2Layout {
    uint32_t width{0}, height{1}, point_step{12};
    uint32_t off_x{0}, off_y{4}, off_z{8};
  };
  static constexpr int kMaxCloudPoints = 1000000;  // main.cpp:54
  // buildCloudInOdomFrame (main.cpp:640-693) then update(): the first
  // kMaxCloudPoints records (main.cpp:655-660), kept when finite and within
  // the max raycast range (main.cpp:672-679), transformed by T (row-major
  // 3x4 [R | t], sensor -> world) and integrated with t as the sensor
  // origin. Returns false, updating nothing, when no point is kept
  // (main.cpp:686-690). Runs on the fused path. Blocks. `data_bytes` is the
  // payload's length (the message's data size): std::invalid_argument when
  // it is shorter than the records read (min(width * height, 1 M) *
  // point_step bytes).
  bool updatePointCloud2(const void *data, const PointCloud2Layout &layout,
                         const float T[12], size_t data_bytes);
  bool updatePointCloud2Device(const void *d_data,
                               const PointCloud2Layout &layout,
                               const float T[12], size_t data_bytes);
  // depthCallback (main.cpp:558-638) then the above: the image (row pitch
  // `step` bytes) is back-projected with K = {fx, fy, cx, cy} into the
  // sensor-frame cloud the driver publishes, invalid pixels as NaN points.
  enum class DepthEncoding { MILLIMETRES_U16, METRES_F32 };
  bool updateDepthImage(const void *data, uint32_t width, uint32_t height,
                        uint32_t step, DepthEncoding encoding,
                        const float K[4], const float T[12]);
  bool updateDepthImageDevice(const void *d_data, uint32_t width,
                              uint32_t height, uint32_t step,
                              DepthEncoding encoding, const float K[4],
                              const float T[12]);

  // The driver's occupancy viz cloud (main.cpp:484-491 fetch, 740-817
  // build) in one launch: the OCCUPIED fetch around robot_pos with radius
  // min(max_viz_radius, maxHorizontalInWindowRadius(robot_pos)) and
  // max_points results, each fetched sample dropped by the global-index
  // subsample (main.cpp:71-84) or kept as (x, y, z, rgb) with the driver's
  // distance colour packed into a float (main.cpp:275-304). Record order is
  // unspecified. Returns the driver's fetched sample count. Blocks.
  struct VizPoint {
    float x, y, z, rgb;
  };
  size_t buildOccupancyVizCloud(const Eigen::Vector3f &robot_pos,
                                float max_viz_radius, int subsample,
                                int max_points,
                                std::vector<VizPoint> &out) const;

  // Map snapshot (no reference counterpart; format in cublox/snapshot.h):
  // geometry, window origin, settings and the occupancy, dense or as its
  // nonzero voxels, whichever is smaller. Blocks.
  std::vector<uint8_t> saveSnapshot() const;
  // Restores a snapshot of a grid with the same half map size, resolution
  // and binning (std::invalid_argument otherwise, or for a damaged
  // snapshot): occupancy, origin, first-run flag, recenter threshold,
  // log-odds parameters and ray-cast range. Blocks.
  void loadSnapshot(const uint8_t *data, size_t bytes);
  void saveSnapshotFile(const std::string &path) const;
  void loadSnapshotFile(const std::string &path);

  // Changed-voxel stream (the reference's modified-voxel compaction,
  // kernels.cu:202-209, which it compiles but never enables). capacity > 0:
  // every later update also records (hash id, new log-odds) of each voxel
  // whose stored value changed, at most `capacity` entries per update; 0
  // turns it off. Recentering is not recorded. On PER_VOXEL_ATOMICS the
  // capacity must be at least the voxel count (the reference's gate,
  // kernels.cu:251-262).
  void setChangedVoxelCapacity(unsigned int capacity);
  unsigned int getChangedVoxelCapacity() const { return delta_capacity_; }
  struct ChangedVoxel {
    int hash_id;
    float log_odds;
  };
  // The changes of the last update, in unspecified order (blocks). Returns
  // how many voxels changed: more than out.size() when the capacity was too
  // small.
  unsigned long long changedVoxels(std::vector<ChangedVoxel> &out) const;
  // Device views of the stream (count, hash ids, values), valid until the
  // next update on this grid's stream.
  const unsigned int *deviceChangedCount() const { return d_delta_count_; }
  const int *deviceChangedIds() const { return d_delta_idx_; }
  const float *deviceChangedValues() const { return d_delta_val_; }

  // Per-kernel device time measured with HIP events on this grid's stream
  // (off by default). update_ms covers the ray cast (and, on the fused
  // path, the fold); apply_ms the separate fold of the per-voxel path.
  // Recenter also accumulates the voxels it zeroed.
  // fetch_ms covers the batched fetch's classification kernel
  // (fetchOccupancyBatch), with the voxels its queries tested and the
  // samples they found.
  struct KernelTimes {
    double update_ms{0.0}, apply_ms{0.0}, recenter_ms{0.0}, fetch_ms{0.0};
    long long update_launches{0}, apply_launches{0}, recenter_launches{0};
    long long recenter_zeroed_voxels{0};
    long long fetch_launches{0}, fetch_tested_voxels{0}, fetch_samples{0};
  };
  void enableKernelTiming(bool on);
  KernelTimes kernelTimes() const;
  void resetKernelTimes() { times_ = KernelTimes{}; }

  // Device views and host copies for integration and tests.
  const float *deviceOccupancy() const { return d_occ_; }
  void copyOccupancyToHost(float *dst) const;
  void copyCountersToHost(int *op, int *hit) const;
  void loadOccupancyFromHost(const float *src);
  // The grid's stream; work enqueued on it after this call is ordered after
  // every call made on the grid so far (a pending recenter's clear
  // included).
  void *stream() const {
    join();
    return stream_;
  }
  float getLogOddsFree() const { return l_free_; }
  float getLogOddsOccupied() const { return l_occupied_; }

private:
  friend class GridBatch;  // batched updates read the grid's state

  bool first_run_{true};

  float storedToLogit(const float s) const noexcept { return s + l_unknown_; }
  bool occOccupied(const float s) const noexcept {
    return storedToLogit(s) >= l_occupied_;
  }
  bool occFree(const float s) const noexcept {
    return storedToLogit(s) < l_free_;
  }
  bool occUnknown(const float s) const noexcept {
    const float l = storedToLogit(s);
    return l >= l_free_ && l < l_occupied_;
  }

  float storedAt(const int hash_id) const;
  void ensureCloudCapacity(const int n);
  void allocateVoxelBuffers();
  void ensureFetchCapacity(const long long n) const;
  void releaseAll() noexcept;
  void pinFirstOrigin(const Eigen::Vector3f &sensor_origin);
  void castRaysWorld(const float *d_x, const float *d_y, const float *d_z,
                     int n, const Eigen::Vector3f &sensor_origin);
  void castRaysSensor(const float *d_xs, const float *d_ys, const float *d_zs,
                      int n, const float T[12]);
  void launchApply();
  void launchIntegrate(const float *d_xs, const float *d_ys, const float *d_zs,
                       int n, const float *T, const Eigen::Vector3f &origin,
                       bool export_counts, int stride = 1, float drop_sq = 0.0f);
  bool ingest(const float *d_xs, const float *d_ys, const float *d_zs,
              int stride, int n, const float T[12]);
  const float *rawFloats(const void *d_data, uint32_t off) const;
  void ensureCounters();
  void ensureWorkspace(int n);
  void sync() const;
  void mark(int ev) const;
  void harvestTimes() const;

  void *stream_{nullptr};

  // A recenter plans its slab clear on the host into zplan_ and returns:
  // the next fused update carries it into its walk launch, whose extra
  // workgroups clear the slabs beside the walk (the folds come after);
  // every other call that reads or writes occ or the counters first
  // launches it on stream_ (join).
  std::vector<unsigned char> zplan_;
  mutable bool zero_pending_{false};
  void join() const;
  void settle() const;  // join, then a host wait (for work on other streams)

  // Timing events: [0,1] ray cast, [2,3] apply, [4,5] recenter (with
  // timing on, a recenter launches its clear at once, between them), [6,7]
  // fetch.
  bool timing_{false};
  void *events_[8]{nullptr, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr, nullptr};
  mutable bool pending_[3]{false, false, false};
  mutable long long pending_zeroed_{0};
  mutable KernelTimes times_{};

  int *d_op_cnt_{nullptr};
  int *d_hit_cnt_{nullptr};
  float *d_occ_{nullptr};

  unsigned int *d_kept_{nullptr};

  unsigned int *d_delta_count_{nullptr};
  int *d_delta_idx_{nullptr};
  float *d_delta_val_{nullptr};
  unsigned int delta_capacity_{0};

  float *d_cloud_x_{nullptr};
  float *d_cloud_y_{nullptr};
  float *d_cloud_z_{nullptr};
  int cloud_capacity_{0};

  mutable unsigned int *d_fetch_count_{nullptr};
  mutable float *d_fetch_pos_{nullptr};
  mutable unsigned char *d_fetch_state_{nullptr};
  mutable void *d_fetch_cfgs_{nullptr};
  mutable unsigned int *d_fetch_counts_{nullptr};
  mutable float *d_fetch_dense_pos_{nullptr};
  mutable unsigned char *d_fetch_dense_state_{nullptr};
  mutable unsigned int *d_fetch_offsets_{nullptr};
  // Pinned staging of host inputs (clouds, payloads, images): a host copy
  // into it, then asynchronous DMA, in pieces so the two overlap.
  void *h_upload_{nullptr};
  void *h_upload_dev_{nullptr};  // its device address (stageColumns)
  mutable bool stage_in_use_{false};  // kernels may read the stage until a sync
  size_t h_upload_bytes_{0};
  void stageColumns(const float *x, const float *y, const float *z, int n,
                    const float **dx, const float **dy, const float **dz);
  uint8_t *readStage(size_t bytes);
  const void *stageHost(const void *src, size_t bytes);
  void uploadHost(void *d_dst, const void *const *h_srcs, const size_t *lens,
                  int parts);
  void uploadColumns(const float *x, const float *y, const float *z, int n);
  mutable void *h_stage_{nullptr};  // pinned staging of downloads and snapshots
  mutable size_t h_stage_bytes_{0};
  uint8_t *pinnedStage(size_t bytes) const;
  void downloadHost(void *const *h_dsts, const void *const *d_srcs, const size_t *lens,
                    int parts) const;  // at most 8 parts
  mutable float *d_viz_{nullptr};
  mutable unsigned int *d_viz_count_{nullptr};
  mutable long long viz_capacity_{0};
  mutable long long fetch_dense_capacity_{0};
  mutable int fetch_offsets_capacity_{0};
  mutable long long fetch_capacity_{0};
  mutable int fetch_cfg_capacity_{0};

  float max_raycast_range_{25.0f};

  UpdatePath update_path_{UpdatePath::FUSED_TILES};
  int near_steps_{8};
  void *d_ws_{nullptr};
  size_t ws_bytes_{0};
  int ws_points_{0};
  float ws_range_{0.0f};
  int ws_near_{0};

  // Defaults of the reference (OccupancyGrid.hpp:162-168).
  float l_hit_{0.847f};
  float l_miss_{-0.405f};
  float l_min_{-1.992f};
  float l_max_{3.476f};
  float l_free_{-0.0004f};
  float l_occupied_{1.7346f};
  float l_unknown_{0.0f};
};

} // namespace cublox
