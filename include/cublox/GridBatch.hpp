// Batched updates of many occupancy grids (extension; the reference updates
// one grid per call, OccupancyGrid.cpp:181-238).
//
// One tick runs the fused update of every listed scan, each on its own
// grid, and then the recenter of every listed grid, as one launch sequence
// each on the batch's stream: every grid's rays, tiles and slabs drain
// through the same kernels, so a tick fills the GPU with any mix of grids.
// Each grid ends the tick exactly as its own update() + recenter() would
// leave it. The per-grid descriptors travel in one host->device copy at the
// head of each sequence; with graph mode on, the sequences are captured
// once and each tick restages and replays them.
//
// Pipelined mode overlaps consecutive ticks: each tick's first half (ray
// walk, record counting, tile scan, placing: no occupancy access) runs on
// one of two streams and workspaces beside the folds of the tick before,
// and every tick's folds and recenters wait for the previous tick's, so
// every grid still sees its updates in tick order.
//
// The online autotuner cycles candidate settings over the first ticks
// (results do not depend on them) and keeps the fastest by device time.
#pragma once

#include <Eigen/Dense>
#include <array>
#include <utility>
#include <vector>

#include <cublox/OccupancyGrid.hpp>

namespace cublox {

struct BatchLimits {
  int max_grids{64};
  int max_points{1 << 22};          // points per tick, all grids
  int max_tiles{1 << 20};           // sum of the grids' tile counts
  int max_segments_per_ray{96};     // max over grids (segmentBound)
  int max_near_steps{12};
  int min_chunk{256};               // smallest hot-tile chunk any tick uses
  int max_slabs{65536};             // 32 KB scratch slabs
  int max_regions{0};               // recenter regions per tick (0: 6 per grid)
  // Grids with a changed-voxel stream (OccupancyGrid::setChangedVoxelCapacity)
  // get it filled by each tick; needs this set (other grids pay nothing,
  // but the folds of every tick take the stream-feeding variants).
  bool changed_voxels{false};
};

// One grid's scan of one tick: device-resident columns, sensor-frame with
// T (row-major 3x4 [R | t] sensor->world) or world-frame from `origin`.
struct GridScan {
  OccupancyGrid *grid{nullptr};
  const float *d_x{nullptr}, *d_y{nullptr}, *d_z{nullptr};
  int n{0};
  bool sensor_frame{true};
  std::array<float, 12> T{};
  Eigen::Vector3f origin{Eigen::Vector3f::Zero()};
  // Point layout and the driver's range drop (cublox_integrate_grid):
  // point k is (d_x[k*s], d_y[k*s], d_z[k*s]), s = point_stride (0: 1), so
  // a PointCloud2 payload is read in place; range_drop_sq > 0 drops
  // sensor-frame points with |p|^2 above it (main.cpp:676-679). The
  // driver's skip of a cloud with no kept point is not applied here.
  int point_stride{0};
  float range_drop_sq{0.0f};
};

// One autotuner setting (none changes any result):
//   near_steps  ray steps counted in the LDS cube around the sensor voxel
//               before rays are cut into tile segments, 1..max_near_steps;
//   chunk       records per work item of a hot (split) tile;
//   lanes       lanes per ray of the walk, i.e. rays per wavefront: 1 (64
//               rays) or 4 (16 rays); 0 picks by the batch capacity;
//   hot_mode    deduplication of a hot tile's chunks: 0 a 32 KB slab of
//               LDS counts per chunk, summed by the fold; 1 one shared
//               op/hit slab pair per tile that the chunks add their
//               nonzero counts to with global atomics.
struct TuneSetting {
  int near_steps{8};
  int chunk{8192};
  int lanes{0};
  int hot_mode{0};
};

class GridBatch {
public:
  explicit GridBatch(const BatchLimits &limits);
  ~GridBatch();
  GridBatch(const GridBatch &) = delete;
  GridBatch &operator=(const GridBatch &) = delete;

  // Per-ray segment bound and tile count of a grid (for the limits).
  static int segmentBound(const OccupancyGrid &g);
  static int tileCount(const OccupancyGrid &g);

  // One tick: update every scan's grid, then recenter each listed grid
  // toward its position (as recenter(pos)). Blocks until done.
  void tick(const std::vector<GridScan> &scans,
            const std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>>
                &recenters);
  // tick() in two halves: submit() stages and launches the tick on stream()
  // and returns at once, with the recentered grids' origins already moved
  // (work the caller enqueues on stream() afterwards sees the grids after
  // the tick); wait() blocks until every submitted tick has finished. Without
  // pipelining every submit() needs its wait() before the next submit();
  // pipelined, a submit() first waits for the tick submitted two before it,
  // so at most two ticks are in flight. The scans' device columns must stay
  // valid until their tick has finished, and the grids must not be used
  // directly while a tick is in flight.
  void submit(const std::vector<GridScan> &scans,
              const std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>>
                  &recenters);
  void wait();
  // Blocks until the oldest submitted tick still in flight has finished
  // (the last one when not pipelined); last*() then describe it.
  void waitOldest();
  bool inFlight() const;

  // Pipelined ticks (a second stream and workspace, allocated when first
  // switched on). Not with changed-voxel streams: a tick's stream restarts
  // in its first half. Autotuning ticks run one at a time either way.
  void setPipelined(bool on);
  bool pipelined() const { return n_slots_ == 2; }

  // Capture the tick's launch sequences in a HIP graph on the next tick and
  // replay it afterwards.
  void setGraphMode(bool on);
  bool graphMode() const { return graph_mode_; }

  // Online autotuning: each candidate runs `ticks_each` ticks (after one
  // warm-up tick per candidate), then the fastest by mean device time is
  // kept. An empty list disables tuning.
  void setTuning(const std::vector<TuneSetting> &candidates, int ticks_each);
  bool tuning() const { return tune_pos_ < tune_.size(); }
  TuneSetting setting() const { return setting_; }
  void setSetting(const TuneSetting &s);
  std::vector<std::pair<TuneSetting, double>> tuneResults() const;

  // Device time of the last finished tick's update and recenter sequences
  // [ms] (span on its stream, overlap included), and the voxels its
  // recenters zeroed.
  double lastUpdateMs() const { return last_update_ms_; }
  double lastRecenterMs() const { return last_recenter_ms_; }
  long long lastZeroedVoxels() const { return last_zeroed_; }
  // Counters of the last tick's update: work items, tile records, overflow
  // flag, split hot tiles, unsplit hot tiles.
  std::array<int, 5> lastCounters() const;
  // The stream of the last submitted tick.
  void *stream() const { return slots_[cur_].stream; }

private:
  // A tick's resources: stream, workspace, staging and captured sequences.
  struct Slot {
    void *stream{nullptr};
    void *d_ws{nullptr};
    void *h_stage{nullptr};   // integrate staging (pinned)
    void *h_rstage{nullptr};  // recenter staging (pinned)
    void *d_rarea{nullptr};   // recenter staging, device copy
    // Captured first halves by walk (lanes per ray 1, 4) and near-field
    // steps (the walk's LDS cube), folds + recenter by hot-tile mode: a
    // staged launch holds the kernels and shapes its tick's setting names.
    void *front[2][13]{};
    void *back[2]{nullptr, nullptr};
    void *ev[4]{nullptr, nullptr, nullptr, nullptr};  // start, folds done, end, order mark
    bool dedicated{false};  // its stream has a hardware queue of its own
    bool in_flight{false};
    bool alone{true};     // no other tick overlapped it
    int tune_idx{-1};     // autotuner candidate it ran, or -1
    int lanes{1};
    int near_steps{8};
    int hot_mode{0};
    long long zeroed{0};  // voxels its recenters zero
  };
  void allocSlot(Slot &s, bool dedicated = false);
  void releaseSlot(Slot &s) noexcept;
  void dropGraphs(Slot &s);
  void stage(Slot &s, const std::vector<GridScan> &scans,
             const std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>>
                 &recenters);
  void launch(int q, int prev);
  void finish(int q);
  void release() noexcept;

  BatchLimits lim_;
  size_t ws_bytes_{0};
  int max_regions_{0};
  Slot slots_[2];
  int n_slots_{1};
  int cur_{0};  // slot of the last submitted tick
  bool graph_mode_{false};

  // Recenter outcomes of the staged tick, applied to the grids after it.
  std::vector<OccupancyGrid *> rc_grids_;
  std::vector<Eigen::Vector3i> rc_origins_;

  TuneSetting setting_{};
  std::vector<TuneSetting> tune_;
  std::vector<double> tune_ms_;
  std::vector<int> tune_n_;
  size_t tune_pos_{0};
  int tune_ticks_{0};
  int tune_count_{0};

  double last_update_ms_{0.0};
  double last_recenter_ms_{0.0};
  long long last_zeroed_{0};
};

}  // namespace cublox
