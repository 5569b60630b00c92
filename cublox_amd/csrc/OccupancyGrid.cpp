// Host orchestration of the occupancy grid (reference
// src/OccupancyGrid.cpp:54-521). Compiled by the host compiler; reaches the
// device only through the C ABI of cublox/capi.h.
#include <cublox/OccupancyGrid.hpp>
#include <cublox/capi.h>
#include <cublox/snapshot.h>
#include <cublox/utils.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>

namespace cublox {

namespace {

// Status -> exception, message shaped as the reference's cudaCheck
// (OccupancyGrid.cpp:32-49).
void statusCheck(int status, const char *expr, const char *file, int line) {
  if (status == CUBLOX_OK) return;
  std::string msg = "cublox::cudaCheck: CUDA error in `";
  msg += expr;
  msg += "`: ";
  msg += cublox_last_error();
  msg += " (";
  msg += file;
  msg += ":";
  msg += std::to_string(line);
  msg += ")";
  throw std::runtime_error(msg);
}

#define CUBLOX_CHECK(call) ::cublox::statusCheck((call), #call, __FILE__, __LINE__)

cublox_int3 i3(const Eigen::Vector3i &v) { return cublox_int3{v.x(), v.y(), v.z()}; }
cublox_float3 f3(const Eigen::Vector3f &v) {
  return cublox_float3{v.x(), v.y(), v.z()};
}

template <typename T> void freeDevice(T *&p) noexcept {
  if (p) {
    cublox_free(p);
    p = nullptr;
  }
}

} // namespace

OccupancyGrid::OccupancyGrid(const Eigen::Vector3i &half_map_size_i,
                             const float resolution,
                             const bool origin_at_center,
                             const std::optional<double> recenter_threshold,
                             const Eigen::Vector3f &origin)
    : Grid(half_map_size_i, resolution,
           recenter_threshold ? std::optional<float>(
                                    static_cast<float>(*recenter_threshold))
                              : std::nullopt,
           origin_at_center) {
  Eigen::Vector3i origin_i;
  posToGlobalIndex(origin, config_.resolution_inv, config_.origin_at_center,
                   origin_i);
  updateOriginAndBound(origin_i);
  allocateVoxelBuffers();
}

OccupancyGrid::~OccupancyGrid() { releaseAll(); }

OccupancyGrid::OccupancyGrid(OccupancyGrid &&o) noexcept : Grid(o) {
  *this = std::move(o);
}

OccupancyGrid &OccupancyGrid::operator=(OccupancyGrid &&o) noexcept {
  if (this == &o) return *this;
  releaseAll();
  static_cast<Grid &>(*this) = static_cast<const Grid &>(o);
  first_run_ = o.first_run_;
  stream_ = std::exchange(o.stream_, nullptr);
  zplan_ = std::move(o.zplan_);
  zero_pending_ = std::exchange(o.zero_pending_, false);
  d_op_cnt_ = std::exchange(o.d_op_cnt_, nullptr);
  d_hit_cnt_ = std::exchange(o.d_hit_cnt_, nullptr);
  d_occ_ = std::exchange(o.d_occ_, nullptr);
  d_cloud_x_ = std::exchange(o.d_cloud_x_, nullptr);
  d_cloud_y_ = std::exchange(o.d_cloud_y_, nullptr);
  d_cloud_z_ = std::exchange(o.d_cloud_z_, nullptr);
  cloud_capacity_ = std::exchange(o.cloud_capacity_, 0);
  d_fetch_count_ = std::exchange(o.d_fetch_count_, nullptr);
  d_fetch_pos_ = std::exchange(o.d_fetch_pos_, nullptr);
  d_fetch_state_ = std::exchange(o.d_fetch_state_, nullptr);
  d_fetch_cfgs_ = std::exchange(o.d_fetch_cfgs_, nullptr);
  d_fetch_counts_ = std::exchange(o.d_fetch_counts_, nullptr);
  d_fetch_dense_pos_ = std::exchange(o.d_fetch_dense_pos_, nullptr);
  d_fetch_dense_state_ = std::exchange(o.d_fetch_dense_state_, nullptr);
  d_fetch_offsets_ = std::exchange(o.d_fetch_offsets_, nullptr);
  h_upload_ = std::exchange(o.h_upload_, nullptr);
  h_upload_dev_ = std::exchange(o.h_upload_dev_, nullptr);
  stage_in_use_ = std::exchange(o.stage_in_use_, false);
  h_upload_bytes_ = std::exchange(o.h_upload_bytes_, 0);
  h_stage_ = std::exchange(o.h_stage_, nullptr);
  h_stage_bytes_ = std::exchange(o.h_stage_bytes_, 0);
  d_viz_ = std::exchange(o.d_viz_, nullptr);
  d_viz_count_ = std::exchange(o.d_viz_count_, nullptr);
  viz_capacity_ = std::exchange(o.viz_capacity_, 0);
  d_kept_ = std::exchange(o.d_kept_, nullptr);
  d_delta_count_ = std::exchange(o.d_delta_count_, nullptr);
  d_delta_idx_ = std::exchange(o.d_delta_idx_, nullptr);
  d_delta_val_ = std::exchange(o.d_delta_val_, nullptr);
  delta_capacity_ = std::exchange(o.delta_capacity_, 0u);
  fetch_dense_capacity_ = std::exchange(o.fetch_dense_capacity_, 0);
  fetch_offsets_capacity_ = std::exchange(o.fetch_offsets_capacity_, 0);
  fetch_capacity_ = std::exchange(o.fetch_capacity_, 0);
  fetch_cfg_capacity_ = std::exchange(o.fetch_cfg_capacity_, 0);
  timing_ = std::exchange(o.timing_, false);
  for (int k = 0; k < 8; ++k) events_[k] = std::exchange(o.events_[k], nullptr);
  for (int k = 0; k < 3; ++k) pending_[k] = std::exchange(o.pending_[k], false);
  pending_zeroed_ = std::exchange(o.pending_zeroed_, 0);
  times_ = o.times_;
  update_path_ = o.update_path_;
  near_steps_ = o.near_steps_;
  d_ws_ = std::exchange(o.d_ws_, nullptr);
  ws_bytes_ = std::exchange(o.ws_bytes_, 0);
  ws_points_ = std::exchange(o.ws_points_, 0);
  ws_range_ = o.ws_range_;
  ws_near_ = o.ws_near_;
  max_raycast_range_ = o.max_raycast_range_;
  l_hit_ = o.l_hit_;
  l_miss_ = o.l_miss_;
  l_min_ = o.l_min_;
  l_max_ = o.l_max_;
  l_free_ = o.l_free_;
  l_occupied_ = o.l_occupied_;
  l_unknown_ = o.l_unknown_;
  return *this;
}

void OccupancyGrid::releaseAll() noexcept {
  if (stream_) cublox_stream_synchronize(stream_);
  stage_in_use_ = false;
  zero_pending_ = false;  // the map goes: its pending clear with it
  freeDevice(d_fetch_state_);
  freeDevice(d_fetch_pos_);
  freeDevice(d_fetch_count_);
  freeDevice(d_fetch_cfgs_);
  freeDevice(d_fetch_counts_);
  freeDevice(d_fetch_dense_pos_);
  freeDevice(d_fetch_dense_state_);
  freeDevice(d_fetch_offsets_);
  fetch_dense_capacity_ = 0;
  fetch_offsets_capacity_ = 0;
  freeDevice(d_delta_count_);
  freeDevice(d_delta_idx_);
  freeDevice(d_delta_val_);
  delta_capacity_ = 0;
  if (h_stage_) cublox_host_free(h_stage_);
  h_stage_ = nullptr;
  h_stage_bytes_ = 0;
  if (h_upload_) cublox_host_free(h_upload_);
  h_upload_ = nullptr;
  h_upload_dev_ = nullptr;
  h_upload_bytes_ = 0;
  freeDevice(d_viz_);
  freeDevice(d_viz_count_);
  viz_capacity_ = 0;
  freeDevice(d_kept_);
  freeDevice(d_op_cnt_);
  freeDevice(d_hit_cnt_);
  freeDevice(d_occ_);
  freeDevice(d_cloud_x_);
  d_cloud_y_ = d_cloud_z_ = nullptr;
  freeDevice(d_ws_);
  ws_bytes_ = 0;
  ws_points_ = 0;
  cloud_capacity_ = 0;
  fetch_capacity_ = 0;
  fetch_cfg_capacity_ = 0;
  for (auto &e : events_) {
    if (e) {
      cublox_event_destroy(e);
      e = nullptr;
    }
  }
  timing_ = false;
  if (stream_) {
    cublox_stream_destroy(stream_);
    stream_ = nullptr;
  }
}

void OccupancyGrid::enableKernelTiming(bool on) {
  if (on && !events_[0]) {
    for (auto &e : events_) CUBLOX_CHECK(cublox_event_create(&e));
  }
  timing_ = on;
}

void OccupancyGrid::mark(int ev) const {
  if (timing_) CUBLOX_CHECK(cublox_event_record(events_[ev], stream_));
}

void OccupancyGrid::join() const {
  if (!zero_pending_) return;
  CUBLOX_CHECK(cublox_launch_recenter_plan(zplan_.data(), d_occ_, d_op_cnt_, d_hit_cnt_,
                                           stream_));
  zero_pending_ = false;
}

void OccupancyGrid::settle() const {
  if (!zero_pending_) return;
  join();
  sync();
}

OccupancyGrid::KernelTimes OccupancyGrid::kernelTimes() const {
  if (timing_) {
    sync();
    harvestTimes();
  }
  return times_;
}

// Called once the events' work is done (after a stream sync; a recenter's
// events are waited for): fold the completed event pairs into times_.
void OccupancyGrid::harvestTimes() const {
  if (!timing_) return;
  if (pending_[2]) CUBLOX_CHECK(cublox_event_synchronize(events_[5]));
  double *acc[3] = {&times_.update_ms, &times_.apply_ms, &times_.recenter_ms};
  long long *cnt[3] = {&times_.update_launches, &times_.apply_launches,
                       &times_.recenter_launches};
  for (int k = 0; k < 3; ++k) {
    if (!pending_[k]) continue;
    float ms = 0.0f;
    CUBLOX_CHECK(cublox_event_elapsed_ms(events_[2 * k], events_[2 * k + 1], &ms));
    *acc[k] += ms;
    *cnt[k] += 1;
    pending_[k] = false;
  }
  times_.recenter_zeroed_voxels += pending_zeroed_;
  pending_zeroed_ = 0;
}

void OccupancyGrid::sync() const {
  CUBLOX_CHECK(cublox_stream_synchronize(stream_));
  stage_in_use_ = false;
}

// OccupancyGrid.cpp:431-480: the occupancy array of V, zeroed; partial
// allocations are rolled back before throwing. The per-voxel op/hit counters
// of the reference are only needed by the per-voxel-atomics path and the
// stage-wise checks, so they are allocated on first use (ensureCounters).
void OccupancyGrid::allocateVoxelBuffers() {
  const size_t n = static_cast<size_t>(config_.voxel_num);
  if (n == 0) return;
  try {
    CUBLOX_CHECK(cublox_stream_create(&stream_));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_occ_), n * sizeof(float)));
    CUBLOX_CHECK(cublox_fill_zero_async(d_occ_, n, stream_));
    sync();
  } catch (...) {
    releaseAll();
    throw std::runtime_error(
        "cublox::OccupancyGrid: failed to allocate voxel buffers");
  }
}

void OccupancyGrid::reset() {
  join();
  if (d_occ_ == nullptr) return;
  const size_t n = static_cast<size_t>(config_.voxel_num);
  CUBLOX_CHECK(cublox_fill_zero_async(d_occ_, n, stream_));
  if (d_op_cnt_) CUBLOX_CHECK(cublox_fill_zero_async(d_op_cnt_, n, stream_));
  if (d_hit_cnt_) CUBLOX_CHECK(cublox_fill_zero_async(d_hit_cnt_, n, stream_));
  sync();
  first_run_ = true;
}

void OccupancyGrid::resetVoxel(const int hash_id) {
  join();
  if (d_occ_ == nullptr || hash_id < 0 || hash_id >= config_.voxel_num) return;
  CUBLOX_CHECK(cublox_memset_async(d_occ_ + hash_id, 0, sizeof(float), stream_));
  if (d_op_cnt_) {
    CUBLOX_CHECK(cublox_memset_async(d_op_cnt_ + hash_id, 0, sizeof(int), stream_));
    CUBLOX_CHECK(cublox_memset_async(d_hit_cnt_ + hash_id, 0, sizeof(int), stream_));
  }
  sync();
}

// OccupancyGrid.cpp:141-179. The plan runs on the host (integer math on the
// config), so there is no device round trip. When the per-voxel counters
// exist they are cleared with occ, as the reference's launchRecenter does:
// the stage-wise calls (castRays* then applyCounters) can leave counts
// pending across a recenter, and those in the cleared slabs must go too.
void OccupancyGrid::recenter(const Eigen::Vector3f &pos) {
  if (!config_.recenter_threshold.has_value() || d_occ_ == nullptr) return;
  if ((pos - origin_f_).squaredNorm() < config_.recenter_threshold_squared) return;

  cublox_recenter_cfg cfg{};
  cfg.pos = f3(pos);
  cfg.origin_f = f3(origin_f_);
  cfg.origin_i = i3(origin_i_);
  cfg.map_size_i = i3(config_.map_size_i);
  cfg.half_map_size_i = i3(config_.half_map_size_i);
  cfg.resolution = config_.resolution;
  cfg.inv_resolution = config_.resolution_inv;
  cfg.recenter_threshold_sq = config_.recenter_threshold_squared;
  cfg.origin_at_center = config_.origin_at_center ? 1 : 0;
  cfg.map_vox_num = config_.voxel_num;

  // The plan is host arithmetic; only an action needs device work. An
  // earlier clear still pending runs first (clears follow each other).
  join();
  if (zplan_.empty()) zplan_.resize(cublox_recenter_plan_bytes());
  cublox_recenter_result result{};
  CUBLOX_CHECK(cublox_recenter_plan_stage(&cfg, zplan_.data(), &result));
  if (result.action == 0) return;
  if (timing_) {
    // Timed: the clear runs now, on its own, between the recenter events.
    if (pending_[2]) {  // the previous clear's times first
      CUBLOX_CHECK(cublox_event_synchronize(events_[5]));
      harvestTimes();
    }
    CUBLOX_CHECK(cublox_event_record(events_[4], stream_));
    CUBLOX_CHECK(cublox_launch_recenter_plan(zplan_.data(), d_occ_, d_op_cnt_, d_hit_cnt_,
                                             stream_));
    CUBLOX_CHECK(cublox_event_record(events_[5], stream_));
    pending_[2] = true;
    pending_zeroed_ += result.zeroed_voxels;
  } else {
    zero_pending_ = true;  // the next update's walk launch, or join()
  }
  updateOriginAndBound(Eigen::Vector3i(result.new_origin_i.x,
                                       result.new_origin_i.y,
                                       result.new_origin_i.z));
}

// OccupancyGrid.cpp:192-198: the first update pins the window origin to the
// sensor voxel.
void OccupancyGrid::pinFirstOrigin(const Eigen::Vector3f &sensor_origin) {
  if (!first_run_) return;
  Eigen::Vector3i origin_i;
  posToGlobalIndex(sensor_origin, config_.resolution_inv,
                   config_.origin_at_center, origin_i);
  updateOriginAndBound(origin_i);
  first_run_ = false;
}

void OccupancyGrid::ensureCounters() {
  if (d_op_cnt_) return;
  const size_t n = static_cast<size_t>(config_.voxel_num);
  try {
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_op_cnt_), n * sizeof(int)));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_hit_cnt_), n * sizeof(int)));
    CUBLOX_CHECK(cublox_fill_zero_async(d_op_cnt_, n, stream_));
    CUBLOX_CHECK(cublox_fill_zero_async(d_hit_cnt_, n, stream_));
  } catch (...) {
    freeDevice(d_op_cnt_);
    freeDevice(d_hit_cnt_);
    throw;
  }
}

// The per-voxel apply keeps the reference's gate (kernels.cu:251-262): it
// records changed voxels only when the capacity covers every voxel.
static void checkDeltaGate(OccupancyGrid::UpdatePath p, unsigned int capacity,
                           int voxel_num) {
  if (p == OccupancyGrid::UpdatePath::PER_VOXEL_ATOMICS && capacity > 0 &&
      capacity < static_cast<unsigned int>(voxel_num)) {
    throw std::invalid_argument(
        "cublox::OccupancyGrid: the per-voxel apply records changed voxels only "
        "with a capacity of at least the voxel count");
  }
}

void OccupancyGrid::setUpdatePath(UpdatePath p) {
  checkDeltaGate(p, delta_capacity_, config_.voxel_num);
  update_path_ = p;
}

void OccupancyGrid::setChangedVoxelCapacity(unsigned int capacity) {
  checkDeltaGate(update_path_, capacity, config_.voxel_num);
  if (capacity == delta_capacity_) return;
  sync();
  freeDevice(d_delta_count_);
  freeDevice(d_delta_idx_);
  freeDevice(d_delta_val_);
  delta_capacity_ = 0;
  if (capacity == 0) return;
  try {
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_delta_count_), sizeof(unsigned int)));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_delta_idx_),
                               static_cast<size_t>(capacity) * sizeof(int)));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_delta_val_),
                               static_cast<size_t>(capacity) * sizeof(float)));
    CUBLOX_CHECK(cublox_fill_zero_async(d_delta_count_, 1, stream_));
  } catch (...) {
    freeDevice(d_delta_count_);
    freeDevice(d_delta_idx_);
    freeDevice(d_delta_val_);
    throw;
  }
  delta_capacity_ = capacity;
}

unsigned long long OccupancyGrid::changedVoxels(std::vector<ChangedVoxel> &out) const {
  out.clear();
  if (delta_capacity_ == 0) return 0;
  unsigned int count = 0;
  CUBLOX_CHECK(cublox_memcpy_d2h_async(&count, d_delta_count_, sizeof(count), stream_));
  sync();
  const unsigned int n = std::min(count, delta_capacity_);
  std::vector<int> ids(n);
  std::vector<float> vals(n);
  if (n > 0) {
    void *const dst[2] = {ids.data(), vals.data()};
    const void *const src[2] = {d_delta_idx_, d_delta_val_};
    const size_t len[2] = {n * sizeof(int), n * sizeof(float)};
    downloadHost(dst, src, len, 2);
  }
  out.resize(n);
  for (unsigned int k = 0; k < n; ++k) out[k] = ChangedVoxel{ids[k], vals[k]};
  return count;
}

void OccupancyGrid::setNearSteps(int k) {
  if (k < 1 || k > 12) {
    throw std::invalid_argument("cublox::OccupancyGrid: near steps must be in [1, 12]");
  }
  near_steps_ = k;
}

void OccupancyGrid::ensureWorkspace(int n) {
  cublox_raycast_cfg cfg{};
  cfg.map_size_i = i3(config_.map_size_i);
  cfg.half_map_size_i = i3(config_.half_map_size_i);
  cfg.resolution = config_.resolution;
  cfg.inv_resolution = config_.resolution_inv;
  cfg.max_range = max_raycast_range_;
  cfg.map_vox_num = config_.voxel_num;
  if (d_ws_ && n <= ws_points_ && ws_range_ == max_raycast_range_ &&
      ws_near_ == near_steps_) {
    return;
  }
  const int pts = std::max(n, ws_points_);
  const size_t bytes = cublox_integrate_workspace_bytes(&cfg, pts, near_steps_);
  sync();
  freeDevice(d_ws_);
  ws_bytes_ = 0;
  ws_points_ = 0;
  CUBLOX_CHECK(cublox_malloc(&d_ws_, bytes));
  CUBLOX_CHECK(cublox_integrate_workspace_init(d_ws_, bytes, stream_));
  ws_bytes_ = bytes;
  ws_points_ = pts;
  ws_range_ = max_raycast_range_;
  ws_near_ = near_steps_;
}

// Fused ray cast + fold (or count export) on the grid's stream.
void OccupancyGrid::launchIntegrate(const float *d_xs, const float *d_ys,
                                    const float *d_zs, int n, const float *T,
                                    const Eigen::Vector3f &origin,
                                    bool export_counts, int stride,
                                    float drop_sq) {
  ensureWorkspace(n);
  if (export_counts) ensureCounters();
  cublox_raycast_cfg cfg{};
  cfg.origin = f3(origin);
  cfg.origin_i = i3(origin_i_);
  cfg.map_size_i = i3(config_.map_size_i);
  cfg.half_map_size_i = i3(config_.half_map_size_i);
  cfg.resolution = config_.resolution;
  cfg.inv_resolution = config_.resolution_inv;
  cfg.max_range = max_raycast_range_;
  cfg.map_vox_num = config_.voxel_num;
  cfg.origin_at_center = config_.origin_at_center ? 1 : 0;
  const bool delta = !export_counts && delta_capacity_ > 0;
  cublox_integrate_grid q{};
  q.cfg = cfg;
  if (T) {
    for (int k = 0; k < 12; ++k) q.T[k] = T[k];
    q.use_T = 1;
  }
  q.d_xs = d_xs;
  q.d_ys = d_ys;
  q.d_zs = d_zs;
  q.n = n;
  q.d_occ = d_occ_;
  q.l_hit = l_hit_;
  q.l_miss = l_miss_;
  q.l_min = l_min_;
  q.l_max = l_max_;
  q.l_unknown = l_unknown_;
  q.near_steps = near_steps_;
  if (export_counts) {
    q.d_export_op = d_op_cnt_;
    q.d_export_hit = d_hit_cnt_;
  }
  if (delta) {
    q.d_delta_count = d_delta_count_;
    q.d_delta_idx = d_delta_idx_;
    q.d_delta_val = d_delta_val_;
    q.delta_capacity = delta_capacity_;
  }
  q.point_stride = stride;
  q.range_drop_sq = drop_sq;
  // A pending recenter clear rides in the walk launch (the folds follow).
  q.zero_plan = zero_pending_ ? zplan_.data() : nullptr;
  q.d_zero_op = d_op_cnt_;
  q.d_zero_hit = d_hit_cnt_;
  mark(0);
  CUBLOX_CHECK(cublox_launch_integrate_grid(&q, d_ws_, ws_bytes_, ws_points_, stream_));
  mark(1);
  zero_pending_ = false;
  pending_[0] = timing_;
}

void OccupancyGrid::castRaysWorld(const float *d_x, const float *d_y,
                                  const float *d_z, int n,
                                  const Eigen::Vector3f &sensor_origin) {
  join();
  ensureCounters();
  cublox_raycast_cfg cfg{};
  cfg.origin = f3(sensor_origin);
  cfg.origin_i = i3(origin_i_);
  cfg.map_size_i = i3(config_.map_size_i);
  cfg.half_map_size_i = i3(config_.half_map_size_i);
  cfg.resolution = config_.resolution;
  cfg.inv_resolution = config_.resolution_inv;
  cfg.max_range = max_raycast_range_;
  cfg.map_vox_num = config_.voxel_num;
  cfg.origin_at_center = config_.origin_at_center ? 1 : 0;
  mark(0);
  CUBLOX_CHECK(cublox_launch_ray_cast_update(&cfg, d_x, d_y, d_z, n, stream_,
                                             d_op_cnt_, d_hit_cnt_));
  mark(1);
  pending_[0] = timing_;
}

void OccupancyGrid::castRaysSensor(const float *d_xs, const float *d_ys,
                                   const float *d_zs, int n,
                                   const float T[12]) {
  join();
  ensureCounters();
  cublox_raycast_cfg cfg{};
  cfg.origin = cublox_float3{T[3], T[7], T[11]};
  cfg.origin_i = i3(origin_i_);
  cfg.map_size_i = i3(config_.map_size_i);
  cfg.half_map_size_i = i3(config_.half_map_size_i);
  cfg.resolution = config_.resolution;
  cfg.inv_resolution = config_.resolution_inv;
  cfg.max_range = max_raycast_range_;
  cfg.map_vox_num = config_.voxel_num;
  cfg.origin_at_center = config_.origin_at_center ? 1 : 0;
  mark(0);
  CUBLOX_CHECK(cublox_launch_ray_cast_update_sensor(&cfg, T, d_xs, d_ys, d_zs, n,
                                                    stream_, d_op_cnt_,
                                                    d_hit_cnt_));
  mark(1);
  pending_[0] = timing_;
}

void OccupancyGrid::launchApply() {
  join();
  ensureCounters();
  checkDeltaGate(UpdatePath::PER_VOXEL_ATOMICS, delta_capacity_, config_.voxel_num);
  const bool delta = delta_capacity_ > 0;
  if (delta) CUBLOX_CHECK(cublox_fill_zero_async(d_delta_count_, 1, stream_));
  mark(2);
  CUBLOX_CHECK(cublox_launch_apply_update(
      d_occ_, d_op_cnt_, d_hit_cnt_, config_.voxel_num, l_hit_, l_miss_, l_min_,
      l_max_, l_unknown_, stream_, delta ? d_delta_count_ : nullptr,
      delta ? d_delta_idx_ : nullptr, delta ? d_delta_val_ : nullptr,
      delta ? delta_capacity_ : 0u));
  mark(3);
  pending_[1] = timing_;
}

// Host bytes to device memory on the grid's stream. A pageable
// hipMemcpyAsync of a scan's columns runs at a few GB/s; instead the parts
// are copied into a grid-owned pinned buffer, and each goes out by DMA
// while the next is copied. The buffer is reused by the next call, so the
// caller syncs the stream before returning (every public update blocks).
void OccupancyGrid::uploadHost(void *d_dst, const void *const *h_srcs,
                              const size_t *lens, int parts) {
  size_t bytes = 0;
  for (int k = 0; k < parts; ++k) bytes += lens[k];
  if (bytes == 0) return;
  if (stage_in_use_) sync();  // kernels of a call that threw may read it
  stage_in_use_ = true;
  if (bytes > h_upload_bytes_) {
    sync();  // an earlier upload may still read the old buffer
    stage_in_use_ = true;
    if (h_upload_) cublox_host_free(h_upload_);
    h_upload_ = nullptr;
    h_upload_dev_ = nullptr;
    h_upload_bytes_ = 0;
    const size_t cap = std::max<size_t>(bytes, size_t{1} << 20);
    CUBLOX_CHECK(cublox_host_alloc(&h_upload_, cap));
    h_upload_bytes_ = cap;
  }
  // One transfer per part: each goes out while the next part is copied in
  // (a transfer's launch costs several microseconds, so parts are not cut
  // finer).
  auto *stage = static_cast<uint8_t *>(h_upload_);
  auto *dst = static_cast<uint8_t *>(d_dst);
  size_t base = 0;
  for (int k = 0; k < parts; ++k) {
    std::memcpy(stage + base, h_srcs[k], lens[k]);
    CUBLOX_CHECK(cublox_memcpy_h2d_async(dst + base, stage + base, lens[k], stream_));
    base += lens[k];
  }
}

// Host columns staged in the grid's pinned buffer for the fused path,
// which reads them from there over the bus (each point is read by one
// walk workgroup): no DMA to HBM before the walk. *dx, *dy, *dz receive
// the columns' device addresses. The buffer is reused by the next call,
// and every public update syncs its stream before returning.
void OccupancyGrid::stageColumns(const float *x, const float *y, const float *z, int n,
                                 const float **dx, const float **dy, const float **dz) {
  const size_t b = static_cast<size_t>(n) * sizeof(float);
  const size_t col = (b + 255) & ~static_cast<size_t>(255);
  uint8_t *stage = readStage(3 * col);
  std::memcpy(stage, x, b);
  std::memcpy(stage + col, y, b);
  std::memcpy(stage + 2 * col, z, b);
  const auto *dev = static_cast<const uint8_t *>(h_upload_dev_);
  *dx = reinterpret_cast<const float *>(dev);
  *dy = reinterpret_cast<const float *>(dev + col);
  *dz = reinterpret_cast<const float *>(dev + 2 * col);
}

// The pinned stage with room for `bytes`, and its device address set; the
// caller's kernels will read it (until the next sync). A call that threw
// after handing the stage to kernels leaves it marked: then wait for them.
uint8_t *OccupancyGrid::readStage(size_t bytes) {
  if (stage_in_use_) sync();
  if (bytes > h_upload_bytes_) {
    sync();  // an earlier call's kernels may still read the old buffer
    if (h_upload_) cublox_host_free(h_upload_);
    h_upload_ = nullptr;
    h_upload_dev_ = nullptr;
    h_upload_bytes_ = 0;
    const size_t cap = std::max<size_t>(bytes, size_t{1} << 20);
    CUBLOX_CHECK(cublox_host_alloc(&h_upload_, cap));
    h_upload_bytes_ = cap;
  }
  if (!h_upload_dev_) CUBLOX_CHECK(cublox_host_device_pointer(&h_upload_dev_, h_upload_));
  stage_in_use_ = true;
  return static_cast<uint8_t *>(h_upload_);
}

// A host payload (PointCloud2 records, a depth image) copied into the
// pinned stage; returns the copy's device address, which the ingest
// kernels read over the bus (no DMA to HBM first).
const void *OccupancyGrid::stageHost(const void *src, size_t bytes) {
  std::memcpy(readStage(bytes), src, bytes);
  return h_upload_dev_;
}

void OccupancyGrid::uploadColumns(const float *x, const float *y, const float *z,
                                  int n) {
  const size_t b = static_cast<size_t>(n) * sizeof(float);
  // d_cloud_y_ and d_cloud_z_ follow d_cloud_x_ at n floats each.
  const void *srcs[3] = {x, y, z};
  const size_t lens[3] = {b, b, b};
  uploadHost(d_cloud_x_, srcs, lens, 3);
}

// One device buffer holds the three columns of the current cloud back to
// back (x, then y at +n, then z at +2n), so one pinned stage uploads them.
// d_cloud_y_ / d_cloud_z_ are set per call by placeCloud.
void OccupancyGrid::ensureCloudCapacity(const int n) {
  if (n > cloud_capacity_) {
    sync();
    freeDevice(d_cloud_x_);
    d_cloud_y_ = d_cloud_z_ = nullptr;
    cloud_capacity_ = 0;
    const size_t bytes = 3 * static_cast<size_t>(n) * sizeof(float);
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_cloud_x_), bytes));
    cloud_capacity_ = n;
  }
  d_cloud_y_ = d_cloud_x_ + n;
  d_cloud_z_ = d_cloud_x_ + 2 * static_cast<size_t>(n);
}

// OccupancyGrid.cpp:181-238: cast, then fold, then one sync (the reference
// syncs after each pass; stream order makes the first unnecessary).
void OccupancyGrid::update(const PointCloud &cloud,
                           const Eigen::Vector3f &sensor_origin) {
  // PointCloud is column-major: each column is a contiguous SoA array.
  updateColumns(cloud.col(0).data(), cloud.col(1).data(), cloud.col(2).data(),
                static_cast<int>(cloud.rows()), sensor_origin);
}

void OccupancyGrid::updateColumns(const float *x, const float *y,
                                  const float *z, int n,
                                  const Eigen::Vector3f &sensor_origin) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(sensor_origin);
  if (update_path_ == UpdatePath::FUSED_TILES) {
    const float *dx, *dy, *dz;
    stageColumns(x, y, z, n, &dx, &dy, &dz);
    launchIntegrate(dx, dy, dz, n, nullptr, sensor_origin, false);
  } else {
    ensureCloudCapacity(n);
    uploadColumns(x, y, z, n);
    castRaysWorld(d_cloud_x_, d_cloud_y_, d_cloud_z_, n, sensor_origin);
    launchApply();
  }
  sync();
  harvestTimes();
}

void OccupancyGrid::updateSensorFrame(const PointCloud &cloud,
                                      const Eigen::Matrix3f &R,
                                      const Eigen::Vector3f &t) {
  float T[12];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[4 * r + c] = R(r, c);
    T[4 * r + 3] = t(r);
  }
  updateSensorFrameColumns(cloud.col(0).data(), cloud.col(1).data(),
                           cloud.col(2).data(), static_cast<int>(cloud.rows()),
                           T);
}

void OccupancyGrid::updateSensorFrameColumns(const float *xs, const float *ys,
                                             const float *zs, int n,
                                             const float T[12]) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(Eigen::Vector3f(T[3], T[7], T[11]));
  if (update_path_ == UpdatePath::FUSED_TILES) {
    const float *dx, *dy, *dz;
    stageColumns(xs, ys, zs, n, &dx, &dy, &dz);
    launchIntegrate(dx, dy, dz, n, T, Eigen::Vector3f(T[3], T[7], T[11]), false);
  } else {
    ensureCloudCapacity(n);
    uploadColumns(xs, ys, zs, n);
    castRaysSensor(d_cloud_x_, d_cloud_y_, d_cloud_z_, n, T);
    launchApply();
  }
  sync();
  harvestTimes();
}

void OccupancyGrid::updateSensorFrameDevice(const float *d_xs,
                                            const float *d_ys,
                                            const float *d_zs, int n,
                                            const float T[12]) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(Eigen::Vector3f(T[3], T[7], T[11]));
  if (update_path_ == UpdatePath::FUSED_TILES) {
    launchIntegrate(d_xs, d_ys, d_zs, n, T, Eigen::Vector3f(T[3], T[7], T[11]),
                    false);
  } else {
    castRaysSensor(d_xs, d_ys, d_zs, n, T);
    launchApply();
  }
  sync();
  harvestTimes();
}

// ---- driver ingest ------------------------------------------------------

const float *OccupancyGrid::rawFloats(const void *d_data, uint32_t off) const {
  return reinterpret_cast<const float *>(static_cast<const unsigned char *>(d_data) + off);
}

// buildCloudInOdomFrame (main.cpp:640-693) fused into the update: the
// driver's filter runs in the ray set-up; the census of kept points decides
// the return value (and, on the first run, whether the origin is pinned).
bool OccupancyGrid::ingest(const float *d_xs, const float *d_ys,
                           const float *d_zs, int stride, int n,
                           const float T[12]) {
  if (update_path_ != UpdatePath::FUSED_TILES) {
    throw std::logic_error("cublox::OccupancyGrid: sensor ingest runs on the fused path");
  }
  if (config_.voxel_num <= 0 || n <= 0) return false;
  if (!d_kept_) {
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_kept_), sizeof(unsigned int)));
  }
  // main.cpp:451: the range filter is the grid's max range, squared in float.
  const float drop_sq = max_raycast_range_ * max_raycast_range_;
  CUBLOX_CHECK(cublox_launch_count_kept_points(d_xs, d_ys, d_zs, stride, n, drop_sq,
                                               d_kept_, stream_));
  unsigned int kept = 0;
  if (first_run_) {
    CUBLOX_CHECK(cublox_memcpy_d2h_async(&kept, d_kept_, sizeof(kept), stream_));
    sync();
    if (kept == 0) return false;
  }
  const Eigen::Vector3f t(T[3], T[7], T[11]);
  pinFirstOrigin(t);
  launchIntegrate(d_xs, d_ys, d_zs, n, T, t, false, stride, drop_sq);
  CUBLOX_CHECK(cublox_memcpy_d2h_async(&kept, d_kept_, sizeof(kept), stream_));
  sync();
  harvestTimes();
  return kept > 0;
}

static void checkLayout(const OccupancyGrid::PointCloud2Layout &l) {
  if (l.point_step == 0 || l.point_step % 4 != 0 || l.off_x % 4 != 0 ||
      l.off_y % 4 != 0 || l.off_z % 4 != 0 || l.off_x + 4 > l.point_step ||
      l.off_y + 4 > l.point_step || l.off_z + 4 > l.point_step) {
    throw std::invalid_argument(
        "cublox::OccupancyGrid: PointCloud2 layout needs a point step and "
        "float32 x/y/z offsets that are multiples of 4 inside the record");
  }
}

static int cappedPoints(uint32_t width, uint32_t height) {
  const unsigned long long n = static_cast<unsigned long long>(width) * height;
  return static_cast<int>(std::min<unsigned long long>(
      n, static_cast<unsigned long long>(OccupancyGrid::kMaxCloudPoints)));
}

// The payload must hold the records read: the first min(width * height,
// kMaxCloudPoints) records of point_step bytes each.
static void checkPayload(const OccupancyGrid::PointCloud2Layout &l, int n,
                         size_t data_bytes) {
  const unsigned long long need = static_cast<unsigned long long>(n) * l.point_step;
  if (data_bytes < need) {
    throw std::invalid_argument(
        "cublox::OccupancyGrid: PointCloud2 payload of " + std::to_string(data_bytes) +
        " bytes is shorter than the " + std::to_string(need) +
        " bytes its width, height and point step describe");
  }
}

bool OccupancyGrid::updatePointCloud2Device(const void *d_data,
                                            const PointCloud2Layout &layout,
                                            const float T[12],
                                            size_t data_bytes) {
  checkLayout(layout);
  const int n = cappedPoints(layout.width, layout.height);
  checkPayload(layout, n, data_bytes);
  if (n == 0) return false;
  return ingest(rawFloats(d_data, layout.off_x), rawFloats(d_data, layout.off_y),
                rawFloats(d_data, layout.off_z),
                static_cast<int>(layout.point_step / 4), n, T);
}

bool OccupancyGrid::updatePointCloud2(const void *data,
                                      const PointCloud2Layout &layout,
                                      const float T[12], size_t data_bytes) {
  checkLayout(layout);
  const int n = cappedPoints(layout.width, layout.height);
  checkPayload(layout, n, data_bytes);
  if (n == 0) return false;
  const size_t bytes = static_cast<size_t>(n) * layout.point_step;
  return updatePointCloud2Device(stageHost(data, bytes), layout, T, bytes);
}

bool OccupancyGrid::updateDepthImageDevice(const void *d_data, uint32_t width,
                                           uint32_t height, uint32_t step,
                                           DepthEncoding encoding,
                                           const float K[4], const float T[12]) {
  const int n = cappedPoints(width, height);
  if (n == 0) return false;
  ensureCloudCapacity(n);
  CUBLOX_CHECK(cublox_launch_depth_to_points(
      d_data, static_cast<int>(width), static_cast<int>(height), step,
      encoding == DepthEncoding::MILLIMETRES_U16 ? 0 : 1, K, n, d_cloud_x_,
      d_cloud_y_, d_cloud_z_, stream_));
  return ingest(d_cloud_x_, d_cloud_y_, d_cloud_z_, 1, n, T);
}

bool OccupancyGrid::updateDepthImage(const void *data, uint32_t width,
                                     uint32_t height, uint32_t step,
                                     DepthEncoding encoding, const float K[4],
                                     const float T[12]) {
  if (width == 0 || height == 0) return false;
  const size_t bytes = static_cast<size_t>(step) * height;
  return updateDepthImageDevice(stageHost(data, bytes), width, height, step, encoding, K, T);
}

void OccupancyGrid::updateDevice(const float *d_x, const float *d_y,
                                 const float *d_z, int n,
                                 const Eigen::Vector3f &sensor_origin) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(sensor_origin);
  if (update_path_ == UpdatePath::FUSED_TILES) {
    launchIntegrate(d_x, d_y, d_z, n, nullptr, sensor_origin, false);
  } else {
    castRaysWorld(d_x, d_y, d_z, n, sensor_origin);
    launchApply();
  }
  sync();
  harvestTimes();
}

void OccupancyGrid::castRaysDevice(const float *d_x, const float *d_y,
                                   const float *d_z, int n,
                                   const Eigen::Vector3f &sensor_origin) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(sensor_origin);
  castRaysWorld(d_x, d_y, d_z, n, sensor_origin);
  sync();
  harvestTimes();
}

void OccupancyGrid::castRaysSensorFrameDevice(const float *d_xs,
                                              const float *d_ys,
                                              const float *d_zs, int n,
                                              const float T[12]) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(Eigen::Vector3f(T[3], T[7], T[11]));
  castRaysSensor(d_xs, d_ys, d_zs, n, T);
  sync();
  harvestTimes();
}

void OccupancyGrid::castRaysFusedSensorFrameDevice(const float *d_xs,
                                                   const float *d_ys,
                                                   const float *d_zs, int n,
                                                   const float T[12]) {
  if (config_.voxel_num <= 0 || n <= 0) return;
  pinFirstOrigin(Eigen::Vector3f(T[3], T[7], T[11]));
  launchIntegrate(d_xs, d_ys, d_zs, n, T, Eigen::Vector3f(T[3], T[7], T[11]),
                  true);
  sync();
  harvestTimes();
}

void OccupancyGrid::counterCensus(unsigned long long out[4]) const {
  join();
  for (int k = 0; k < 4; ++k) out[k] = 0;
  if (!d_occ_ || !d_op_cnt_) return;
  unsigned long long *d = nullptr;
  CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d), 4 * sizeof(unsigned long long)));
  int st = cublox_memset_async(d, 0, 4 * sizeof(unsigned long long), stream_);
  if (st == CUBLOX_OK)
    st = cublox_counter_census(d_op_cnt_, d_hit_cnt_, config_.voxel_num, d, stream_);
  if (st == CUBLOX_OK)
    st = cublox_memcpy_d2h_async(out, d, 4 * sizeof(unsigned long long), stream_);
  if (st == CUBLOX_OK) st = cublox_stream_synchronize(stream_);
  cublox_free(d);
  CUBLOX_CHECK(st);
}

void OccupancyGrid::applyCounters() {
  if (config_.voxel_num <= 0 || !d_op_cnt_) return;
  launchApply();
  sync();
  harvestTimes();
}

void OccupancyGrid::ensureFetchCapacity(const long long n) const {
  if (n <= fetch_capacity_) return;
  sync();
  freeDevice(d_fetch_state_);
  freeDevice(d_fetch_pos_);
  fetch_capacity_ = 0;
  if (d_fetch_count_ == nullptr) {
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_count_),
                               sizeof(unsigned int)));
  }
  const size_t cap = static_cast<size_t>(n);
  CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_pos_),
                             cap * 3 * sizeof(float)));
  CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_state_), cap));
  fetch_capacity_ = n;
}

// OccupancyGrid.cpp:240-356
std::vector<OccupancyGrid::OccupancySample> OccupancyGrid::fetchOccupancyAround(
    const Eigen::Vector3f &center, const float radius,
    const std::optional<VoxelState> state_filter, const int max_results) const {
  join();
  std::vector<OccupancySample> out;
  if (d_occ_ == nullptr || config_.voxel_num <= 0 || radius <= 0.0f) return out;

  cublox_fetch_cfg cfg{};
  long long est = 0;
  unsigned int output_cap = 0;
  const int filter = state_filter ? static_cast<int>(*state_filter) : -1;
  if (!cublox_fetch_plan(i3(origin_i_), i3(config_.map_size_i),
                         i3(config_.half_map_size_i), config_.resolution,
                         config_.resolution_inv,
                         config_.origin_at_center ? 1 : 0, l_free_,
                         l_occupied_, l_unknown_, f3(center), radius, filter,
                         max_results, &cfg, &est, &output_cap)) {
    return out;
  }
  const long long buf = std::min<long long>(output_cap, est);
  ensureFetchCapacity(buf);
  CUBLOX_CHECK(cublox_memset_async(d_fetch_count_, 0, sizeof(unsigned int), stream_));
  const unsigned int max_out =
      max_results > 0 ? std::min(static_cast<unsigned int>(max_results), output_cap)
                      : output_cap;
  CUBLOX_CHECK(cublox_launch_fetch_occupancy_around(
      &cfg, d_occ_, d_fetch_count_, d_fetch_pos_, d_fetch_state_,
      static_cast<unsigned int>(buf), max_out, stream_));
  unsigned int n = 0;
  CUBLOX_CHECK(cublox_memcpy_d2h_async(&n, d_fetch_count_, sizeof(n), stream_));
  sync();
  n = std::min(n, output_cap);
  if (max_results > 0) n = std::min(n, static_cast<unsigned int>(max_results));
  n = std::min<unsigned int>(n, static_cast<unsigned int>(buf));
  if (n == 0) return out;

  std::vector<float> pos(static_cast<size_t>(n) * 3);
  std::vector<unsigned char> st(n);
  {
    void *const dst[2] = {pos.data(), st.data()};
    const void *const src[2] = {d_fetch_pos_, d_fetch_state_};
    const size_t len[2] = {pos.size() * sizeof(float), st.size()};
    downloadHost(dst, src, len, 2);
  }
  out.reserve(n);
  for (unsigned int i = 0; i < n; ++i) {
    out.push_back(OccupancySample{
        Eigen::Vector3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
        static_cast<VoxelState>(st[i])});
  }
  return out;
}

size_t OccupancyGrid::buildOccupancyVizCloud(const Eigen::Vector3f &robot_pos,
                                             const float max_viz_radius,
                                             const int subsample,
                                             const int max_points,
                                             std::vector<VizPoint> &out) const {
  join();
  out.clear();
  if (d_occ_ == nullptr || config_.voxel_num <= 0) return 0;
  // main.cpp:484-485; fetchOccupancyAround returns nothing for r <= 0.
  const float radius = std::min(max_viz_radius, maxHorizontalInWindowRadius(robot_pos));
  if (!(radius > 0.0f)) return 0;
  cublox_fetch_cfg cfg{};
  long long est = 0;
  unsigned int output_cap = 0;
  if (!cublox_fetch_plan(i3(origin_i_), i3(config_.map_size_i),
                         i3(config_.half_map_size_i), config_.resolution,
                         config_.resolution_inv, config_.origin_at_center ? 1 : 0,
                         l_free_, l_occupied_, l_unknown_, f3(robot_pos), radius,
                         static_cast<int>(VoxelState::OCCUPIED), max_points, &cfg,
                         &est, &output_cap)) {
    return 0;
  }
  const long long buf = std::min<long long>(output_cap, est);
  if (buf <= 0) return 0;
  const unsigned int max_out =
      max_points > 0 ? std::min(static_cast<unsigned int>(max_points), output_cap)
                     : output_cap;
  const long long out_cap = std::min<long long>(buf, max_out);
  if (out_cap > viz_capacity_) {
    sync();
    freeDevice(d_viz_);
    viz_capacity_ = 0;
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_viz_),
                               static_cast<size_t>(out_cap) * 4 * sizeof(float)));
    viz_capacity_ = out_cap;
  }
  if (!d_viz_count_) {
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_viz_count_), 2 * sizeof(unsigned int)));
  }
  cublox_viz_cfg viz{};
  viz.robot_pos = f3(robot_pos);
  viz.max_viz_radius = max_viz_radius;
  viz.inv_resolution = config_.resolution_inv;
  viz.origin_at_center = config_.origin_at_center ? 1 : 0;
  viz.subsample = subsample;
  viz.max_points = static_cast<int>(max_out);
  CUBLOX_CHECK(cublox_launch_viz_cloud(&cfg, d_occ_, &viz, static_cast<unsigned int>(buf),
                                       d_viz_count_, d_viz_count_ + 1, d_viz_,
                                       static_cast<unsigned int>(out_cap), stream_));
  unsigned int counts[2] = {0, 0};
  CUBLOX_CHECK(cublox_memcpy_d2h_async(counts, d_viz_count_, sizeof(counts), stream_));
  sync();
  const unsigned int written =
      std::min<unsigned int>(counts[1], static_cast<unsigned int>(out_cap));
  if (written > 0) {
    out.resize(written);
    void *const dst[1] = {out.data()};
    const void *const src[1] = {d_viz_};
    const size_t len[1] = {static_cast<size_t>(written) * sizeof(VizPoint)};
    downloadHost(dst, src, len, 1);
  }
  // fetchOccupancyAround's sample count (OccupancyGrid.cpp:335-356).
  unsigned int n = std::min(counts[0], output_cap);
  if (max_points > 0) n = std::min(n, static_cast<unsigned int>(max_points));
  return std::min<unsigned int>(n, static_cast<unsigned int>(buf));
}

// ---- map snapshots ----------------------------------------------------

// Device bytes to host memory through the pinned stage (pageable copies
// run at a few GB/s): every part goes out by DMA into the stage, then one
// host copy each. Syncs the stream.
void OccupancyGrid::downloadHost(void *const *h_dsts, const void *const *d_srcs,
                                 const size_t *lens, int parts) const {
  if (parts < 0 || parts > 8) throw std::logic_error("downloadHost: at most 8 parts");
  size_t off[8];
  size_t total = 0;
  for (int k = 0; k < parts; ++k) {
    off[k] = total;
    total = (total + lens[k] + 255) & ~static_cast<size_t>(255);
  }
  if (total == 0) return;
  uint8_t *stage = pinnedStage(total);
  for (int k = 0; k < parts; ++k) {
    if (lens[k]) CUBLOX_CHECK(cublox_memcpy_d2h_async(stage + off[k], d_srcs[k], lens[k], stream_));
  }
  sync();
  for (int k = 0; k < parts; ++k) {
    if (lens[k]) std::memcpy(h_dsts[k], stage + off[k], lens[k]);
  }
}

uint8_t *OccupancyGrid::pinnedStage(size_t bytes) const {
  if (bytes > h_stage_bytes_) {
    if (h_stage_) cublox_host_free(h_stage_);
    h_stage_ = nullptr;
    h_stage_bytes_ = 0;
    CUBLOX_CHECK(cublox_host_alloc(&h_stage_, bytes));
    h_stage_bytes_ = bytes;
  }
  return static_cast<uint8_t *>(h_stage_);
}

std::vector<uint8_t> OccupancyGrid::saveSnapshot() const {
  join();
  cublox_snapshot_header h{};
  std::memcpy(h.magic, "CBLXSNAP", 8);
  h.version = CUBLOX_SNAPSHOT_VERSION;
  for (int a = 0; a < 3; ++a) {
    h.half_map_size[a] = config_.half_map_size_i[a];
    h.origin_i[a] = origin_i_[a];
    h.origin_f[a] = origin_f_[a];
  }
  h.resolution = config_.resolution;
  h.origin_at_center = config_.origin_at_center ? 1 : 0;
  h.has_threshold = config_.recenter_threshold.has_value() ? 1 : 0;
  h.recenter_threshold = config_.recenter_threshold.value_or(0.0f);
  h.first_run = first_run_ ? 1 : 0;
  h.l_hit = l_hit_;
  h.l_miss = l_miss_;
  h.l_min = l_min_;
  h.l_max = l_max_;
  h.l_free = l_free_;
  h.l_occupied = l_occupied_;
  h.l_unknown = l_unknown_;
  h.max_raycast_range = max_raycast_range_;
  h.voxel_num = config_.voxel_num;
  const size_t V = static_cast<size_t>(std::max(config_.voxel_num, 0));
  std::vector<uint8_t> out;
  if (V == 0 || d_occ_ == nullptr) {
    h.encoding = CUBLOX_SNAPSHOT_DENSE;
    h.payload_bytes = 0;
    out.resize(sizeof(h));
    std::memcpy(out.data(), &h, sizeof(h));
    return out;
  }
  // Count the nonzero voxels; sparse when 8 bytes each beat 4 per voxel.
  const size_t scratch = cublox_sparse_scratch_bytes(config_.voxel_num);
  void *d_scratch = nullptr;
  int *d_ids = nullptr;
  float *d_vals = nullptr;
  try {
    CUBLOX_CHECK(cublox_malloc(&d_scratch, scratch + 256));
    unsigned int *d_count = reinterpret_cast<unsigned int *>(
        static_cast<char *>(d_scratch) + scratch);
    CUBLOX_CHECK(cublox_launch_sparse_compact(d_occ_, config_.voxel_num, d_scratch,
                                              scratch, nullptr, nullptr, 0u,
                                              d_count, stream_));
    unsigned int nnz = 0;
    CUBLOX_CHECK(cublox_memcpy_d2h_async(&nnz, d_count, sizeof(nnz), stream_));
    sync();
    if (static_cast<size_t>(nnz) * 8 + 8 < V * 4) {
      h.encoding = CUBLOX_SNAPSHOT_SPARSE;
      h.payload_bytes = 8 + static_cast<uint64_t>(nnz) * 8;
      // Device -> pinned staging -> the output, one host copy.
      uint8_t *p = pinnedStage(h.payload_bytes);
      const uint32_t head[2] = {nnz, 0u};
      std::memcpy(p, head, 8);
      if (nnz > 0) {
        CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_ids), nnz * sizeof(int)));
        CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_vals), nnz * sizeof(float)));
        CUBLOX_CHECK(cublox_launch_sparse_compact(d_occ_, config_.voxel_num, d_scratch,
                                                  scratch, d_ids, d_vals, nnz,
                                                  d_count, stream_));
        CUBLOX_CHECK(cublox_memcpy_d2h_async(p + 8, d_ids, nnz * sizeof(int), stream_));
        CUBLOX_CHECK(cublox_memcpy_d2h_async(p + 8 + static_cast<size_t>(nnz) * 4, d_vals,
                                             nnz * sizeof(float), stream_));
        sync();
      }
      out.reserve(sizeof(h) + h.payload_bytes);
      out.resize(sizeof(h));
      out.insert(out.end(), p, p + h.payload_bytes);
    } else {
      h.encoding = CUBLOX_SNAPSHOT_DENSE;
      h.payload_bytes = V * 4;
      uint8_t *p = pinnedStage(h.payload_bytes);
      CUBLOX_CHECK(cublox_memcpy_d2h_async(p, d_occ_, V * 4, stream_));
      sync();
      out.reserve(sizeof(h) + h.payload_bytes);
      out.resize(sizeof(h));
      out.insert(out.end(), p, p + h.payload_bytes);
    }
  } catch (...) {
    freeDevice(d_ids);
    freeDevice(d_vals);
    freeDevice(d_scratch);
    throw;
  }
  freeDevice(d_ids);
  freeDevice(d_vals);
  freeDevice(d_scratch);
  std::memcpy(out.data(), &h, sizeof(h));
  return out;
}

void OccupancyGrid::loadSnapshot(const uint8_t *data, size_t bytes) {
  join();
  auto bad = [](const char *why) {
    throw std::invalid_argument(std::string("cublox::OccupancyGrid::loadSnapshot: ") + why);
  };
  if (data == nullptr || bytes < sizeof(cublox_snapshot_header)) bad("truncated header");
  cublox_snapshot_header h;
  std::memcpy(&h, data, sizeof(h));
  if (std::memcmp(h.magic, "CBLXSNAP", 8) != 0) bad("not a cublox snapshot");
  if (h.version != CUBLOX_SNAPSHOT_VERSION) bad("unsupported version");
  for (int a = 0; a < 3; ++a) {
    if (h.half_map_size[a] != config_.half_map_size_i[a]) bad("half map size differs");
  }
  if (h.resolution != config_.resolution) bad("resolution differs");
  if ((h.origin_at_center != 0) != config_.origin_at_center) bad("binning differs");
  if (h.voxel_num != config_.voxel_num) bad("voxel count differs");
  if (h.payload_bytes != bytes - sizeof(h)) bad("payload size differs from the header");
  const size_t V = static_cast<size_t>(std::max(config_.voxel_num, 0));
  const uint8_t *p = data + sizeof(h);
  uint32_t nnz = 0;
  if (h.encoding == CUBLOX_SNAPSHOT_DENSE) {
    if (h.payload_bytes != V * 4 && !(h.payload_bytes == 0 && d_occ_ == nullptr)) {
      bad("dense payload is not one float per voxel");
    }
  } else if (h.encoding == CUBLOX_SNAPSHOT_SPARSE) {
    if (h.payload_bytes < 8) bad("sparse payload without a count");
    std::memcpy(&nnz, p, 4);
    if (h.payload_bytes != 8 + static_cast<uint64_t>(nnz) * 8 || nnz > V) {
      bad("sparse payload size differs from its count");
    }
  } else {
    bad("unknown encoding");
  }
  if (V > 0 && d_occ_ != nullptr) {
    sync();
    if (h.encoding == CUBLOX_SNAPSHOT_DENSE) {
      CUBLOX_CHECK(cublox_memcpy_h2d_async(d_occ_, p, V * 4, stream_));
      sync();
    } else {
      int *d_ids = nullptr;
      float *d_vals = nullptr;
      try {
        if (nnz > 0) {
          CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_ids), nnz * sizeof(int)));
          CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_vals), nnz * sizeof(float)));
          CUBLOX_CHECK(cublox_memcpy_h2d_async(d_ids, p + 8, nnz * sizeof(int), stream_));
          CUBLOX_CHECK(cublox_memcpy_h2d_async(d_vals, p + 8 + static_cast<size_t>(nnz) * 4,
                                               nnz * sizeof(float), stream_));
        }
        CUBLOX_CHECK(cublox_launch_sparse_restore(d_occ_, config_.voxel_num, d_ids, d_vals,
                                                  nnz, stream_));
        sync();
      } catch (...) {
        freeDevice(d_ids);
        freeDevice(d_vals);
        throw;
      }
      freeDevice(d_ids);
      freeDevice(d_vals);
    }
  }
  // Counters (per-voxel path) start from zero after a restore.
  if (d_op_cnt_) {
    CUBLOX_CHECK(cublox_fill_zero_async(d_op_cnt_, V, stream_));
    CUBLOX_CHECK(cublox_fill_zero_async(d_hit_cnt_, V, stream_));
    sync();
  }
  updateOriginAndBound(Eigen::Vector3i(h.origin_i[0], h.origin_i[1], h.origin_i[2]));
  first_run_ = h.first_run != 0;
  if (h.has_threshold) {
    config_.recenter_threshold = h.recenter_threshold;
    config_.recenter_threshold_squared = h.recenter_threshold * h.recenter_threshold;
  } else {
    config_.recenter_threshold.reset();
    config_.recenter_threshold_squared = 0.0f;
  }
  l_hit_ = h.l_hit;
  l_miss_ = h.l_miss;
  l_min_ = h.l_min;
  l_max_ = h.l_max;
  l_free_ = h.l_free;
  l_occupied_ = h.l_occupied;
  l_unknown_ = h.l_unknown;
  max_raycast_range_ = h.max_raycast_range;
}

void OccupancyGrid::saveSnapshotFile(const std::string &path) const {
  const std::vector<uint8_t> b = saveSnapshot();
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  f.write(reinterpret_cast<const char *>(b.data()), static_cast<std::streamsize>(b.size()));
  if (!f) throw std::runtime_error("cublox::OccupancyGrid::saveSnapshotFile: cannot write " + path);
}

void OccupancyGrid::loadSnapshotFile(const std::string &path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cublox::OccupancyGrid::loadSnapshotFile: cannot read " + path);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<uint8_t> b(static_cast<size_t>(std::max<std::streamsize>(n, 0)));
  f.read(reinterpret_cast<char *>(b.data()), n);
  if (!f) throw std::runtime_error("cublox::OccupancyGrid::loadSnapshotFile: cannot read " + path);
  loadSnapshot(b.data(), b.size());
}

void OccupancyGrid::fetchOccupancyBatch(
    const std::vector<Eigen::Vector3f> &centers, const float radius,
    const std::optional<VoxelState> state_filter, const int per_query_cap,
    std::vector<unsigned int> &counts, std::vector<float> &positions,
    std::vector<unsigned char> &states) const {
  join();
  const int nq = static_cast<int>(centers.size());
  counts.assign(static_cast<size_t>(nq), 0u);
  positions.clear();
  states.clear();
  if (nq == 0 || d_occ_ == nullptr || per_query_cap <= 0) return;
  const int filter = state_filter ? static_cast<int>(*state_filter) : -1;
  std::vector<cublox_fetch_cfg> cfgs(static_cast<size_t>(nq));
  int max_diam = 0, max_nz = 0;
  for (int q = 0; q < nq; ++q) {
    cublox_fetch_cfg &c = cfgs[static_cast<size_t>(q)];
    long long est = 0;
    unsigned int cap = 0;
    if (cublox_fetch_plan(i3(origin_i_), i3(config_.map_size_i),
                          i3(config_.half_map_size_i), config_.resolution,
                          config_.resolution_inv,
                          config_.origin_at_center ? 1 : 0, l_free_,
                          l_occupied_, l_unknown_, f3(centers[q]), radius,
                          filter, per_query_cap, &c, &est, &cap)) {
      max_diam = std::max(max_diam, 2 * c.disk_radius_vox + 1);
      max_nz = std::max(max_nz, c.gz_max - c.gz_min + 1);
    } else {
      c = cublox_fetch_cfg{};
      c.disk_radius_vox = -1;  // no work for this query
      c.gz_min = 1;
      c.gz_max = 0;
    }
  }
  if (max_diam == 0) return;
  const long long cap_total = static_cast<long long>(nq) * per_query_cap;
  ensureFetchCapacity(cap_total);
  if (nq > fetch_cfg_capacity_) {
    sync();
    freeDevice(d_fetch_cfgs_);
    freeDevice(d_fetch_counts_);
    fetch_cfg_capacity_ = 0;
    CUBLOX_CHECK(cublox_malloc(&d_fetch_cfgs_, sizeof(cublox_fetch_cfg) * nq));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_counts_),
                               sizeof(unsigned int) * nq));
    fetch_cfg_capacity_ = nq;
  }
  // Transfers go through the grid's pinned stage (pageable copies run at a
  // few GB/s): the configurations and the counts first, then (after the
  // sync below) the offsets and the samples.
  const size_t cfg_bytes = sizeof(cublox_fetch_cfg) * static_cast<size_t>(nq);
  const size_t cnt_off = (cfg_bytes + 255) & ~static_cast<size_t>(255);
  uint8_t *stage = pinnedStage(cnt_off + sizeof(unsigned int) * static_cast<size_t>(nq));
  std::memcpy(stage, cfgs.data(), cfg_bytes);
  CUBLOX_CHECK(cublox_memcpy_h2d_async(d_fetch_cfgs_, stage, cfg_bytes, stream_));
  CUBLOX_CHECK(cublox_memset_async(d_fetch_counts_, 0, sizeof(unsigned int) * nq,
                                   stream_));
  mark(6);
  CUBLOX_CHECK(cublox_launch_fetch_occupancy_batch(
      static_cast<const cublox_fetch_cfg *>(d_fetch_cfgs_), nq, max_diam, max_nz,
      d_occ_, d_fetch_counts_, d_fetch_pos_, d_fetch_state_,
      static_cast<unsigned int>(per_query_cap), stream_));
  mark(7);
  CUBLOX_CHECK(cublox_memcpy_d2h_async(stage + cnt_off, d_fetch_counts_,
                                       sizeof(unsigned int) * nq, stream_));
  sync();
  std::memcpy(counts.data(), stage + cnt_off, sizeof(unsigned int) * nq);
  if (timing_) {
    float ms = 0.0f;
    CUBLOX_CHECK(cublox_event_elapsed_ms(events_[6], events_[7], &ms));
    times_.fetch_ms += ms;
    times_.fetch_launches += 1;
    for (const cublox_fetch_cfg &c : cfgs) {
      if (c.disk_radius_vox < 0) continue;
      const long long d = 2LL * c.disk_radius_vox + 1;
      times_.fetch_tested_voxels += d * d * (c.gz_max - c.gz_min + 1);
    }
    for (int q = 0; q < nq; ++q) {
      times_.fetch_samples += std::min(counts[q], static_cast<unsigned int>(per_query_cap));
    }
  }
  // Compact on the device: query q's samples go to offsets[q] of a dense
  // buffer, so only the real samples cross PCIe.
  std::vector<unsigned int> offsets(static_cast<size_t>(nq) + 1, 0u);
  unsigned int max_count = 0;
  for (int q = 0; q < nq; ++q) {
    counts[q] = std::min(counts[q], static_cast<unsigned int>(per_query_cap));
    offsets[q + 1] = offsets[q] + counts[q];
    max_count = std::max(max_count, counts[q]);
  }
  const size_t total = offsets[nq];
  if (total == 0) return;
  if (static_cast<long long>(total) > fetch_dense_capacity_ ||
      nq + 1 > fetch_offsets_capacity_) {
    sync();
    freeDevice(d_fetch_dense_pos_);
    freeDevice(d_fetch_dense_state_);
    freeDevice(d_fetch_offsets_);
    fetch_dense_capacity_ = 0;
    fetch_offsets_capacity_ = 0;
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_dense_pos_),
                               total * 3 * sizeof(float)));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_dense_state_), total));
    CUBLOX_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_fetch_offsets_),
                               sizeof(unsigned int) * (nq + 1)));
    fetch_dense_capacity_ = static_cast<long long>(total);
    fetch_offsets_capacity_ = nq + 1;
  }
  const size_t off_bytes = sizeof(unsigned int) * (static_cast<size_t>(nq) + 1);
  const size_t pos_off = (off_bytes + 255) & ~static_cast<size_t>(255);
  const size_t st_off = (pos_off + total * 3 * sizeof(float) + 255) & ~static_cast<size_t>(255);
  stage = pinnedStage(st_off + total);  // nothing in flight reads the stage now
  std::memcpy(stage, offsets.data(), off_bytes);
  CUBLOX_CHECK(cublox_memcpy_h2d_async(d_fetch_offsets_, stage, off_bytes, stream_));
  CUBLOX_CHECK(cublox_launch_fetch_gather(
      d_fetch_counts_, d_fetch_offsets_, nq,
      static_cast<unsigned int>(per_query_cap), max_count, d_fetch_pos_,
      d_fetch_state_, d_fetch_dense_pos_, d_fetch_dense_state_, stream_));
  CUBLOX_CHECK(cublox_memcpy_d2h_async(stage + pos_off, d_fetch_dense_pos_,
                                       total * 3 * sizeof(float), stream_));
  CUBLOX_CHECK(cublox_memcpy_d2h_async(stage + st_off, d_fetch_dense_state_, total,
                                       stream_));
  sync();
  positions.resize(total * 3);
  states.resize(total);
  std::memcpy(positions.data(), stage + pos_off, total * 3 * sizeof(float));
  std::memcpy(states.data(), stage + st_off, total);
}

float OccupancyGrid::storedAt(const int hash_id) const {
  join();
  if (d_occ_ == nullptr || hash_id < 0 || hash_id >= config_.voxel_num) return 0.0f;
  float v = 0.0f;
  CUBLOX_CHECK(cublox_memcpy_d2h_async(&v, d_occ_ + hash_id, sizeof(float), stream_));
  sync();
  return v;
}

bool OccupancyGrid::isOccupied(const Eigen::Vector3f &pos) const {
  if (!inside(pos)) return false;
  return occOccupied(storedAt(posToHashIndex(pos, config_.map_size_i,
                                             config_.half_map_size_i,
                                             config_.resolution_inv,
                                             config_.origin_at_center)));
}

bool OccupancyGrid::isUnknown(const Eigen::Vector3f &pos) const {
  if (!inside(pos)) return true;
  return occUnknown(storedAt(posToHashIndex(pos, config_.map_size_i,
                                            config_.half_map_size_i,
                                            config_.resolution_inv,
                                            config_.origin_at_center)));
}

bool OccupancyGrid::isFree(const Eigen::Vector3f &pos) const {
  if (!inside(pos)) return false;
  return occFree(storedAt(posToHashIndex(pos, config_.map_size_i,
                                         config_.half_map_size_i,
                                         config_.resolution_inv,
                                         config_.origin_at_center)));
}

bool OccupancyGrid::isOccupied(const Eigen::Vector3i &id_g) const {
  if (!inside(id_g)) return false;
  return occOccupied(storedAt(
      globalIndexToHashId(id_g, config_.map_size_i, config_.half_map_size_i)));
}

bool OccupancyGrid::isUnknown(const Eigen::Vector3i &id_g) const {
  if (!inside(id_g)) return true;
  return occUnknown(storedAt(
      globalIndexToHashId(id_g, config_.map_size_i, config_.half_map_size_i)));
}

bool OccupancyGrid::isFree(const Eigen::Vector3i &id_g) const {
  if (!inside(id_g)) return false;
  return occFree(storedAt(
      globalIndexToHashId(id_g, config_.map_size_i, config_.half_map_size_i)));
}

bool OccupancyGrid::isOccupied(const int hash_id) const {
  return occOccupied(storedAt(hash_id));
}
bool OccupancyGrid::isUnknown(const int hash_id) const {
  return occUnknown(storedAt(hash_id));
}
bool OccupancyGrid::isFree(const int hash_id) const {
  return occFree(storedAt(hash_id));
}

void OccupancyGrid::copyOccupancyToHost(float *dst) const {
  join();
  if (!d_occ_) return;
  CUBLOX_CHECK(cublox_memcpy_d2h_async(
      dst, d_occ_, static_cast<size_t>(config_.voxel_num) * sizeof(float), stream_));
  sync();
}

void OccupancyGrid::copyCountersToHost(int *op, int *hit) const {
  join();
  if (!d_occ_) return;
  const size_t bytes = static_cast<size_t>(config_.voxel_num) * sizeof(int);
  if (!d_op_cnt_) {
    if (op) std::fill(op, op + config_.voxel_num, 0);
    if (hit) std::fill(hit, hit + config_.voxel_num, 0);
    return;
  }
  if (op) CUBLOX_CHECK(cublox_memcpy_d2h_async(op, d_op_cnt_, bytes, stream_));
  if (hit) CUBLOX_CHECK(cublox_memcpy_d2h_async(hit, d_hit_cnt_, bytes, stream_));
  sync();
}

void OccupancyGrid::loadOccupancyFromHost(const float *src) {
  join();
  if (!d_occ_) return;
  CUBLOX_CHECK(cublox_memcpy_h2d_async(
      d_occ_, src, static_cast<size_t>(config_.voxel_num) * sizeof(float), stream_));
  sync();
}

} // namespace cublox
