// Batched updates of many occupancy grids: host orchestration over the
// batched C ABI (cublox_integrate_batch_*, cublox_recenter_batch_*).
#include <cublox/GridBatch.hpp>
#include <cublox/capi.h>
#include <cublox/utils.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>

namespace cublox {

namespace {

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

cublox_raycast_cfg raycastCfg(const OccupancyGrid &g, float max_range) {
  cublox_raycast_cfg c{};
  c.origin_i = i3(g.origin_i_);
  c.map_size_i = i3(g.config_.map_size_i);
  c.half_map_size_i = i3(g.config_.half_map_size_i);
  c.resolution = g.config_.resolution;
  c.inv_resolution = g.config_.resolution_inv;
  c.max_range = max_range;
  c.map_vox_num = g.config_.voxel_num;
  c.origin_at_center = g.config_.origin_at_center ? 1 : 0;
  return c;
}

} // namespace

static cublox_batch_limits abiLimits(const BatchLimits &l) {
  cublox_batch_limits c{};
  c.max_grids = l.max_grids;
  c.max_points = l.max_points;
  c.max_tiles = l.max_tiles;
  c.max_segments_per_ray = l.max_segments_per_ray;
  c.max_near_steps = l.max_near_steps;
  c.chunk = l.min_chunk;
  c.max_slabs = l.max_slabs;
  c.delta = l.changed_voxels ? 1 : 0;
  return c;
}

GridBatch::GridBatch(const BatchLimits &limits) : lim_(limits) {
  if (lim_.max_grids < 1 || lim_.max_points < 1 || lim_.max_tiles < 1 ||
      lim_.max_segments_per_ray < 1 || lim_.max_near_steps < 1 ||
      lim_.max_near_steps > 12) {
    throw std::invalid_argument("cublox::GridBatch: invalid limits");
  }
  setting_.near_steps = std::min(setting_.near_steps, lim_.max_near_steps);
  setting_.chunk = std::max(setting_.chunk, lim_.min_chunk);
  max_regions_ = lim_.max_regions > 0 ? lim_.max_regions : 6 * lim_.max_grids;
  const cublox_batch_limits c = abiLimits(lim_);
  ws_bytes_ = cublox_integrate_batch_workspace_bytes(&c);
  try {
    allocSlot(slots_[0]);
  } catch (...) {
    release();
    throw;
  }
}

GridBatch::~GridBatch() { release(); }

void GridBatch::allocSlot(Slot &s, bool dedicated) {
  const cublox_batch_limits c = abiLimits(lim_);
  if (dedicated) {
    CUBLOX_CHECK(cublox_stream_create_dedicated(&s.stream));
  } else {
    CUBLOX_CHECK(cublox_stream_create(&s.stream));
  }
  s.dedicated = dedicated;
  for (auto &e : s.ev) CUBLOX_CHECK(cublox_event_create(&e));
  CUBLOX_CHECK(cublox_malloc(&s.d_ws, ws_bytes_));
  CUBLOX_CHECK(cublox_integrate_workspace_init(s.d_ws, ws_bytes_, s.stream));
  CUBLOX_CHECK(cublox_host_alloc(&s.h_stage, cublox_integrate_batch_staging_bytes(&c)));
  const size_t rb = cublox_recenter_batch_staging_bytes(max_regions_);
  CUBLOX_CHECK(cublox_host_alloc(&s.h_rstage, rb));
  CUBLOX_CHECK(cublox_malloc(&s.d_rarea, rb));
  CUBLOX_CHECK(cublox_stream_synchronize(s.stream));
}

void GridBatch::dropGraphs(Slot &s) {
  auto drop = [](void *&g) {
    if (g) cublox_graph_destroy(g);
    g = nullptr;
  };
  for (auto &by_near : s.front) {
    for (void *&g : by_near) drop(g);
  }
  for (void *&g : s.back) drop(g);
}

void GridBatch::releaseSlot(Slot &s) noexcept {
  if (s.stream) cublox_stream_synchronize(s.stream);
  dropGraphs(s);
  if (s.d_ws) cublox_free(s.d_ws);
  if (s.d_rarea) cublox_free(s.d_rarea);
  if (s.h_stage) cublox_host_free(s.h_stage);
  if (s.h_rstage) cublox_host_free(s.h_rstage);
  s.d_ws = s.d_rarea = s.h_stage = s.h_rstage = nullptr;
  for (auto &e : s.ev) {
    if (e) cublox_event_destroy(e);
    e = nullptr;
  }
  if (s.stream) cublox_stream_destroy(s.stream);
  s.stream = nullptr;
  s.in_flight = false;
}

void GridBatch::release() noexcept {
  // Both streams first: a tick's folds may wait on the other slot.
  for (auto &s : slots_) {
    if (s.stream) cublox_stream_synchronize(s.stream);
  }
  for (auto &s : slots_) releaseSlot(s);
}

std::array<int, 5> GridBatch::lastCounters() const {
  std::array<int, 5> out{};
  const cublox_batch_limits c = abiLimits(lim_);
  const Slot &s = slots_[cur_];
  CUBLOX_CHECK(cublox_integrate_batch_counters(s.d_ws, ws_bytes_, &c, out.data(),
                                               s.stream));
  return out;
}

int GridBatch::segmentBound(const OccupancyGrid &g) {
  const cublox_raycast_cfg c = raycastCfg(g, g.getMaxRaycastRange());
  return cublox_integrate_segment_bound(&c);
}

int GridBatch::tileCount(const OccupancyGrid &g) {
  const cublox_raycast_cfg c = raycastCfg(g, g.getMaxRaycastRange());
  return cublox_integrate_tile_count(&c);
}

void GridBatch::setGraphMode(bool on) {
  if (!on) {
    wait();
    for (auto &s : slots_) {
      if (!s.stream) continue;
      CUBLOX_CHECK(cublox_stream_synchronize(s.stream));
      dropGraphs(s);
    }
  }
  graph_mode_ = on;
}

void GridBatch::setPipelined(bool on) {
  if (on && lim_.changed_voxels) {
    throw std::invalid_argument(
        "cublox::GridBatch: pipelined ticks do not feed changed-voxel streams");
  }
  wait();
  if (on) {
    // Overlapped ticks need their two streams on different hardware
    // queues: both slots take streams with queues of their own.
    if (!slots_[0].dedicated) {
      void *old = slots_[0].stream;
      CUBLOX_CHECK(cublox_stream_synchronize(old));
      CUBLOX_CHECK(cublox_stream_create_dedicated(&slots_[0].stream));
      slots_[0].dedicated = true;
      cublox_stream_destroy(old);
    }
    if (!slots_[1].stream) {
      try {
        allocSlot(slots_[1], true);
      } catch (...) {
        releaseSlot(slots_[1]);
        throw;
      }
    }
    n_slots_ = 2;
  } else {
    releaseSlot(slots_[1]);
    n_slots_ = 1;
    cur_ = 0;
  }
}

bool GridBatch::inFlight() const {
  return slots_[0].in_flight || slots_[1].in_flight;
}

static bool settingValid(const TuneSetting &s, const BatchLimits &lim) {
  return s.near_steps >= 1 && s.near_steps <= lim.max_near_steps &&
         s.chunk >= lim.min_chunk && s.chunk <= 65535 &&
         (s.lanes == 0 || s.lanes == 1 || s.lanes == 4) &&
         (s.hot_mode == 0 || s.hot_mode == 1);
}

void GridBatch::setSetting(const TuneSetting &s) {
  if (!settingValid(s, lim_)) {
    throw std::invalid_argument("cublox::GridBatch: setting outside the limits");
  }
  setting_ = s;
}

void GridBatch::setTuning(const std::vector<TuneSetting> &candidates,
                          int ticks_each) {
  for (const auto &s : candidates) {
    if (!settingValid(s, lim_)) {
      throw std::invalid_argument("cublox::GridBatch: candidate outside the limits");
    }
  }
  tune_ = candidates;
  tune_ms_.assign(candidates.size(), 0.0);
  tune_n_.assign(candidates.size(), 0);
  tune_pos_ = 0;
  tune_ticks_ = std::max(ticks_each, 1);
  tune_count_ = 0;
}

std::vector<std::pair<TuneSetting, double>> GridBatch::tuneResults() const {
  std::vector<std::pair<TuneSetting, double>> out;
  for (size_t k = 0; k < tune_.size(); ++k) {
    out.emplace_back(tune_[k], tune_n_[k] ? tune_ms_[k] / tune_n_[k] : 0.0);
  }
  return out;
}

// Host staging of one tick into slot s: the update descriptors (pinning
// each grid's first origin, as update() does) and the recenter plans.
void GridBatch::stage(
    Slot &slot, const std::vector<GridScan> &scans,
    const std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>> &recenters) {
  std::vector<cublox_integrate_grid> items(scans.size());
  for (const GridScan &s : scans) {
    if (s.grid && s.grid->delta_capacity_ > 0 && !lim_.changed_voxels) {
      throw std::invalid_argument(
          "cublox::GridBatch: a grid records changed voxels; set "
          "BatchLimits::changed_voxels");
    }
  }
  for (size_t k = 0; k < scans.size(); ++k) {
    const GridScan &s = scans[k];
    if (!s.grid) throw std::invalid_argument("cublox::GridBatch: null grid");
    OccupancyGrid &g = *s.grid;
    g.settle();  // a plain recenter's clear still running on the grid's side stream
    const Eigen::Vector3f sensor =
        s.sensor_frame ? Eigen::Vector3f(s.T[3], s.T[7], s.T[11]) : s.origin;
    g.pinFirstOrigin(sensor);
    cublox_integrate_grid &q = items[k];
    q.cfg = raycastCfg(g, g.max_raycast_range_);
    q.cfg.origin = f3(sensor);
    for (int i = 0; i < 12; ++i) q.T[i] = s.T[static_cast<size_t>(i)];
    q.use_T = s.sensor_frame ? 1 : 0;
    q.d_xs = s.d_x;
    q.d_ys = s.d_y;
    q.d_zs = s.d_z;
    q.n = s.n;
    q.d_occ = g.d_occ_;
    q.l_hit = g.l_hit_;
    q.l_miss = g.l_miss_;
    q.l_min = g.l_min_;
    q.l_max = g.l_max_;
    q.l_unknown = g.l_unknown_;
    q.near_steps = setting_.near_steps;
    q.d_export_op = nullptr;
    q.d_export_hit = nullptr;
    q.point_stride = s.point_stride;
    q.range_drop_sq = s.range_drop_sq;
    if (g.delta_capacity_ > 0) {
      q.d_delta_count = g.d_delta_count_;
      q.d_delta_idx = g.d_delta_idx_;
      q.d_delta_val = g.d_delta_val_;
      q.delta_capacity = g.delta_capacity_;
    }
  }
  const cublox_batch_limits c = abiLimits(lim_);
  cublox_batch_tuning tune{};
  tune.chunk = setting_.chunk;
  tune.lanes_per_ray = setting_.lanes;
  tune.hot_mode = setting_.hot_mode;
  CUBLOX_CHECK(cublox_integrate_batch_stage_tuned(
      items.data(), static_cast<int>(items.size()), &c, &tune, slot.h_stage));
  slot.lanes = tune.lanes_per_ray;
  slot.near_steps = setting_.near_steps;
  slot.hot_mode = setting_.hot_mode;

  // OccupancyGrid.cpp:141-179 per grid: threshold pre-check on the host,
  // then the plan; origins move once the tick is launched.
  std::vector<cublox_recenter_item> ritems;
  std::vector<OccupancyGrid *> rgrids;
  for (const auto &[gp, pos] : recenters) {
    if (!gp) throw std::invalid_argument("cublox::GridBatch: null grid");
    OccupancyGrid &g = *gp;
    g.settle();
    if (!g.config_.recenter_threshold.has_value() || g.d_occ_ == nullptr) continue;
    if ((pos - g.origin_f_).squaredNorm() < g.config_.recenter_threshold_squared) continue;
    cublox_recenter_item it{};
    it.cfg.pos = f3(pos);
    it.cfg.origin_f = f3(g.origin_f_);
    it.cfg.origin_i = i3(g.origin_i_);
    it.cfg.map_size_i = i3(g.config_.map_size_i);
    it.cfg.half_map_size_i = i3(g.config_.half_map_size_i);
    it.cfg.resolution = g.config_.resolution;
    it.cfg.inv_resolution = g.config_.resolution_inv;
    it.cfg.recenter_threshold_sq = g.config_.recenter_threshold_squared;
    it.cfg.origin_at_center = g.config_.origin_at_center ? 1 : 0;
    it.cfg.map_vox_num = g.config_.voxel_num;
    it.d_occ = g.d_occ_;
    ritems.push_back(it);
    rgrids.push_back(gp);
  }
  std::vector<cublox_recenter_result> res(ritems.size());
  CUBLOX_CHECK(cublox_recenter_batch_stage(ritems.data(), static_cast<int>(ritems.size()),
                                           max_regions_, slot.h_rstage, res.data()));
  rc_grids_.clear();
  rc_origins_.clear();
  slot.zeroed = 0;
  for (size_t k = 0; k < res.size(); ++k) {
    if (res[k].action == 0) continue;
    rc_grids_.push_back(rgrids[k]);
    rc_origins_.emplace_back(res[k].new_origin_i.x, res[k].new_origin_i.y,
                             res[k].new_origin_i.z);
    slot.zeroed += res[k].zeroed_voxels;
  }
}

// Enqueues slot q's staged tick on its stream. prev >= 0: the slot of the
// tick submitted before it, on the other stream; its stream is marked now
// (after everything enqueued there so far, that tick's folds and recenter
// and any work the caller added) and this tick's folds wait for the mark.
// The first half does not touch occupancy, so it runs beside that work.
void GridBatch::launch(int q, int prev) {
  Slot &s = slots_[q];
  const cublox_batch_limits c = abiLimits(lim_);
  auto front = [&]() {
    // Both staging copies first, so the recenter sequence is one launch.
    CUBLOX_CHECK(cublox_recenter_batch_upload(s.h_rstage, s.d_rarea, max_regions_, s.stream));
    CUBLOX_CHECK(cublox_launch_integrate_batch_staged_part(s.h_stage, s.d_ws, ws_bytes_, &c, 1,
                                                           s.stream));
  };
  auto back = [&]() {
    CUBLOX_CHECK(cublox_launch_integrate_batch_staged_part(s.h_stage, s.d_ws, ws_bytes_, &c, 2,
                                                           s.stream));
    if (!graph_mode_) CUBLOX_CHECK(cublox_event_record(s.ev[1], s.stream));
    CUBLOX_CHECK(cublox_launch_recenter_batch_uploaded(s.d_rarea, max_regions_, s.stream));
  };
  // Runs a part directly, or as its captured graph (captured on first use).
  auto run = [&](auto &&part, void *&graph) {
    if (!graph_mode_) {
      part();
      return;
    }
    if (!graph) {
      CUBLOX_CHECK(cublox_graph_begin_capture(s.stream));
      try {
        part();
      } catch (...) {
        cublox_graph g = nullptr;
        cublox_graph_end_capture(s.stream, &g);  // leave capture mode
        if (g) cublox_graph_destroy(g);
        throw;
      }
      CUBLOX_CHECK(cublox_graph_end_capture(s.stream, &graph));
    }
    CUBLOX_CHECK(cublox_graph_launch(graph, s.stream));
  };
  if (prev >= 0) CUBLOX_CHECK(cublox_event_record(slots_[prev].ev[3], slots_[prev].stream));
  CUBLOX_CHECK(cublox_event_record(s.ev[0], s.stream));
  run(front, s.front[s.lanes == 4 ? 1 : 0][s.near_steps]);
  if (prev >= 0) CUBLOX_CHECK(cublox_stream_wait_event(s.stream, slots_[prev].ev[3]));
  run(back, s.back[s.hot_mode ? 1 : 0]);
  CUBLOX_CHECK(cublox_event_record(s.ev[2], s.stream));
}

void GridBatch::finish(int q) {
  Slot &s = slots_[q];
  s.in_flight = false;
  CUBLOX_CHECK(cublox_stream_synchronize(s.stream));
  float total = 0.0f, upd = 0.0f;
  CUBLOX_CHECK(cublox_event_elapsed_ms(s.ev[0], s.ev[2], &total));
  if (!graph_mode_) {
    CUBLOX_CHECK(cublox_event_elapsed_ms(s.ev[0], s.ev[1], &upd));
  } else {
    upd = total;
  }
  last_update_ms_ = upd;
  last_recenter_ms_ = graph_mode_ ? 0.0 : static_cast<double>(total - upd);
  last_zeroed_ = s.zeroed;
  if (s.tune_idx < 0 || !tuning() || static_cast<size_t>(s.tune_idx) != tune_pos_) return;
  // One warm-up tick per candidate, then ticks_each measured ones; a tick
  // that shared the GPU with another tick is not measured.
  if (tune_count_ > 0 && s.alone) {
    tune_ms_[tune_pos_] += total;
    tune_n_[tune_pos_] += 1;
  }
  if (++tune_count_ > tune_ticks_) {
    tune_count_ = 0;
    if (++tune_pos_ == tune_.size()) {
      size_t best = 0;
      for (size_t k = 1; k < tune_.size(); ++k) {
        if (tune_ms_[k] / std::max(tune_n_[k], 1) <
            tune_ms_[best] / std::max(tune_n_[best], 1)) {
          best = k;
        }
      }
      setting_ = tune_[best];
    }
  }
}

void GridBatch::submit(
    const std::vector<GridScan> &scans,
    const std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>> &recenters) {
  if (n_slots_ == 1 && slots_[0].in_flight) {
    throw std::logic_error("cublox::GridBatch: submit() before wait() of the last tick");
  }
  if (static_cast<int>(scans.size()) > lim_.max_grids) {
    throw std::invalid_argument("cublox::GridBatch: more scans than max_grids");
  }
  int tune_idx = -1;
  if (tuning()) {
    wait();  // an autotuning tick runs alone
    if (tuning()) {
      setting_ = tune_[tune_pos_];
      tune_idx = static_cast<int>(tune_pos_);
    }
  }
  const int q = n_slots_ == 1 ? 0 : 1 - cur_;
  if (slots_[q].in_flight) finish(q);  // the tick two submits back
  Slot &s = slots_[q];
  stage(s, scans, recenters);
  const int prev = q != cur_ ? cur_ : -1;
  const bool overlap = prev >= 0 && slots_[prev].in_flight;
  if (overlap) slots_[prev].alone = false;
  launch(q, prev);
  s.in_flight = true;
  s.alone = !overlap;
  s.tune_idx = tune_idx;
  cur_ = q;
  // The grids' host state describes them after all work enqueued so far.
  for (size_t k = 0; k < rc_grids_.size(); ++k) {
    rc_grids_[k]->updateOriginAndBound(rc_origins_[k]);
  }
}

void GridBatch::wait() {
  // The older tick first.
  if (n_slots_ == 2 && slots_[1 - cur_].in_flight) finish(1 - cur_);
  if (slots_[cur_].in_flight) finish(cur_);
}

void GridBatch::waitOldest() {
  if (n_slots_ == 2 && slots_[1 - cur_].in_flight) {
    finish(1 - cur_);
  } else if (slots_[cur_].in_flight) {
    finish(cur_);
  }
}

void GridBatch::tick(
    const std::vector<GridScan> &scans,
    const std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>> &recenters) {
  submit(scans, recenters);
  wait();
}

}  // namespace cublox
