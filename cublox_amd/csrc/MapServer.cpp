// Occupancy map server: sessions, HBM-budget admission, continuous batching
// of scans over GridBatch, fetches in submission order.
#include <cublox/MapServer.hpp>
#include <cublox/capi.h>

#include <algorithm>
#include <deque>
#include <cstring>
#include <iterator>
#include <stdexcept>
#include <string>

namespace cublox {

namespace {

void statusCheck(int status, const char *expr) {
  if (status == CUBLOX_OK) return;
  throw std::runtime_error(std::string("cublox::MapServer: `") + expr +
                           "` failed: " + cublox_last_error());
}
#define MS_CHECK(call) statusCheck((call), #call)

// Fetch output reserve of a session: the reference's default cap of 500000
// samples at 12 B position + 1 B state, plus the count word.
constexpr std::size_t kFetchReserve = 500000ull * 13ull + 256ull;

// Pinned upload stages (each one largest scan's columns).
constexpr int kUploadStages = 8;

cublox_raycast_cfg shapeCfg(const SessionSpec &s) {
  cublox_raycast_cfg c{};
  c.map_size_i = cublox_int3{2 * s.half_map_size_i.x() + 1, 2 * s.half_map_size_i.y() + 1,
                             2 * s.half_map_size_i.z() + 1};
  c.half_map_size_i =
      cublox_int3{s.half_map_size_i.x(), s.half_map_size_i.y(), s.half_map_size_i.z()};
  c.resolution = s.resolution;
  c.inv_resolution = 1.0f / s.resolution;
  c.max_range = s.max_range;
  c.map_vox_num = c.map_size_i.x * c.map_size_i.y * c.map_size_i.z;
  c.origin_at_center = s.origin_at_center ? 1 : 0;
  return c;
}

}  // namespace

std::size_t MapServer::sessionBytes(const SessionSpec &s) {
  const std::size_t v = static_cast<std::size_t>(2 * s.half_map_size_i.x() + 1) *
                        static_cast<std::size_t>(2 * s.half_map_size_i.y() + 1) *
                        static_cast<std::size_t>(2 * s.half_map_size_i.z() + 1);
  return v * sizeof(float) + kFetchReserve;
}

MapServer::MapServer(const ServerLimits &limits) : lim_(limits) {
  if (lim_.max_sessions < 1 || lim_.max_points_per_scan < 1 ||
      lim_.max_points_per_tick < lim_.max_points_per_scan || lim_.max_grids_per_tick < 1 ||
      lim_.max_tiles_per_tick < 1 || lim_.max_segments_per_ray < 1) {
    throw std::invalid_argument("cublox::MapServer: invalid limits");
  }
  BatchLimits bl;
  bl.max_grids = lim_.max_grids_per_tick;
  bl.max_points = lim_.max_points_per_tick;
  bl.max_tiles = lim_.max_tiles_per_tick;
  bl.max_segments_per_ray = lim_.max_segments_per_ray;
  bl.max_near_steps = 12;
  bl.min_chunk = 512;
  if (lim_.point_arena_bytes < 3ull * sizeof(float) * lim_.max_points_per_scan + 256) {
    throw std::invalid_argument("cublox::MapServer: point arena below one largest scan");
  }
  batch_ = std::make_unique<GridBatch>(bl);
  batch_->setGraphMode(lim_.graph);
  if (lim_.pipelined) batch_->setPipelined(true);
  TuneSetting ts = batch_->setting();
  ts.chunk = std::max(lim_.chunk, bl.min_chunk);
  batch_->setSetting(ts);
  try {
    MS_CHECK(cublox_malloc(reinterpret_cast<void **>(&d_arena_), lim_.point_arena_bytes));
    const std::size_t stage_bytes =
        3 * sizeof(float) * static_cast<std::size_t>(lim_.max_points_per_scan);
    for (int k = 0; k < kUploadStages; ++k) {
      UploadStage st;
      stages_.push_back(st);
      MS_CHECK(cublox_host_alloc(&stages_.back().host, stage_bytes));
      MS_CHECK(cublox_stream_create(&stages_.back().stream));
      free_stages_.push_back(k);
    }
    // Fetch output arenas sized up front (they still grow on demand): a
    // first growth inside a tick's path frees and allocates device and
    // pinned memory, a stall of milliseconds.
    for (auto &a : farena_) {
      growFetchArena(a, static_cast<std::size_t>(lim_.fetch_samples_per_tick) / 2, 512);
    }
  } catch (...) {
    for (auto &a : farena_) growFetchArena(a, 0, 0);
    if (d_arena_) cublox_free(d_arena_);
    d_arena_ = nullptr;
    for (auto &st : stages_) {
      if (st.host) cublox_host_free(st.host);
      if (st.stream) cublox_stream_destroy(st.stream);
    }
    stages_.clear();
    throw;
  }
  arena_free_[0] = lim_.point_arena_bytes;
  stats_.workspace_bytes = lim_.point_arena_bytes;
  t0_ = std::chrono::steady_clock::now();
  worker_ = std::thread([this] { run(); });
}

MapServer::~MapServer() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (worker_.joinable()) worker_.join();
  {
    std::lock_guard<std::mutex> lk(arena_mu_);
    arena_closed_ = true;
  }
  arena_cv_.notify_all();
  batch_.reset();
  for (auto &a : farena_) growFetchArena(a, 0, 0);  // frees
  if (d_arena_) cublox_free(d_arena_);
  for (auto &st : stages_) {
    if (st.host) cublox_host_free(st.host);
    if (st.stream) cublox_stream_destroy(st.stream);
  }
  sessions_.clear();  // frees the grids
}

double MapServer::nowMs() const {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_)
      .count();
}

// First fit over the free list; waits while no range fits.
std::size_t MapServer::allocPoints(std::size_t bytes) {
  std::unique_lock<std::mutex> lk(arena_mu_);
  while (true) {
    if (arena_closed_) throw std::runtime_error("cublox::MapServer: shutting down");
    for (auto it = arena_free_.begin(); it != arena_free_.end(); ++it) {
      if (it->second < bytes) continue;
      const std::size_t off = it->first, size = it->second;
      arena_free_.erase(it);
      if (size > bytes) arena_free_[off + bytes] = size - bytes;
      return off;
    }
    arena_cv_.wait(lk);
  }
}

void MapServer::releasePoints(std::size_t off, std::size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(arena_mu_);
    auto next = arena_free_.lower_bound(off);
    if (next != arena_free_.end() && off + bytes == next->first) {  // merge right
      bytes += next->second;
      next = arena_free_.erase(next);
    }
    if (next != arena_free_.begin()) {  // merge left
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        bytes = 0;
      }
    }
    if (bytes) arena_free_[off] = bytes;
  }
  arena_cv_.notify_all();
}

int MapServer::openSession(const SessionSpec &spec) {
  const cublox_raycast_cfg c = shapeCfg(spec);
  Session s;
  s.spec = spec;
  s.bytes = sessionBytes(spec);
  s.tiles = cublox_integrate_tile_count(&c);
  s.seg_bound = cublox_integrate_segment_bound(&c);
  if (spec.resolution <= 0.0f || spec.half_map_size_i.minCoeff() < 0) {
    throw std::invalid_argument("cublox::MapServer: invalid session shape");
  }
  if (s.seg_bound > lim_.max_segments_per_ray || s.tiles > lim_.max_tiles_per_tick ||
      s.bytes > lim_.hbm_budget_bytes) {
    throw std::invalid_argument(
        "cublox::MapServer: session shape above the server limits (segments per ray, "
        "tiles per tick or HBM budget)");
  }
  std::lock_guard<std::mutex> lk(mu_);
  if (worker_error_) std::rethrow_exception(worker_error_);
  const int sid = next_sid_++;
  s.in_waiting = true;
  sessions_.emplace(sid, std::move(s));
  waiting_.push_back(sid);
  ++gen_;
  cv_work_.notify_all();
  return sid;
}

void MapServer::closeSession(int sid) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sessions_.find(sid);
  if (it == sessions_.end()) throw std::invalid_argument("cublox::MapServer: no such session");
  it->second.closing = true;
  ++gen_;
  cv_work_.notify_all();
}

void MapServer::evictSession(int sid) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sessions_.find(sid);
  if (it == sessions_.end()) throw std::invalid_argument("cublox::MapServer: no such session");
  it->second.evict_requested = true;
  ++gen_;
  cv_work_.notify_all();
}

std::vector<uint8_t> MapServer::evictedSnapshot(int sid) const {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sessions_.find(sid);
  if (it == sessions_.end()) throw std::invalid_argument("cublox::MapServer: no such session");
  return it->second.snapshot;
}

bool MapServer::admitted(int sid) const {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sessions_.find(sid);
  return it != sessions_.end() && it->second.admitted;
}

long long MapServer::submitScan(int sid, const float *xs, const float *ys, const float *zs,
                                int n, const float T[12], const Eigen::Vector3f &base) {
  if (n < 0 || n > lim_.max_points_per_scan) {
    throw std::invalid_argument("cublox::MapServer: scan size outside [0, max_points_per_scan]");
  }
  Request r;
  r.submit_ms = nowMs();
  r.kind = Kind::SCAN;
  r.n = n;
  std::copy(T, T + 12, r.T.begin());
  r.pos = base;
  {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = sessions_.find(sid);
    if (it == sessions_.end() || it->second.closing) {
      throw std::invalid_argument("cublox::MapServer: scan for a closed or unknown session");
    }
  }
  if (n > 0) {
    r.arena_bytes = (3 * sizeof(float) * static_cast<std::size_t>(n) + 255) & ~std::size_t{255};
    r.arena_off = allocPoints(r.arena_bytes);
    r.d_pts = reinterpret_cast<float *>(d_arena_ + r.arena_off);
    try {
      // The three columns go into a free pinned stage (one host copy each)
      // and on to the arena in one transfer; a pageable copy runs at a few
      // GB/s. Stages let several client threads upload at once.
      int k = -1;
      {
        std::unique_lock<std::mutex> lk(stage_mu_);
        stage_cv_.wait(lk, [&] { return !free_stages_.empty(); });
        k = free_stages_.back();
        free_stages_.pop_back();
      }
      struct Release {
        MapServer *s;
        int k;
        ~Release() {
          {
            std::lock_guard<std::mutex> lk(s->stage_mu_);
            s->free_stages_.push_back(k);
          }
          s->stage_cv_.notify_one();
        }
      } release{this, k};
      const std::size_t b = sizeof(float) * static_cast<std::size_t>(n);
      auto *stage = static_cast<uint8_t *>(stages_[static_cast<std::size_t>(k)].host);
      std::memcpy(stage, xs, b);
      std::memcpy(stage + b, ys, b);
      std::memcpy(stage + 2 * b, zs, b);
      void *us = stages_[static_cast<std::size_t>(k)].stream;
      MS_CHECK(cublox_memcpy_h2d_async(r.d_pts, stage, 3 * b, us));
      MS_CHECK(cublox_stream_synchronize(us));
    } catch (...) {
      releasePoints(r.arena_off, r.arena_bytes);
      throw;
    }
  }
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sessions_.find(sid);
  if (worker_error_ || it == sessions_.end() || it->second.closing) {
    if (r.arena_bytes) releasePoints(r.arena_off, r.arena_bytes);
    if (worker_error_) std::rethrow_exception(worker_error_);
    throw std::invalid_argument("cublox::MapServer: scan for a closed or unknown session");
  }
  r.id = next_id_++;
  it->second.queue.push_back(std::move(r));
  if (!it->second.admitted && !it->second.in_waiting) {  // evicted: readmit
    it->second.in_waiting = true;
    waiting_.push_back(sid);
  }
  ++gen_;
  cv_work_.notify_all();
  return next_id_ - 1;
}

long long MapServer::submitFetch(int sid, const Eigen::Vector3f &center, float radius,
                                 std::optional<OccupancyGrid::VoxelState> filter,
                                 int max_results) {
  Request r;
  r.kind = Kind::FETCH;
  r.pos = center;
  r.radius = radius;
  r.filter = filter;
  r.max_results = max_results;
  std::lock_guard<std::mutex> lk(mu_);
  if (worker_error_) std::rethrow_exception(worker_error_);
  auto it = sessions_.find(sid);
  if (it == sessions_.end() || it->second.closing) {
    throw std::invalid_argument("cublox::MapServer: fetch for a closed or unknown session");
  }
  r.id = next_id_++;
  r.submit_ms = nowMs();
  it->second.queue.push_back(std::move(r));
  if (!it->second.admitted && !it->second.in_waiting) {  // evicted: readmit
    it->second.in_waiting = true;
    waiting_.push_back(sid);
  }
  ++gen_;
  cv_work_.notify_all();
  return next_id_ - 1;
}

bool MapServer::quiescent() const {
  for (const auto &[sid, s] : sessions_) {
    if (s.busy > 0) return false;
    if (s.admitted && !s.queue.empty()) return false;
    if (s.closing && s.admitted) return false;  // not yet retired
  }
  for (const auto &[sid, s] : sessions_) {
    if (s.evict_requested && s.admitted && !s.closing) return false;
  }
  // The worker would admit someone (the same walk as admitWaiting).
  for (int sid : waiting_) {
    const Session &s = sessions_.at(sid);
    if (admittable(s)) return false;
    if (lim_.evict_idle && !s.queue.empty() && idleVictim()) return false;
    if (!(lim_.evict_idle && s.queue.empty())) break;  // FIFO: blocked here
  }
  return true;
}

void MapServer::pause() {
  std::lock_guard<std::mutex> lk(mu_);
  paused_ = true;
}

void MapServer::resume() {
  std::lock_guard<std::mutex> lk(mu_);
  paused_ = false;
  ++gen_;
  cv_work_.notify_all();
}

void MapServer::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  if (paused_) throw std::logic_error("cublox::MapServer: drain() while paused");
  ++gen_;
  cv_work_.notify_all();
  cv_idle_.wait(lk, [&] { return worker_error_ || (quiescent() && !worker_busy_); });
  if (worker_error_) std::rethrow_exception(worker_error_);
}

std::vector<MapServer::Completion> MapServer::takeCompletions() {
  std::lock_guard<std::mutex> lk(mu_);
  std::vector<Completion> out;
  out.swap(done_);
  return out;
}

void MapServer::keepFetchResults(bool on) {
  std::lock_guard<std::mutex> lk(mu_);
  keep_fetch_ = on;
}

std::vector<OccupancyGrid::OccupancySample> MapServer::takeFetchResult(long long id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = fetch_results_.find(id);
  if (it == fetch_results_.end()) return {};
  auto v = std::move(it->second);
  fetch_results_.erase(it);
  return v;
}

const OccupancyGrid &MapServer::grid(int sid) const {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = sessions_.find(sid);
  if (it == sessions_.end() || !it->second.grid) {
    throw std::invalid_argument("cublox::MapServer: session has no resident grid");
  }
  return *it->second.grid;
}

MapServer::Stats MapServer::stats() const {
  std::lock_guard<std::mutex> lk(mu_);
  Stats s = stats_;
  s.resident_bytes = resident_;
  s.admitted = admitted_count_;
  s.waiting = static_cast<int>(waiting_.size());
  s.snapshot_bytes = 0;
  for (const auto &kv : sessions_) s.snapshot_bytes += kv.second.snapshot.size();
  return s;
}

// ---------------------------------------------------------------------------
// worker
// ---------------------------------------------------------------------------

bool MapServer::admittable(const Session &s) const {
  return admitted_count_ < lim_.max_sessions && resident_ + s.bytes <= lim_.hbm_budget_bytes;
}

void MapServer::retireClosed() {
  for (auto it = sessions_.begin(); it != sessions_.end();) {
    Session &s = it->second;
    if (s.closing && s.busy == 0 && (s.queue.empty() || !s.admitted)) {
      for (const Request &r : s.queue) {  // dropped with a never-admitted session
        if (r.arena_bytes) releasePoints(r.arena_off, r.arena_bytes);
      }
      if (s.admitted) {
        resident_ -= s.bytes;
        --admitted_count_;
      }
      if (s.in_waiting) {
        waiting_.erase(std::remove(waiting_.begin(), waiting_.end(), it->first),
                       waiting_.end());
      }
      it = sessions_.erase(it);  // frees the grid
    } else {
      ++it;
    }
  }
}

void MapServer::admitWaiting() {
  for (auto it = waiting_.begin(); it != waiting_.end();) {
    Session &s = sessions_.at(*it);
    if (!admittable(s)) {
      // FIFO: later sessions wait behind it. With idle eviction, an idle
      // session makes room for one with queued requests, and a waiting
      // session with nothing queued does not hold up the ones behind it.
      if (lim_.evict_idle && !s.queue.empty() && evictIdle()) continue;
      if (lim_.evict_idle && s.queue.empty()) {
        ++it;
        continue;
      }
      break;
    }
    const SessionSpec &p = s.spec;
    s.grid = std::make_unique<OccupancyGrid>(p.half_map_size_i, p.resolution,
                                             p.origin_at_center, p.recenter_threshold);
    s.grid->setMaxRaycastRange(p.max_range);
    s.grid->setLogOddsParams(p.log_odds[0], p.log_odds[1], p.log_odds[2], p.log_odds[3],
                             p.log_odds[4], p.log_odds[5]);
    if (!s.snapshot.empty()) {  // an evicted session comes back as it left
      s.grid->loadSnapshot(s.snapshot.data(), s.snapshot.size());
      std::vector<uint8_t>().swap(s.snapshot);
      ++stats_.restores;
    }
    s.admitted = true;
    s.in_waiting = false;
    resident_ += s.bytes;
    ++admitted_count_;
    it = waiting_.erase(it);
  }
}

// An idle session: resident, with nothing queued, in flight or closing.
bool MapServer::idleVictim() const {
  for (const auto &kv : sessions_) {
    const Session &s = kv.second;
    if (s.admitted && !s.closing && s.busy == 0 && s.queue.empty()) return true;
  }
  return false;
}

void MapServer::evict(int sid, Session &s) {
  s.snapshot = s.grid->saveSnapshot();  // nothing of it is in flight
  s.grid.reset();
  s.admitted = false;
  s.evict_requested = false;
  resident_ -= s.bytes;
  --admitted_count_;
  ++stats_.evictions;
  if (!s.queue.empty() && !s.in_waiting) {
    s.in_waiting = true;
    waiting_.push_back(sid);
  }
}

// The least recently used idle session.
bool MapServer::evictIdle() {
  int best = -1;
  long long best_tick = 0;
  for (auto &[sid, s] : sessions_) {
    if (!(s.admitted && !s.closing && s.busy == 0 && s.queue.empty())) continue;
    if (best < 0 || s.last_tick < best_tick) {
      best = sid;
      best_tick = s.last_tick;
    }
  }
  if (best < 0) return false;
  evict(best, sessions_.at(best));
  return true;
}

void MapServer::evictRequested() {
  for (auto &[sid, s] : sessions_) {
    if (s.evict_requested && s.admitted && !s.closing && s.busy == 0) evict(sid, s);
    if (s.evict_requested && (s.closing || !s.admitted)) s.evict_requested = false;
  }
}

// Whether the admitted sessions' queued scans hold at least half a tick of
// points (lock held): requests are backing up, not just arriving.
bool MapServer::backlog() const {
  long long points = 0;
  const long long half = lim_.max_points_per_tick / 2;
  for (const auto &kv : sessions_) {
    const Session &s = kv.second;
    if (!s.admitted || s.closing) continue;
    for (const Request &r : s.queue) {
      if (r.kind != Kind::SCAN) continue;
      points += r.n;
      if (points >= half) return true;
    }
  }
  return false;
}

// Takes the tick's requests: per admitted session (round-robin from rr_),
// a leading run of fetches, or one scan and the fetches right after it,
// while the scan fits the tick's caps.
bool MapServer::formTick(Tick &t) {
  long long points = 0, slots = 0;
  int tiles = 0, grids = 0;
  const long long slot_cap =
      static_cast<long long>(lim_.max_points_per_tick) * lim_.max_segments_per_ray;
  std::vector<int> ids;
  ids.reserve(sessions_.size());
  for (const auto &kv : sessions_) ids.push_back(kv.first);
  if (ids.empty()) return false;
  const std::size_t start =
      static_cast<std::size_t>(std::lower_bound(ids.begin(), ids.end(), rr_) - ids.begin());
  for (std::size_t k = 0; k < ids.size(); ++k) {
    const int sid = ids[(start + k) % ids.size()];
    Session &s = sessions_.at(sid);
    if (!s.admitted || s.queue.empty()) continue;
    if (s.queue.front().kind == Kind::SCAN) {
      const Request &r = s.queue.front();
      const long long need = static_cast<long long>(r.n) * s.seg_bound;
      if (grids + 1 > lim_.max_grids_per_tick || points + r.n > lim_.max_points_per_tick ||
          tiles + s.tiles > lim_.max_tiles_per_tick || slots + need > slot_cap) {
        continue;
      }
      grids += 1;
      points += r.n;
      tiles += s.tiles;
      slots += need;
      t.scans.emplace_back(sid, std::move(s.queue.front()));
      s.queue.pop_front();
      ++s.busy;
      s.last_tick = ticks_;
    }
    while (!s.queue.empty() && s.queue.front().kind == Kind::FETCH) {
      t.fetches.emplace_back(sid, std::move(s.queue.front()));
      s.queue.pop_front();
      ++s.busy;
      s.last_tick = ticks_;
    }
  }
  if (t.scans.empty() && t.fetches.empty()) return false;
  rr_ = ids[(start + 1) % ids.size()];
  return true;
}

void MapServer::submitTick(Tick &t) {
  std::vector<GridScan> scans;
  std::vector<std::pair<OccupancyGrid *, Eigen::Vector3f>> rc;
  std::vector<OccupancyGrid *> grids;
  {
    std::lock_guard<std::mutex> lk(mu_);  // a busy session's grid stays put
    for (auto &[sid, r] : t.scans) grids.push_back(sessions_.at(sid).grid.get());
  }
  std::size_t k = 0;
  for (auto &[sid, r] : t.scans) {
    OccupancyGrid *g = grids[k++];
    GridScan s;
    s.grid = g;
    s.d_x = r.d_pts;
    s.d_y = r.d_pts ? r.d_pts + r.n : nullptr;
    s.d_z = r.d_pts ? r.d_pts + 2 * static_cast<std::size_t>(r.n) : nullptr;
    s.n = r.n;
    s.sensor_frame = true;
    s.T = r.T;
    scans.push_back(s);
    rc.emplace_back(g, r.pos);
  }
  batch_->submit(scans, rc);
}

// Grows (samples > 0) or frees (0, 0) the fetch output arena. Called
// between ticks only, when no device work uses it.
void MapServer::growFetchArena(FetchArena &a, std::size_t samples, std::size_t fetches) {
  if (samples <= a.cap && fetches <= a.cnt_cap && (samples > 0 || fetches > 0)) return;
  if (a.d_pos) cublox_free(a.d_pos);
  if (a.d_st) cublox_free(a.d_st);
  if (a.d_cnt) cublox_free(a.d_cnt);
  if (a.h_pos) cublox_host_free(a.h_pos);
  if (a.h_st) cublox_host_free(a.h_st);
  if (a.h_cnt) cublox_host_free(a.h_cnt);
  a = FetchArena{};
  if (samples == 0 && fetches == 0) return;
  const std::size_t s = std::max<std::size_t>(samples, 1 << 16) * 2;  // headroom
  const std::size_t f = std::max<std::size_t>(fetches, 256) * 2;
  MS_CHECK(cublox_malloc(reinterpret_cast<void **>(&a.d_pos), 3 * sizeof(float) * s));
  MS_CHECK(cublox_malloc(reinterpret_cast<void **>(&a.d_st), s));
  MS_CHECK(cublox_malloc(reinterpret_cast<void **>(&a.d_cnt), sizeof(unsigned) * f));
  MS_CHECK(cublox_host_alloc(reinterpret_cast<void **>(&a.h_pos), 3 * sizeof(float) * s));
  MS_CHECK(cublox_host_alloc(reinterpret_cast<void **>(&a.h_st), s));
  MS_CHECK(cublox_host_alloc(reinterpret_cast<void **>(&a.h_cnt), sizeof(unsigned) * f));
  a.cap = s;
  a.cnt_cap = f;
}

// Plans the tick's fetches against the grids as the tick leaves them (the
// host plan of OccupancyGrid::fetchOccupancyAround) and enqueues them on the
// batch stream after the tick's update and recenter, with their outputs
// copied to pinned memory: no host wait per fetch.
void MapServer::enqueueFetches(Tick &t) {
  if (t.fetches.empty()) return;
  std::vector<const OccupancyGrid *> grids;
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto &[sid, r] : t.fetches) grids.push_back(sessions_.at(sid).grid.get());
  }
  std::vector<cublox_fetch_cfg> cfgs(t.fetches.size());
  t.jobs.assign(t.fetches.size(), FetchJob{});
  std::size_t total = 0;
  for (std::size_t k = 0; k < t.fetches.size(); ++k) {
    const Request &r = t.fetches[k].second;
    const OccupancyGrid &g = *grids[k];
    FetchJob &j = t.jobs[k];
    if (g.deviceOccupancy() == nullptr || r.radius <= 0.0f) continue;
    long long est = 0;
    unsigned output_cap = 0;
    const auto &c = g.config_;
    j.live = cublox_fetch_plan(
                 cublox_int3{g.origin_i_.x(), g.origin_i_.y(), g.origin_i_.z()},
                 cublox_int3{c.map_size_i.x(), c.map_size_i.y(), c.map_size_i.z()},
                 cublox_int3{c.half_map_size_i.x(), c.half_map_size_i.y(), c.half_map_size_i.z()},
                 c.resolution, c.resolution_inv, c.origin_at_center ? 1 : 0,
                 g.getLogOddsFree(), g.getLogOddsOccupied(), 0.0f,
                 cublox_float3{r.pos.x(), r.pos.y(), r.pos.z()}, r.radius,
                 r.filter ? static_cast<int>(*r.filter) : -1, r.max_results, &cfgs[k], &est,
                 &output_cap) != 0;
    if (!j.live) continue;
    j.cap = static_cast<unsigned>(std::min<long long>(output_cap, est));
    j.limit = r.max_results > 0 ? std::min(static_cast<unsigned>(r.max_results), output_cap)
                                : output_cap;
    j.off = total;
    total += j.cap;
  }
  FetchArena &a = farena_[t.buf];
  growFetchArena(a, total, t.fetches.size());
  void *s = batch_->stream();
  MS_CHECK(cublox_fill_zero_async(a.d_cnt, t.fetches.size(), s));
  for (std::size_t k = 0; k < t.fetches.size(); ++k) {
    const FetchJob &j = t.jobs[k];
    if (!j.live || j.cap == 0) continue;
    MS_CHECK(cublox_launch_fetch_occupancy_around(&cfgs[k], grids[k]->deviceOccupancy(),
                                                  a.d_cnt + k, a.d_pos + 3 * j.off,
                                                  a.d_st + j.off, j.cap, j.limit, s));
  }
  MS_CHECK(cublox_memcpy_d2h_async(a.h_cnt, a.d_cnt, sizeof(unsigned) * t.fetches.size(), s));
  if (total > 0) {
    MS_CHECK(cublox_memcpy_d2h_async(a.h_pos, a.d_pos, 3 * sizeof(float) * total, s));
    MS_CHECK(cublox_memcpy_d2h_async(a.h_st, a.d_st, total, s));
  }
}

// Blocks until the tick's device work (update, recenter, fetches) is done.
// Ticks are waited for in submission order, so a tick with scans is the
// batch's oldest in flight; its fetches follow it on its stream.
void MapServer::waitTick(Tick &t) {
  if (!t.scans.empty()) {
    batch_->waitOldest();
    t.device_ms = batch_->lastUpdateMs() + batch_->lastRecenterMs();
  }
  MS_CHECK(cublox_stream_synchronize(t.stream));
  t.done_ms = nowMs();
}

// Records a finished tick's completions and fetch results and frees its
// points (runs while the next tick is on the device).
void MapServer::finishTick(Tick &t) {
  const double done = t.done_ms;
  const FetchArena &a = farena_[t.buf];
  std::vector<Completion> out;
  for (auto &[sid, r] : t.scans) {
    out.push_back(Completion{r.id, sid, Kind::SCAN, r.submit_ms, t.start_ms, done, t.index,
                             static_cast<long long>(r.n)});
  }
  bool keep = false;
  {
    std::lock_guard<std::mutex> lk(mu_);
    keep = keep_fetch_;
  }
  std::vector<std::pair<long long, std::vector<OccupancyGrid::OccupancySample>>> kept;
  for (std::size_t k = 0; k < t.fetches.size(); ++k) {
    const Request &r = t.fetches[k].second;
    const FetchJob &j = t.jobs[k];
    // OccupancyGrid.cpp: n = min(count, output cap, max_results, buffer).
    const unsigned n = j.live ? std::min({a.h_cnt[k], j.limit, j.cap}) : 0u;
    out.push_back(Completion{r.id, t.fetches[k].first, Kind::FETCH, r.submit_ms, t.start_ms,
                             done, t.index, static_cast<long long>(n)});
    if (!keep) continue;
    std::vector<OccupancyGrid::OccupancySample> v;
    v.reserve(n);
    for (unsigned i = 0; i < n; ++i) {
      const float *p = a.h_pos + 3 * (j.off + i);
      v.push_back(OccupancyGrid::OccupancySample{
          Eigen::Vector3f(p[0], p[1], p[2]),
          static_cast<OccupancyGrid::VoxelState>(a.h_st[j.off + i])});
    }
    kept.emplace_back(r.id, std::move(v));
  }
  for (auto &[sid, r] : t.scans) {
    if (r.arena_bytes) releasePoints(r.arena_off, r.arena_bytes);
  }
  std::lock_guard<std::mutex> lk(mu_);
  for (auto &[sid, r] : t.scans) {
    --sessions_.at(sid).busy;
    stats_.points += r.n;
  }
  for (auto &[sid, r] : t.fetches) --sessions_.at(sid).busy;
  stats_.scans += static_cast<long long>(t.scans.size());
  stats_.fetches += static_cast<long long>(t.fetches.size());
  if (!t.scans.empty()) {
    ++stats_.ticks;
    stats_.device_ms += t.device_ms;
  }
  for (auto &kv : kept) fetch_results_[kv.first] = std::move(kv.second);
  done_.insert(done_.end(), out.begin(), out.end());
}

// The next tick is formed the moment the device finishes the current one,
// from everything queued by then; the finished tick's completions are then
// recorded while the new tick runs.
void MapServer::run() {
  // Submitted ticks not yet finished, oldest first. One at a time; when
  // pipelined, two while requests back up (a formed tick left at least
  // half a tick of points queued), so a lighter load keeps the latency of
  // single ticks. A finished
  // tick's host work (completions, fetch results, freeing its points) runs
  // while the next ticks are on the device.
  std::deque<Tick> flight;
  const int arenas = lim_.pipelined ? 3 : 2;
  std::size_t depth = 1;
  int buf = 0;
  long long seen = -1;
  try {
    while (true) {
      while (flight.size() > depth) {  // the backlog cleared: back to one
        waitTick(flight.front());
        finishTick(flight.front());
        flight.pop_front();
      }
      const bool full = flight.size() >= depth;
      if (full) waitTick(flight.front());
      Tick cur;
      bool have_cur = false;
      {
        std::unique_lock<std::mutex> lk(mu_);
        retireClosed();
        evictRequested();
        admitWaiting();
        if (flight.empty()) {
          worker_busy_ = false;
          cv_idle_.notify_all();
          cv_work_.wait(lk, [&] { return stop_ || (gen_ != seen && !paused_); });
          seen = gen_;
          retireClosed();
          evictRequested();
          admitWaiting();
          worker_busy_ = true;
        }
        have_cur = !paused_ && formTick(cur);
        if (have_cur) cur.index = ticks_++;
        if (!have_cur && flight.empty() && stop_) break;
        depth = lim_.pipelined && have_cur && backlog() ? 2 : 1;
      }
      if (have_cur) {
        cur.buf = buf;
        buf = (buf + 1) % arenas;
        cur.start_ms = nowMs();
        if (!cur.scans.empty()) submitTick(cur);
        cur.stream = batch_->stream();
        enqueueFetches(cur);
        cur.launched_ms = nowMs();
        if (full) {
          std::lock_guard<std::mutex> lk(mu_);
          stats_.launch_ms += cur.launched_ms - flight.front().done_ms;
        }
      }
      if (!full && !have_cur && !flight.empty()) waitTick(flight.front());
      if (full || (!have_cur && !flight.empty())) {
        finishTick(flight.front());
        flight.pop_front();
      }
      if (have_cur) flight.push_back(std::move(cur));
      {
        std::lock_guard<std::mutex> lk(mu_);
        cv_idle_.notify_all();
      }
    }
  } catch (...) {
    std::lock_guard<std::mutex> lk(mu_);
    worker_error_ = std::current_exception();
    worker_busy_ = false;
    cv_idle_.notify_all();
    return;
  }
  std::lock_guard<std::mutex> lk(mu_);
  worker_busy_ = false;
  cv_idle_.notify_all();
}

}  // namespace cublox
