// Flat C binding of cublox::MapServer (include/cublox/ms_capi.h). Every
// entry point catches C++ exceptions and returns a status; the message of
// the last failure on the calling thread is in cublox_og_last_error().
#include <cublox/MapServer.hpp>
#include <cublox/capi.h>
#include <cublox/ms_capi.h>
#include <cublox/og_capi.h>
#include <cublox/snapshot.h>

#include <algorithm>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace cublox_ffi {
void set_last_error(const char *msg);  // og_capi.cpp
}

namespace {

template <typename F> int guarded(F &&f) {
  try {
    f();
    return CUBLOX_OK;
  } catch (const std::exception &e) {
    cublox_ffi::set_last_error(e.what());
    return CUBLOX_ERR_RUNTIME;
  } catch (...) {
    cublox_ffi::set_last_error("unknown C++ exception");
    return CUBLOX_ERR_RUNTIME;
  }
}

using MS = cublox::MapServer;

struct Handle {
  std::unique_ptr<MS> server;
  std::vector<MS::Completion> pending;  // taken from the server, not yet read
};

cublox::SessionSpec spec_of(const int half[3], float resolution, int origin_at_center,
                            int has_threshold, double threshold, float max_range,
                            const float log_odds[6]) {
  cublox::SessionSpec s;
  s.half_map_size_i = Eigen::Vector3i(half[0], half[1], half[2]);
  s.resolution = resolution;
  s.origin_at_center = origin_at_center != 0;
  if (has_threshold) s.recenter_threshold = threshold;
  s.max_range = max_range;
  if (log_odds) std::copy(log_odds, log_odds + 6, s.log_odds.begin());
  return s;
}

}  // namespace

#define H reinterpret_cast<Handle *>(h)

namespace cublox_ffi {
cublox::MapServer &server_of(cublox_ms *h) { return *H->server; }
}  // namespace cublox_ffi

extern "C" {

cublox_ms *cublox_ms_create(const long long limits[12]) {
  cublox_ms *out = nullptr;
  guarded([&] {
    cublox::ServerLimits l;
    l.hbm_budget_bytes = static_cast<std::size_t>(limits[0]);
    l.max_sessions = static_cast<int>(limits[1]);
    l.max_points_per_scan = static_cast<int>(limits[2]);
    l.max_points_per_tick = static_cast<int>(limits[3]);
    l.max_grids_per_tick = static_cast<int>(limits[4]);
    l.max_tiles_per_tick = static_cast<int>(limits[5]);
    l.max_segments_per_ray = static_cast<int>(limits[6]);
    l.graph = limits[7] != 0;
    if (limits[8] > 0) l.point_arena_bytes = static_cast<std::size_t>(limits[8]);
    l.evict_idle = limits[9] != 0;
    l.pipelined = limits[10] != 0;
    if (limits[11] > 0) l.chunk = static_cast<int>(limits[11]);
    auto *hd = new Handle;
    try {
      hd->server = std::make_unique<MS>(l);
    } catch (...) {
      delete hd;
      throw;
    }
    out = reinterpret_cast<cublox_ms *>(hd);
  });
  return out;
}

void cublox_ms_destroy(cublox_ms *h) { delete H; }

long long cublox_ms_session_bytes(const int half[3]) {
  cublox::SessionSpec s;
  s.half_map_size_i = Eigen::Vector3i(half[0], half[1], half[2]);
  return static_cast<long long>(MS::sessionBytes(s));
}

int cublox_ms_open(cublox_ms *h, const int half[3], float resolution, int origin_at_center,
                   int has_threshold, double threshold, float max_range,
                   const float log_odds[6], int *sid) {
  return guarded([&] {
    *sid = H->server->openSession(spec_of(half, resolution, origin_at_center, has_threshold,
                                          threshold, max_range, log_odds));
  });
}

int cublox_ms_close(cublox_ms *h, int sid) {
  return guarded([&] { H->server->closeSession(sid); });
}

int cublox_ms_evict(cublox_ms *h, int sid) {
  return guarded([&] { H->server->evictSession(sid); });
}

int cublox_ms_admitted(cublox_ms *h, int sid, int *out) {
  return guarded([&] { *out = H->server->admitted(sid) ? 1 : 0; });
}

int cublox_ms_submit_scan(cublox_ms *h, int sid, const float *xs, const float *ys,
                          const float *zs, int n, const float T[12], const float base[3],
                          long long *id) {
  return guarded([&] {
    *id = H->server->submitScan(sid, xs, ys, zs, n, T,
                                Eigen::Vector3f(base[0], base[1], base[2]));
  });
}

int cublox_ms_submit_fetch(cublox_ms *h, int sid, const float center[3], float radius,
                           int state_filter, int max_results, long long *id) {
  return guarded([&] {
    std::optional<cublox::OccupancyGrid::VoxelState> f;
    if (state_filter >= 0) f = static_cast<cublox::OccupancyGrid::VoxelState>(state_filter);
    *id = H->server->submitFetch(sid, Eigen::Vector3f(center[0], center[1], center[2]),
                                 radius, f, max_results);
  });
}

int cublox_ms_drain(cublox_ms *h) {
  return guarded([&] { H->server->drain(); });
}

int cublox_ms_pause(cublox_ms *h) {
  return guarded([&] { H->server->pause(); });
}

int cublox_ms_resume(cublox_ms *h) {
  return guarded([&] { H->server->resume(); });
}

int cublox_ms_take_completions(cublox_ms *h, long long *ids, int *sessions, int *kinds,
                               double *times, long long *ticks, long long *results,
                               int cap, int *n_out) {
  return guarded([&] {
    auto fresh = H->server->takeCompletions();
    auto &p = H->pending;
    p.insert(p.end(), fresh.begin(), fresh.end());
    const int n = std::min<int>(cap, static_cast<int>(p.size()));
    for (int k = 0; k < n; ++k) {
      const auto &c = p[static_cast<size_t>(k)];
      ids[k] = c.id;
      sessions[k] = c.session;
      kinds[k] = static_cast<int>(c.kind);
      times[3 * k] = c.submit_ms;
      times[3 * k + 1] = c.start_ms;
      times[3 * k + 2] = c.done_ms;
      ticks[k] = c.tick;
      results[k] = c.result;
    }
    p.erase(p.begin(), p.begin() + n);
    *n_out = n;
  });
}

int cublox_ms_keep_fetch_results(cublox_ms *h, int on) {
  return guarded([&] { H->server->keepFetchResults(on != 0); });
}

int cublox_ms_fetch_result(cublox_ms *h, long long id, float *pos, unsigned char *state,
                           int cap, int *n) {
  return guarded([&] {
    const auto v = H->server->takeFetchResult(id);
    const int m = std::min<int>(cap, static_cast<int>(v.size()));
    for (int k = 0; k < m; ++k) {
      const auto &s = v[static_cast<size_t>(k)];
      pos[3 * k] = s.position.x();
      pos[3 * k + 1] = s.position.y();
      pos[3 * k + 2] = s.position.z();
      state[k] = static_cast<unsigned char>(s.state);
    }
    *n = static_cast<int>(v.size());
  });
}

// An evicted session's map, read from its snapshot (cublox/snapshot.h).
static cublox_snapshot_header evictedMap(const cublox::MapServer &srv, int sid,
                                         float *dst) {
  const std::vector<uint8_t> b = srv.evictedSnapshot(sid);
  if (b.size() < sizeof(cublox_snapshot_header)) {
    throw std::invalid_argument("cublox::MapServer: session has no map yet");
  }
  cublox_snapshot_header hd;
  std::memcpy(&hd, b.data(), sizeof(hd));
  if (dst) {
    const std::size_t V = static_cast<std::size_t>(std::max(hd.voxel_num, 0));
    const uint8_t *p = b.data() + sizeof(hd);
    if (hd.encoding == CUBLOX_SNAPSHOT_DENSE) {
      if (hd.payload_bytes == V * 4) std::memcpy(dst, p, V * 4);
      else std::fill(dst, dst + V, 0.0f);
    } else {
      uint32_t n = 0;
      std::memcpy(&n, p, 4);
      std::fill(dst, dst + V, 0.0f);
      for (uint32_t k = 0; k < n; ++k) {
        int32_t id;
        float v;
        std::memcpy(&id, p + 8 + 4ull * k, 4);
        std::memcpy(&v, p + 8 + 4ull * n + 4ull * k, 4);
        if (id >= 0 && static_cast<std::size_t>(id) < V) dst[id] = v;
      }
    }
  }
  return hd;
}

int cublox_ms_copy_occupancy(cublox_ms *h, int sid, float *dst) {
  return guarded([&] {
    if (H->server->admitted(sid)) {
      H->server->grid(sid).copyOccupancyToHost(dst);
    } else {
      evictedMap(*H->server, sid, dst);
    }
  });
}

int cublox_ms_origin(cublox_ms *h, int sid, int origin_i[3]) {
  return guarded([&] {
    if (H->server->admitted(sid)) {
      const auto &g = H->server->grid(sid);
      const Eigen::Vector3i o = g.getOriginIndex();
      origin_i[0] = o.x();
      origin_i[1] = o.y();
      origin_i[2] = o.z();
    } else {
      const cublox_snapshot_header hd = evictedMap(*H->server, sid, nullptr);
      for (int a = 0; a < 3; ++a) origin_i[a] = hd.origin_i[a];
    }
  });
}

int cublox_ms_stats(cublox_ms *h, double out[14]) {
  return guarded([&] {
    const MS::Stats s = H->server->stats();
    out[0] = static_cast<double>(s.ticks);
    out[1] = static_cast<double>(s.scans);
    out[2] = static_cast<double>(s.fetches);
    out[3] = static_cast<double>(s.points);
    out[4] = s.device_ms;
    out[5] = static_cast<double>(s.resident_bytes);
    out[6] = static_cast<double>(s.workspace_bytes);
    out[7] = s.admitted;
    out[8] = s.waiting;
    out[9] = H->server->nowMs();
    out[10] = s.launch_ms;
    out[11] = static_cast<double>(s.evictions);
    out[12] = static_cast<double>(s.restores);
    out[13] = static_cast<double>(s.snapshot_bytes);
  });
}

}  // extern "C"
