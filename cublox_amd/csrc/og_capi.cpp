// Flat C binding of the drop-in cublox::OccupancyGrid for FFI callers
// (Python ctypes in this package; cgo/JNI would bind the same symbols).
// Every entry point catches C++ exceptions and returns a status; the message
// of the last failure on the calling thread is in cublox_og_last_error().
#include <cublox/GridBatch.hpp>
#include <cublox/OccupancyGrid.hpp>
#include <cublox/capi.h>
#include <cublox/og_capi.h>
#include <cublox/utils.hpp>

#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>

namespace {
thread_local std::string g_err;

template <typename F> int guarded(F &&f) {
  try {
    f();
    return CUBLOX_OK;
  } catch (const std::exception &e) {
    g_err = e.what();
    return CUBLOX_ERR_RUNTIME;
  } catch (...) {
    g_err = "unknown C++ exception";
    return CUBLOX_ERR_RUNTIME;
  }
}

using OG = cublox::OccupancyGrid;

Eigen::Vector3f v3(const float *p) { return Eigen::Vector3f(p[0], p[1], p[2]); }

std::optional<OG::VoxelState> state_of(int filter) {
  if (filter < 0) return std::nullopt;
  return static_cast<OG::VoxelState>(filter);
}
} // namespace

namespace cublox_ffi {
// Shared with the MapServer binding (ms_capi.cpp).
void set_last_error(const char *msg) { g_err = msg; }
} // namespace cublox_ffi

extern "C" {

const char *cublox_og_last_error(void) { return g_err.c_str(); }

cublox_og *cublox_og_create(int hx, int hy, int hz, float resolution,
                            int origin_at_center, int has_threshold,
                            double threshold, const float origin[3]) {
  OG *g = nullptr;
  const int st = guarded([&] {
    g = new OG(Eigen::Vector3i(hx, hy, hz), resolution, origin_at_center != 0,
               has_threshold ? std::optional<double>(threshold) : std::nullopt,
               origin ? v3(origin) : Eigen::Vector3f::Zero());
  });
  if (st != CUBLOX_OK) return nullptr;
  return reinterpret_cast<cublox_og *>(g);
}

void cublox_og_destroy(cublox_og *h) { delete reinterpret_cast<OG *>(h); }

#define OGP reinterpret_cast<OG *>(h)
#define OGC reinterpret_cast<const OG *>(h)

int cublox_og_set_max_range(cublox_og *h, float r) {
  return guarded([&] { OGP->setMaxRaycastRange(r); });
}

int cublox_og_set_log_odds(cublox_og *h, float l_hit, float l_miss,
                           float l_min, float l_max, float l_free,
                           float l_occupied) {
  return guarded([&] {
    OGP->setLogOddsParams(l_hit, l_miss, l_min, l_max, l_free, l_occupied);
  });
}

int cublox_og_update(cublox_og *h, const float *x, const float *y,
                     const float *z, int n, const float origin[3]) {
  return guarded([&] { OGP->updateColumns(x, y, z, n, v3(origin)); });
}

int cublox_og_update_sensor(cublox_og *h, const float *xs, const float *ys,
                            const float *zs, int n, const float T[12]) {
  return guarded([&] { OGP->updateSensorFrameColumns(xs, ys, zs, n, T); });
}

int cublox_og_update_sensor_device(cublox_og *h, const float *d_xs,
                                   const float *d_ys, const float *d_zs,
                                   int n, const float T[12]) {
  return guarded([&] { OGP->updateSensorFrameDevice(d_xs, d_ys, d_zs, n, T); });
}

int cublox_og_update_device(cublox_og *h, const float *d_x, const float *d_y,
                            const float *d_z, int n, const float origin[3]) {
  return guarded([&] { OGP->updateDevice(d_x, d_y, d_z, n, v3(origin)); });
}

int cublox_og_cast_rays_device(cublox_og *h, const float *d_x,
                               const float *d_y, const float *d_z, int n,
                               const float origin[3]) {
  return guarded([&] { OGP->castRaysDevice(d_x, d_y, d_z, n, v3(origin)); });
}

int cublox_og_cast_rays_sensor_device(cublox_og *h, const float *d_xs,
                                      const float *d_ys, const float *d_zs,
                                      int n, const float T[12]) {
  return guarded([&] { OGP->castRaysSensorFrameDevice(d_xs, d_ys, d_zs, n, T); });
}

// path: 0 fused tiles, 1 per-voxel atomics.
int cublox_og_set_update_path(cublox_og *h, int path) {
  return guarded([&] {
    OGP->setUpdatePath(path == 0 ? OG::UpdatePath::FUSED_TILES
                                 : OG::UpdatePath::PER_VOXEL_ATOMICS);
  });
}

int cublox_og_set_near_steps(cublox_og *h, int k) {
  return guarded([&] { OGP->setNearSteps(k); });
}

int cublox_og_update_cloud2(cublox_og *h, const void *data, int on_device,
                            unsigned long long data_bytes,
                            const unsigned int layout[6], const float T[12],
                            int *updated) {
  return guarded([&] {
    OG::PointCloud2Layout l;
    l.width = layout[0];
    l.height = layout[1];
    l.point_step = layout[2];
    l.off_x = layout[3];
    l.off_y = layout[4];
    l.off_z = layout[5];
    const size_t b = static_cast<size_t>(data_bytes);
    const bool u = on_device ? OGP->updatePointCloud2Device(data, l, T, b)
                             : OGP->updatePointCloud2(data, l, T, b);
    if (updated) *updated = u ? 1 : 0;
  });
}

int cublox_og_update_depth(cublox_og *h, const void *data, int on_device,
                           unsigned int width, unsigned int height,
                           unsigned int step, int encoding, const float K[4],
                           const float T[12], int *updated) {
  return guarded([&] {
    if (encoding != 0 && encoding != 1) {
      throw std::invalid_argument("cublox_og_update_depth: encoding must be 0 or 1");
    }
    const OG::DepthEncoding e = encoding == 0 ? OG::DepthEncoding::MILLIMETRES_U16
                                              : OG::DepthEncoding::METRES_F32;
    const bool u = on_device
                       ? OGP->updateDepthImageDevice(data, width, height, step, e, K, T)
                       : OGP->updateDepthImage(data, width, height, step, e, K, T);
    if (updated) *updated = u ? 1 : 0;
  });
}

int cublox_og_viz_cloud(const cublox_og *h, const float robot_pos[3],
                        float max_viz_radius, int subsample, int max_points,
                        float *out4, unsigned int out_cap, unsigned int *written,
                        unsigned long long *fetched) {
  return guarded([&] {
    std::vector<OG::VizPoint> v;
    const size_t f = OGC->buildOccupancyVizCloud(v3(robot_pos), max_viz_radius,
                                                 subsample, max_points, v);
    const size_t n = std::min<size_t>(v.size(), out_cap);
    for (size_t k = 0; k < n; ++k) {
      out4[4 * k + 0] = v[k].x;
      out4[4 * k + 1] = v[k].y;
      out4[4 * k + 2] = v[k].z;
      out4[4 * k + 3] = v[k].rgb;
    }
    if (written) *written = static_cast<unsigned int>(n);
    if (fetched) *fetched = f;
  });
}

// Snapshot buffers handed to the caller: the vectors themselves, found by
// their data pointer when the caller frees them (no copy on the way out).
static std::mutex g_buf_mu;
static std::unordered_map<void *, std::vector<uint8_t> *> g_bufs;

int cublox_og_save_snapshot(const cublox_og *h, void **out, size_t *bytes) {
  return guarded([&] {
    if (!out || !bytes) throw std::invalid_argument("cublox_og_save_snapshot: null output");
    *out = nullptr;
    *bytes = 0;
    auto v = std::make_unique<std::vector<uint8_t>>(OGC->saveSnapshot());
    void *p = v->data();
    const size_t n = v->size();
    {
      std::lock_guard<std::mutex> lk(g_buf_mu);
      g_bufs.emplace(p, v.get());
    }
    v.release();
    *out = p;
    *bytes = n;
  });
}

int cublox_og_load_snapshot(cublox_og *h, const void *data, size_t bytes) {
  return guarded(
      [&] { OGP->loadSnapshot(static_cast<const uint8_t *>(data), bytes); });
}

void cublox_og_free_buffer(void *p) {
  if (!p) return;
  std::vector<uint8_t> *v = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_buf_mu);
    auto it = g_bufs.find(p);
    if (it == g_bufs.end()) return;
    v = it->second;
    g_bufs.erase(it);
  }
  delete v;
}

int cublox_og_save_snapshot_file(const cublox_og *h, const char *path) {
  return guarded([&] { OGC->saveSnapshotFile(path ? path : ""); });
}

int cublox_og_load_snapshot_file(cublox_og *h, const char *path) {
  return guarded([&] { OGP->loadSnapshotFile(path ? path : ""); });
}

int cublox_og_set_changed_capacity(cublox_og *h, unsigned int capacity) {
  return guarded([&] { OGP->setChangedVoxelCapacity(capacity); });
}

int cublox_og_changed_voxels(const cublox_og *h, int *ids, float *vals,
                             unsigned int out_cap, unsigned long long *count,
                             unsigned int *written) {
  return guarded([&] {
    std::vector<OG::ChangedVoxel> v;
    const unsigned long long c = OGC->changedVoxels(v);
    const size_t n = std::min<size_t>(v.size(), out_cap);
    for (size_t k = 0; k < n; ++k) {
      ids[k] = v[k].hash_id;
      vals[k] = v[k].log_odds;
    }
    if (count) *count = c;
    if (written) *written = static_cast<unsigned int>(n);
  });
}

int cublox_og_cast_rays_fused_sensor_device(cublox_og *h, const float *d_xs,
                                            const float *d_ys,
                                            const float *d_zs, int n,
                                            const float T[12]) {
  return guarded(
      [&] { OGP->castRaysFusedSensorFrameDevice(d_xs, d_ys, d_zs, n, T); });
}

int cublox_og_census(const cublox_og *h, unsigned long long out[4]) {
  return guarded([&] { OGC->counterCensus(out); });
}

int cublox_og_apply_counters(cublox_og *h) {
  return guarded([&] { OGP->applyCounters(); });
}

// *moved = 1 when the window origin changed.
int cublox_og_recenter(cublox_og *h, const float pos[3], int *moved) {
  return guarded([&] {
    const Eigen::Vector3i before = OGP->getOriginIndex();
    OGP->recenter(v3(pos));
    if (moved) *moved = OGP->getOriginIndex() != before ? 1 : 0;
  });
}

int cublox_og_fetch(const cublox_og *h, const float center[3], float radius,
                    int filter, int max_results, float *out_pos,
                    unsigned char *out_state, int cap, int *n_out) {
  return guarded([&] {
    const auto samples =
        OGC->fetchOccupancyAround(v3(center), radius, state_of(filter), max_results);
    const int n = static_cast<int>(samples.size());
    for (int i = 0; i < n && i < cap; ++i) {
      out_pos[3 * i + 0] = samples[i].position.x();
      out_pos[3 * i + 1] = samples[i].position.y();
      out_pos[3 * i + 2] = samples[i].position.z();
      out_state[i] = static_cast<unsigned char>(samples[i].state);
    }
    *n_out = n;
  });
}

// counts[q] per query; out_pos/out_state sized n_queries * per_query_cap
// (x3 for positions) receive the samples densely, query after query.
int cublox_og_fetch_batch(const cublox_og *h, const float *centers,
                          int n_queries, float radius, int filter,
                          int per_query_cap, unsigned int *counts,
                          float *out_pos, unsigned char *out_state) {
  return guarded([&] {
    std::vector<Eigen::Vector3f> c(static_cast<size_t>(n_queries));
    for (int q = 0; q < n_queries; ++q) c[q] = v3(centers + 3 * q);
    std::vector<unsigned int> cnt;
    std::vector<float> pos;
    std::vector<unsigned char> st;
    OGC->fetchOccupancyBatch(c, radius, state_of(filter), per_query_cap, cnt,
                             pos, st);
    std::memcpy(counts, cnt.data(), sizeof(unsigned int) * cnt.size());
    if (out_pos && !pos.empty()) std::memcpy(out_pos, pos.data(), sizeof(float) * pos.size());
    if (out_state && !st.empty()) std::memcpy(out_state, st.data(), st.size());
  });
}

int cublox_og_occ_to_host(const cublox_og *h, float *dst) {
  return guarded([&] { OGC->copyOccupancyToHost(dst); });
}

int cublox_og_counters_to_host(const cublox_og *h, int *op, int *hit) {
  return guarded([&] { OGC->copyCountersToHost(op, hit); });
}

int cublox_og_load_occ(cublox_og *h, const float *src) {
  return guarded([&] { OGP->loadOccupancyFromHost(src); });
}

int cublox_og_reset(cublox_og *h) {
  return guarded([&] { OGP->reset(); });
}

int cublox_og_reset_voxel(cublox_og *h, int hash_id) {
  return guarded([&] { OGP->resetVoxel(hash_id); });
}

// Geometry: map[3], half[3], origin_i[3], origin_f[3], bound_min_i[3],
// bound_max_i[3], voxel_num, resolution, resolution_inv.
int cublox_og_geometry(const cublox_og *h, int map[3], int half[3],
                       int origin_i[3], float origin_f[3], int bmin[3],
                       int bmax[3], int *voxel_num, float *res,
                       float *res_inv) {
  return guarded([&] {
    const OG &g = *OGC;
    for (int a = 0; a < 3; ++a) {
      map[a] = g.config_.map_size_i(a);
      half[a] = g.config_.half_map_size_i(a);
      origin_i[a] = g.origin_i_(a);
      origin_f[a] = g.origin_f_(a);
      bmin[a] = g.bound_min_i_(a);
      bmax[a] = g.bound_max_i_(a);
    }
    *voxel_num = g.config_.voxel_num;
    *res = g.config_.resolution;
    *res_inv = g.config_.resolution_inv;
  });
}

// kind: 0 occupied, 1 free, 2 unknown; by: 0 position, 1 global index,
// 2 hash id (uses q[0]).
int cublox_og_query(const cublox_og *h, int kind, int by, const float *qf,
                    const int *qi, int *result) {
  return guarded([&] {
    const OG &g = *OGC;
    bool r = false;
    if (by == 0) {
      const Eigen::Vector3f p = v3(qf);
      r = kind == 0 ? g.isOccupied(p) : kind == 1 ? g.isFree(p) : g.isUnknown(p);
    } else if (by == 1) {
      const Eigen::Vector3i p(qi[0], qi[1], qi[2]);
      r = kind == 0 ? g.isOccupied(p) : kind == 1 ? g.isFree(p) : g.isUnknown(p);
    } else {
      r = kind == 0 ? g.isOccupied(qi[0]) : kind == 1 ? g.isFree(qi[0]) : g.isUnknown(qi[0]);
    }
    *result = r ? 1 : 0;
  });
}

int cublox_og_max_horizontal_radius(const cublox_og *h, const float c[3],
                                    float *r) {
  return guarded([&] { *r = OGC->maxHorizontalInWindowRadius(v3(c)); });
}

int cublox_og_enable_timing(cublox_og *h, int on) {
  return guarded([&] { OGP->enableKernelTiming(on != 0); });
}

// out_ms[3] = ray cast, apply, recenter device time; out_n[4] = launches of
// each, then voxels zeroed by recenters.
int cublox_og_kernel_times(const cublox_og *h, double out_ms[4],
                           long long out_n[7]) {
  return guarded([&] {
    const auto t = OGC->kernelTimes();
    out_ms[0] = t.update_ms;
    out_ms[1] = t.apply_ms;
    out_ms[2] = t.recenter_ms;
    out_ms[3] = t.fetch_ms;
    out_n[0] = t.update_launches;
    out_n[1] = t.apply_launches;
    out_n[2] = t.recenter_launches;
    out_n[3] = t.recenter_zeroed_voxels;
    out_n[4] = t.fetch_launches;
    out_n[5] = t.fetch_tested_voxels;
    out_n[6] = t.fetch_samples;
  });
}

int cublox_og_reset_kernel_times(cublox_og *h) {
  return guarded([&] { OGP->resetKernelTimes(); });
}

#define GBC reinterpret_cast<const cublox::GridBatch *>(b)
#define GBP reinterpret_cast<cublox::GridBatch *>(b)

cublox_gb *cublox_gb_create(const int limits[9]) {
  cublox_gb *out = nullptr;
  g_err.clear();
  guarded([&] {
    cublox::BatchLimits l;
    l.max_grids = limits[0];
    l.max_points = limits[1];
    l.max_tiles = limits[2];
    l.max_segments_per_ray = limits[3];
    l.max_near_steps = limits[4];
    l.min_chunk = limits[5];
    l.max_slabs = limits[6];
    l.max_regions = limits[7];
    l.changed_voxels = limits[8] != 0;
    out = reinterpret_cast<cublox_gb *>(new cublox::GridBatch(l));
  });
  return out;
}

void cublox_gb_destroy(cublox_gb *b) { delete GBP; }

int cublox_gb_segment_bound(const cublox_og *h) {
  return cublox::GridBatch::segmentBound(*OGC);
}

int cublox_gb_tile_count(const cublox_og *h) {
  return cublox::GridBatch::tileCount(*OGC);
}

namespace {
// The tick arguments of cublox_gb_tick / cublox_gb_submit as GridBatch takes them.
struct TickArgs {
  std::vector<cublox::GridScan> scans;
  std::vector<std::pair<OG *, Eigen::Vector3f>> rc;
};
TickArgs tick_args(int n_scans, cublox_og *const *grids, const float *const *d_x,
                   const float *const *d_y, const float *const *d_z, const int *n,
                   const float *T, const int *point_stride, const float *range_drop_sq,
                   int n_recenter, cublox_og *const *rgrids, const float *rpos) {
  TickArgs a;
  a.scans.resize(static_cast<size_t>(std::max(n_scans, 0)));
  for (int k = 0; k < n_scans; ++k) {
    cublox::GridScan &s = a.scans[static_cast<size_t>(k)];
    s.grid = reinterpret_cast<OG *>(grids[k]);
    s.d_x = d_x[k];
    s.d_y = d_y[k];
    s.d_z = d_z[k];
    s.n = n[k];
    s.sensor_frame = true;
    for (int i = 0; i < 12; ++i) s.T[static_cast<size_t>(i)] = T[12 * k + i];
    if (point_stride) s.point_stride = point_stride[k];
    if (range_drop_sq) s.range_drop_sq = range_drop_sq[k];
  }
  for (int k = 0; k < n_recenter; ++k) {
    a.rc.emplace_back(reinterpret_cast<OG *>(rgrids[k]), v3(rpos + 3 * k));
  }
  return a;
}
}  // namespace

int cublox_gb_tick(cublox_gb *b, int n_scans, cublox_og *const *grids,
                   const float *const *d_x, const float *const *d_y,
                   const float *const *d_z, const int *n, const float *T,
                   const int *point_stride, const float *range_drop_sq,
                   int n_recenter, cublox_og *const *rgrids,
                   const float *rpos) {
  return guarded([&] {
    const TickArgs a = tick_args(n_scans, grids, d_x, d_y, d_z, n, T, point_stride,
                                 range_drop_sq, n_recenter, rgrids, rpos);
    GBP->tick(a.scans, a.rc);
  });
}

int cublox_gb_submit(cublox_gb *b, int n_scans, cublox_og *const *grids,
                     const float *const *d_x, const float *const *d_y,
                     const float *const *d_z, const int *n, const float *T,
                     const int *point_stride, const float *range_drop_sq,
                     int n_recenter, cublox_og *const *rgrids,
                     const float *rpos) {
  return guarded([&] {
    const TickArgs a = tick_args(n_scans, grids, d_x, d_y, d_z, n, T, point_stride,
                                 range_drop_sq, n_recenter, rgrids, rpos);
    GBP->submit(a.scans, a.rc);
  });
}

int cublox_gb_wait(cublox_gb *b) {
  return guarded([&] { GBP->wait(); });
}

int cublox_gb_set_pipelined(cublox_gb *b, int on) {
  return guarded([&] { GBP->setPipelined(on != 0); });
}

int cublox_gb_set_graph_mode(cublox_gb *b, int on) {
  return guarded([&] { GBP->setGraphMode(on != 0); });
}

int cublox_gb_set_tuning(cublox_gb *b, int n, const int *settings,
                         int ticks_each) {
  return guarded([&] {
    std::vector<cublox::TuneSetting> c;
    for (int k = 0; k < n; ++k) {
      const int *v = settings + 4 * k;
      c.push_back(cublox::TuneSetting{v[0], v[1], v[2], v[3]});
    }
    GBP->setTuning(c, ticks_each);
  });
}

int cublox_gb_tuning(const cublox_gb *b, int *active) {
  return guarded([&] { *active = GBC->tuning() ? 1 : 0; });
}

int cublox_gb_setting(const cublox_gb *b, int out[4]) {
  return guarded([&] {
    const auto s = GBC->setting();
    out[0] = s.near_steps;
    out[1] = s.chunk;
    out[2] = s.lanes;
    out[3] = s.hot_mode;
  });
}

int cublox_gb_set_setting(cublox_gb *b, int near_steps, int chunk, int lanes,
                          int hot_mode) {
  return guarded([&] {
    GBP->setSetting(cublox::TuneSetting{near_steps, chunk, lanes, hot_mode});
  });
}

int cublox_gb_tune_results(const cublox_gb *b, double *out_ms, int cap) {
  return guarded([&] {
    const auto r = GBC->tuneResults();
    for (int k = 0; k < cap && k < static_cast<int>(r.size()); ++k) {
      out_ms[k] = r[static_cast<size_t>(k)].second;
    }
  });
}

int cublox_gb_last(const cublox_gb *b, double out[3]) {
  return guarded([&] {
    out[0] = GBC->lastUpdateMs();
    out[1] = GBC->lastRecenterMs();
    out[2] = static_cast<double>(GBC->lastZeroedVoxels());
  });
}

int cublox_gb_counters(const cublox_gb *b, int out[5]) {
  return guarded([&] {
    const auto c = GBC->lastCounters();
    for (int k = 0; k < 5; ++k) out[k] = c[static_cast<size_t>(k)];
  });
}

const float *cublox_og_device_occ(const cublox_og *h) {
  return OGC->deviceOccupancy();
}

void *cublox_og_stream(const cublox_og *h) { return OGC->stream(); }

} // extern "C"

// Host index math of include/cublox/utils.hpp, exported for FFI checks.
extern "C" {

int cublox_utils_pos_to_global(float p, float inv_res, int center) {
  return cublox::posToGlobalIndex(p, inv_res, center != 0);
}
float cublox_utils_global_to_pos(int g, float res, int center) {
  return cublox::globalIndexToPos(g, res, center != 0);
}
int cublox_utils_global_to_local(int g, int map, int half) {
  return cublox::globalIndexToLocalIndex(g, map, half);
}
int cublox_utils_local_to_global(int l, int map, int half, int origin_g) {
  return cublox::localIndexToGlobalIndex(l, map, half, origin_g);
}
int cublox_utils_local_to_hash(const int l[3], const int map[3],
                               const int half[3]) {
  return cublox::localIndexToHashId(Eigen::Vector3i(l[0], l[1], l[2]),
                                    Eigen::Vector3i(map[0], map[1], map[2]),
                                    Eigen::Vector3i(half[0], half[1], half[2]));
}
void cublox_utils_hash_to_local(int h, const int map[3], const int half[3],
                                int l[3]) {
  Eigen::Vector3i out;
  cublox::hashIdToLocalIndex(h, Eigen::Vector3i(map[0], map[1], map[2]),
                             Eigen::Vector3i(half[0], half[1], half[2]), out);
  for (int a = 0; a < 3; ++a) l[a] = out(a);
}
void cublox_utils_hash_to_pos(int h, const int map[3], const int half[3],
                              const int origin[3], float res, int center,
                              float pos[3]) {
  Eigen::Vector3f out;
  cublox::hashIdToPos(h, Eigen::Vector3i(map[0], map[1], map[2]),
                      Eigen::Vector3i(half[0], half[1], half[2]),
                      Eigen::Vector3i(origin[0], origin[1], origin[2]), res,
                      center != 0, out);
  for (int a = 0; a < 3; ++a) pos[a] = out(a);
}
int cublox_utils_pos_to_hash(const float pos[3], const int map[3],
                             const int half[3], float inv_res, int center) {
  return cublox::posToHashIndex(Eigen::Vector3f(pos[0], pos[1], pos[2]),
                                Eigen::Vector3i(map[0], map[1], map[2]),
                                Eigen::Vector3i(half[0], half[1], half[2]),
                                inv_res, center != 0);
}
float cublox_logit(float p) { return cublox::logit(p); }

} // extern "C"
