/*
 * Flat C binding of the drop-in cublox::OccupancyGrid (libcublox.so), for
 * FFI callers that cannot bind C++ (Python ctypes, cgo, JNI, N-API).
 *
 * Each entry point wraps one method of the C++ API in
 * include/cublox/OccupancyGrid.hpp, itself the reference class
 * (reference include/cublox/OccupancyGrid.hpp:29-98), and behaves as the
 * method does, blocking included. Every call catches C++ exceptions and
 * returns CUBLOX_OK or CUBLOX_ERR_RUNTIME; cublox_og_last_error() holds the
 * message of the calling thread's last failure, in the reference's
 * "cublox::cudaCheck: CUDA error in `expr`: msg (file:line)" shape for
 * device errors (reference src/OccupancyGrid.cpp:32-49).
 *
 * Host arrays are plain pointers; d_-prefixed ones are device pointers
 * (from cublox_malloc in capi.h or any HIP allocation on the same device).
 * A handle is not thread-safe, as the reference object is not.
 */
#ifndef CUBLOX_OG_CAPI_H
#define CUBLOX_OG_CAPI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cublox_og cublox_og;

const char *cublox_og_last_error(void);

/* OccupancyGrid(half_map_size_i, resolution, origin_at_center,
 * recenter_threshold) (reference OccupancyGrid.hpp:42-45); has_threshold = 0
 * is std::nullopt. `origin` (may be NULL) sets an initial origin; NULL keeps
 * the reference's first-update origin pinning. Returns NULL on failure. */
cublox_og *cublox_og_create(int hx, int hy, int hz, float resolution,
                            int origin_at_center, int has_threshold,
                            double threshold, const float origin[3]);
void cublox_og_destroy(cublox_og *h); /* ~OccupancyGrid (hpp:47) */

/* setMaxRaycastRange (hpp:68-70), setLogOddsParams (hpp:74-85). */
int cublox_og_set_max_range(cublox_og *h, float r);
int cublox_og_set_log_odds(cublox_og *h, float l_hit, float l_miss,
                           float l_min, float l_max, float l_free,
                           float l_occupied);

/* update(input_cloud, sensor_origin) (hpp:53-54) with the cloud as three
 * host columns of n floats. */
int cublox_og_update(cublox_og *h, const float *x, const float *y,
                     const float *z, int n, const float origin[3]);
/* The driver's scan transform (main.cpp:640-693) fused on device: sensor-
 * frame host columns and T = sensor->world, row-major 3x4 [R | t]. */
int cublox_og_update_sensor(cublox_og *h, const float *xs, const float *ys,
                            const float *zs, int n, const float T[12]);
/* The same two with device-resident columns (no host copy). */
int cublox_og_update_sensor_device(cublox_og *h, const float *d_xs,
                                   const float *d_ys, const float *d_zs,
                                   int n, const float T[12]);
int cublox_og_update_device(cublox_og *h, const float *d_x, const float *d_y,
                            const float *d_z, int n, const float origin[3]);

/* The halves of update (reference launchRayCastUpdate, launchApplyUpdate,
 * kernels.hpp:54-76) for stage-wise parity: cast_rays leaves per-voxel
 * counters populated, apply_counters folds and zeroes them. The fused cast
 * exports the counts the fused path would fold. */
int cublox_og_cast_rays_device(cublox_og *h, const float *d_x,
                               const float *d_y, const float *d_z, int n,
                               const float origin[3]);
int cublox_og_cast_rays_sensor_device(cublox_og *h, const float *d_xs,
                                      const float *d_ys, const float *d_zs,
                                      int n, const float T[12]);
int cublox_og_cast_rays_fused_sensor_device(cublox_og *h, const float *d_xs,
                                            const float *d_ys,
                                            const float *d_zs, int n,
                                            const float T[12]);
int cublox_og_apply_counters(cublox_og *h);
/* out: voxels with op != 0, sum of op, sum of hit, voxels with hit != 0. */
int cublox_og_census(const cublox_og *h, unsigned long long out[4]);
/* 0: fused tile-binned update (default); 1: per-voxel atomic counters. */
int cublox_og_set_update_path(cublox_og *h, int path);
/* Near-field steps of the fused update, 1..12. */
int cublox_og_set_near_steps(cublox_og *h, int k);
/* Driver ingest (reference main.cpp:558-693, run on the device):
 * OccupancyGrid::updatePointCloud2 / updateDepthImage. layout = {width,
 * height, point_step, off_x, off_y, off_z}; on_device: data is device
 * memory. encoding 0: 16-bit millimetres, 1: 32-bit float metres; K = {fx,
 * fy, cx, cy}. *updated = 1 when some point was kept (the grid updated).
 * data_bytes: the payload's length; CUBLOX_ERR_INVALID when it is shorter
 * than min(width * height, 1 M) * point_step. */
int cublox_og_update_cloud2(cublox_og *h, const void *data, int on_device,
                            unsigned long long data_bytes,
                            const unsigned int layout[6], const float T[12],
                            int *updated);
int cublox_og_update_depth(cublox_og *h, const void *data, int on_device,
                           unsigned int width, unsigned int height,
                           unsigned int step, int encoding, const float K[4],
                           const float T[12], int *updated);
/* The driver's occupancy viz cloud (OccupancyGrid::buildOccupancyVizCloud,
 * main.cpp:484-491 and 740-817): up to out_cap (x, y, z, rgb) records to
 * out4; *written records, *fetched the driver's sample count. */
int cublox_og_viz_cloud(const cublox_og *h, const float robot_pos[3],
                        float max_viz_radius, int subsample, int max_points,
                        float *out4, unsigned int out_cap, unsigned int *written,
                        unsigned long long *fetched);
/* Map snapshots (OccupancyGrid::saveSnapshot / loadSnapshot; format in
 * cublox/snapshot.h). save: *out points to a buffer owned by the library
 * until passed to cublox_og_free_buffer; *bytes is its size. */
int cublox_og_save_snapshot(const cublox_og *h, void **out, size_t *bytes);
int cublox_og_load_snapshot(cublox_og *h, const void *data, size_t bytes);
void cublox_og_free_buffer(void *p);
int cublox_og_save_snapshot_file(const cublox_og *h, const char *path);
int cublox_og_load_snapshot_file(cublox_og *h, const char *path);
/* Changed-voxel stream (no reference counterpart as a call; it enables the
 * compaction of kernels.cu:202-209): capacity > 0 records the changes of
 * each later update, 0 turns it off. */
int cublox_og_set_changed_capacity(cublox_og *h, unsigned int capacity);
/* Changes of the last update: *count = voxels that changed; the first
 * min(*count, capacity, out_cap) go to ids/vals, *written of them. */
int cublox_og_changed_voxels(const cublox_og *h, int *ids, float *vals,
                             unsigned int out_cap, unsigned long long *count,
                             unsigned int *written);

/* recenter(pos) (hpp:57); *moved = 1 when the origin changed. */
int cublox_og_recenter(cublox_og *h, const float pos[3], int *moved);

/* fetchOccupancyAround(center, radius, state_filter, max_results)
 * (hpp:62-65); filter < 0 is std::nullopt, else VoxelState (hpp:31-35).
 * Writes up to cap samples (xyz float triples and states). */
int cublox_og_fetch(const cublox_og *h, const float center[3], float radius,
                    int filter, int max_results, float *out_pos,
                    unsigned char *out_state, int cap, int *n_out);
/* Many queries in one launch (no reference counterpart): counts[q] samples
 * of query q, densely packed in query order in out_pos / out_state; each
 * query keeps at most per_query_cap samples. */
int cublox_og_fetch_batch(const cublox_og *h, const float *centers,
                          int n_queries, float radius, int filter,
                          int per_query_cap, unsigned int *counts,
                          float *out_pos, unsigned char *out_state);

/* Voxel data access (occupancy log-odds and counters, hash order). */
int cublox_og_occ_to_host(const cublox_og *h, float *dst);
int cublox_og_counters_to_host(const cublox_og *h, int *op, int *hit);
int cublox_og_load_occ(cublox_og *h, const float *src);
const float *cublox_og_device_occ(const cublox_og *h);
void *cublox_og_stream(const cublox_og *h);

/* reset() / resetVoxel(hash_id) (hpp:49-50). */
int cublox_og_reset(cublox_og *h);
int cublox_og_reset_voxel(cublox_og *h, int hash_id);

/* Grid::Config and window state (reference Grid.hpp:22-60). */
int cublox_og_geometry(const cublox_og *h, int map[3], int half[3],
                       int origin_i[3], float origin_f[3], int bmin[3],
                       int bmax[3], int *voxel_num, float *res,
                       float *res_inv);
/* isOccupied / isFree / isUnknown (hpp:90-98): kind 0/1/2 = occupied /
 * free / unknown; by 0/1/2 = position qf[3] / global index qi[3] / hash
 * id qi[0]. */
int cublox_og_query(const cublox_og *h, int kind, int by, const float *qf,
                    const int *qi, int *result);
/* Grid::maxHorizontalInWindowRadius (Grid.hpp:42). */
int cublox_og_max_horizontal_radius(const cublox_og *h, const float c[3],
                                    float *r);

/* Per-kernel device timing (no reference counterpart): out_ms = ray cast
 * (or fused update), apply, recenter and batched-fetch kernel device-time
 * totals; out_n = the first three launch counts, the voxels zeroed by
 * recenters, the fetch launches, the voxels those fetches tested and the
 * samples they returned. */
int cublox_og_enable_timing(cublox_og *h, int on);
int cublox_og_kernel_times(const cublox_og *h, double out_ms[4],
                           long long out_n[7]);
int cublox_og_reset_kernel_times(cublox_og *h);

/* Batched updates of many grids (cublox::GridBatch, no reference
 * counterpart). limits: max_grids, max_points, max_tiles,
 * max_segments_per_ray, max_near_steps, min_chunk, max_slabs, max_regions,
 * changed_voxels (9 ints; see include/cublox/GridBatch.hpp). */
typedef struct cublox_gb cublox_gb;
cublox_gb *cublox_gb_create(const int limits[9]);
void cublox_gb_destroy(cublox_gb *b);
int cublox_gb_segment_bound(const cublox_og *g);
int cublox_gb_tile_count(const cublox_og *g);
/* One tick: n_scans scans (grids[k]: device columns d_x/d_y/d_z[k] of
 * n[k] points, sensor-frame with T[12k..12k+11]; point_stride[k] and
 * range_drop_sq[k] as in GridScan, either array may be NULL), then the
 * recenter of rgrids[k] toward rpos[3k..3k+2]. Blocks until done. */
int cublox_gb_tick(cublox_gb *b, int n_scans, cublox_og *const *grids,
                   const float *const *d_x, const float *const *d_y,
                   const float *const *d_z, const int *n, const float *T,
                   const int *point_stride, const float *range_drop_sq,
                   int n_recenter, cublox_og *const *rgrids,
                   const float *rpos);
/* cublox_gb_tick in two halves (GridBatch::submit / wait): submit returns
 * once the tick is launched; with pipelining it first waits for the tick
 * submitted two before it. wait blocks until every submitted tick ends. */
int cublox_gb_submit(cublox_gb *b, int n_scans, cublox_og *const *grids,
                     const float *const *d_x, const float *const *d_y,
                     const float *const *d_z, const int *n, const float *T,
                     const int *point_stride, const float *range_drop_sq,
                     int n_recenter, cublox_og *const *rgrids,
                     const float *rpos);
int cublox_gb_wait(cublox_gb *b);
/* Pipelined ticks (GridBatch::setPipelined): consecutive ticks overlap on
 * two streams and workspaces, every grid's folds in tick order. */
int cublox_gb_set_pipelined(cublox_gb *b, int on);
int cublox_gb_set_graph_mode(cublox_gb *b, int on);
/* settings: n quadruples (near_steps, chunk, lanes per ray 0/1/4, hot-tile
 * mode 0/1) as cublox::TuneSetting; ticks_each measured ticks each. */
int cublox_gb_set_tuning(cublox_gb *b, int n, const int *settings,
                         int ticks_each);
int cublox_gb_tuning(const cublox_gb *b, int *active);
int cublox_gb_setting(const cublox_gb *b, int out[4]);
int cublox_gb_set_setting(cublox_gb *b, int near_steps, int chunk, int lanes,
                          int hot_mode);
/* out_ms[k]: mean device ms of candidate k (cap entries). */
int cublox_gb_tune_results(const cublox_gb *b, double *out_ms, int cap);
/* out: update ms, recenter ms (0 in graph mode: all in update), zeroed. */
int cublox_gb_last(const cublox_gb *b, double out[3]);
/* Counters of the last tick's update (cublox_integrate_batch_counters). */
int cublox_gb_counters(const cublox_gb *b, int out[5]);

/* Index math (reference include/cublox/utils.hpp:24-181). */
int cublox_utils_pos_to_global(float p, float inv_res, int center);
float cublox_utils_global_to_pos(int g, float res, int center);
int cublox_utils_global_to_local(int g, int map, int half);
int cublox_utils_local_to_global(int l, int map, int half, int origin_g);
int cublox_utils_local_to_hash(const int l[3], const int map[3],
                               const int half[3]);
void cublox_utils_hash_to_local(int h, const int map[3], const int half[3],
                                int l[3]);
void cublox_utils_hash_to_pos(int h, const int map[3], const int half[3],
                              const int origin[3], float res, int center,
                              float pos[3]);
int cublox_utils_pos_to_hash(const float pos[3], const int map[3],
                             const int half[3], float inv_res, int center);

#ifdef __cplusplus
}
#endif

#endif /* CUBLOX_OG_CAPI_H */
