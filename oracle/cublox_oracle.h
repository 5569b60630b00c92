/*
 * cublox oracle — CPU restatement of the reference occupancy-update path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness checker for the
 * HIP product path and the timed CPU baseline. Only the test suite, the
 * smoke check and the CPU-baseline leg of bench.py may load it. The product
 * library never links, loads or calls anything here.
 *
 * Every function follows the reference semantics (file:line cited at each
 * definition in cublox_oracle.cpp) with plain FP32 arithmetic evaluated left
 * to right, compiled with -ffp-contract=off and without -ffast-math.
 *
 * Parity pins (what checks this restatement, see tests/test_oracle_*.py):
 *   index math     — the reference's five round-trip invariants with its three
 *                    fixtures, plus exhaustive tiny-map checks against Python
 *                    integer modular arithmetic;
 *   DDA            — brute-force segment/voxel intersection on tiny grids,
 *                    6-connectivity, L1-monotone walk, axis-aligned runs,
 *                    skip/clip special cases;
 *   apply          — exhaustive closed form over small (stored, op, hit);
 *   recenter       — brute-force window-membership model of the slab plan;
 *   fetch          — brute-force enumeration of the window (multiset);
 *   transform      — closed form R·p + t in a fixed operand order;
 *   ingest         — hand-worked decode/filter/back-projection cases (exact
 *                    special values, range boundary, the 1 M cap, field
 *                    offsets, invalid-pixel rules) in tests/test_oracle_ingest.py.
 * The reference CUDA cannot be compiled here (no nvcc, CUDA headers, Eigen),
 * so there is no oracle/_ref build.
 */
#ifndef CUBLOX_ORACLE_H
#define CUBLOX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- index math (utils.cuh:24-176) ---- */
int orc_pos_to_global(float p, float inv_res, int origin_at_center);
float orc_global_to_pos(int g, float res, int origin_at_center);
int orc_global_to_local(int g, int map, int half);
int orc_local_to_global(int l, int map, int half, int origin_g);
int orc_local_to_hash(const int l[3], const int map[3], const int half[3]);
void orc_hash_to_local(int h, const int map[3], const int half[3], int l[3]);
int orc_global_to_hash(const int g[3], const int map[3], const int half[3]);
float orc_logit(float p);

/* ---- one DDA ray, for brute-force checks (kernels.cu:39-151) ----
 * Writes the hash ids the reference kernel would increment, in order, into
 * out_h (the endpoint last, flagged by *endpoint_counted), up to cap entries.
 * Returns the number of op increments (may exceed cap), or -1 if the ray is
 * skipped. *clipped is set to whether the ray was range-clipped. */
int orc_trace_ray(const float origin[3], const float endpoint[3],
                  const int origin_i[3], const int map[3], const int half[3],
                  float res, float inv_res, float max_range, int map_vox_num,
                  int origin_at_center, int *out_h, int cap, int *clipped,
                  int *endpoint_counted);

/* ---- recenter plan (kernels.cu:313-398) ----
 * Returns action (0/1/2). new_origin[3] gets the planned origin. slices[a]
 * receives the raw slice list for axis a (exits, then fixed-local entries, in
 * reference order; duplicates kept); n_slices[a] its length. Each slices[a]
 * buffer must hold 2*max(map) ints (an axis may shift by up to the
 * largest extent minus one). */
int orc_plan_recenter(const float pos[3], const float origin_f[3],
                      const int origin_i[3], const int map[3],
                      const int half[3], float inv_res, float thr_sq,
                      int origin_at_center, int new_origin[3], int *slices_x,
                      int *slices_y, int *slices_z, int n_slices[3]);

/* ---- stateful grid mirroring OccupancyGrid.cpp host logic ---- */
typedef struct orc_grid orc_grid;

orc_grid *orc_grid_create(int hx, int hy, int hz, float res,
                          int origin_at_center, int has_thr, double thr,
                          float ox, float oy, float oz);
void orc_grid_destroy(orc_grid *g);
void orc_grid_set_threads(orc_grid *g, int n_threads);
void orc_grid_set_max_range(orc_grid *g, float r);
void orc_grid_set_log_odds(orc_grid *g, float l_hit, float l_miss, float l_min,
                           float l_max, float l_free, float l_occupied);
/* OccupancyGrid::update (OccupancyGrid.cpp:181-238): world-frame points. */
void orc_grid_update(orc_grid *g, const float *x, const float *y,
                     const float *z, int n, const float sensor_origin[3]);
/* Scan transform (main.cpp:640-693 minus the driver's range drop) then
 * update. T is row-major 3x4 [R | t] taking sensor frame to world frame; the
 * sensor origin is t. Non-finite points are dropped as the driver does. */
void orc_grid_update_sensor(orc_grid *g, const float *xs, const float *ys,
                            const float *zs, int n, const float T[12]);
/* The two halves of update, for stage-wise comparison. raycast leaves the
 * counters populated; apply folds and zeroes them. */
void orc_grid_raycast_world(orc_grid *g, const float *x, const float *y,
                            const float *z, int n, const float sensor_origin[3]);
void orc_grid_apply(orc_grid *g);
/* OccupancyGrid::recenter (OccupancyGrid.cpp:141-179). Returns the action. */
int orc_grid_recenter(orc_grid *g, const float pos[3]);
/* OccupancyGrid::fetchOccupancyAround (OccupancyGrid.cpp:240-356).
 * filter: -1 all, else VoxelState. Writes up to cap samples; returns the
 * number the reference would return. Order: x-major, y, z (deterministic). */
int orc_grid_fetch(const orc_grid *g, const float center[3], float radius,
                   int filter, int max_results, float *out_pos,
                   uint8_t *out_state, int cap);
void orc_grid_reset(orc_grid *g);
void orc_grid_reset_voxel(orc_grid *g, int hash_id);

float *orc_grid_occ(orc_grid *g);
int *orc_grid_op(orc_grid *g);
int *orc_grid_hit(orc_grid *g);
int orc_grid_voxel_num(const orc_grid *g);
void orc_grid_origin(const orc_grid *g, int origin_i[3], float origin_f[3]);
void orc_grid_load_occ(orc_grid *g, const float *occ);

/* ---- driver ingest (main.cpp:558-693) ---- */
/* buildCloudInOdomFrame (main.cpp:640-693) over a PointCloud2 payload of
 * n_records records of point_step bytes with float32 x/y/z at byte offsets
 * off_x/off_y/off_z: the first min(n_records, 1000000) records, non-finite
 * and ||p||^2 > max_range_sq dropped, the rest transformed by T (row-major
 * 3x4 [R | t]) into xw/yw/zw (capacity min(n_records, 1000000)). Returns
 * the rows written; 0 means the driver skips the update. */
int orc_build_cloud(const uint8_t *data, long long n_records, int point_step,
                    int off_x, int off_y, int off_z, const float T[12],
                    float max_range_sq, float *xw, float *yw, float *zw);
/* depthCallback (main.cpp:558-638): width*height sensor-frame points in
 * row-major order, NaN for invalid pixels. encoding 0: 16-bit millimetres,
 * 1: 32-bit float metres. K = {fx, fy, cx, cy}; step = row pitch in bytes. */
void orc_depth_to_cloud(const uint8_t *img, int width, int height,
                        long long step, int encoding, const float K[4],
                        float *xs, float *ys, float *zs);

/* publishOccupancyCloudAsync (main.cpp:740-817) over fetched samples
 * (pos: 3 floats each, the fetch's order): in order, while fewer than
 * max_pts are written, a sample is dropped by the global-index subsample
 * (main.cpp:71-84), else written to out4 as (x, y, z, rgb) with rgb the
 * distance colour of main.cpp:275-304 packed into a float. Returns the
 * records written. */
int orc_viz_points(const float *pos, int n, const float robot[3], float d_max,
                   float inv_res, int origin_at_center, int step, int max_pts,
                   float *out4);

/* Transform only (a1), used by the transform tests. */
void orc_transform_points(const float T[12], const float *xs, const float *ys,
                          const float *zs, int n, float *xw, float *yw,
                          float *zw);

#ifdef __cplusplus
}
#endif

#endif
