/*
 * cublox C ABI — the host/device boundary of the MI355X occupancy path.
 *
 * This header replaces the reference launcher seam include/cublox/kernels.hpp
 * (C++ linkage, CUDA vector types, cudaStream_t) with plain C: POD structs of
 * float/int fields, an opaque stream handle, and an int status on every
 * call. It includes no HIP or CUDA header, so host C++ built by g++ (or any
 * FFI: ctypes, cgo, JNI) can bind it. All kernels are hand-written HIP for
 * gfx950 in libcublox_hip.so.
 *
 * Conventions
 *  - Launchers never allocate device memory; the caller owns every buffer
 *    (kernels.hpp:52-53). They are asynchronous on `stream` (NULL = the
 *    device's null stream) and enqueue nothing for empty work.
 *  - Invalid arguments that the reference launchers silently ignore
 *    (kernels.cu:223-225, 244-246, 288-290, 570-574, 599-601, 633-636) return
 *    CUBLOX_OK with nothing enqueued, so the C++ layer keeps the reference's
 *    no-op behaviour. A HIP runtime failure returns CUBLOX_ERR_RUNTIME and
 *    cublox_last_error() holds the message for the calling thread.
 *  - Per-launch configuration travels by value in kernel arguments, never
 *    through module-global __constant__ symbols, so launches for different
 *    grids on different streams do not race (reference kernels.cu:21, 301,
 *    504 share one symbol per kernel).
 */
#ifndef CUBLOX_CAPI_H
#define CUBLOX_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUBLOX_OK 0
#define CUBLOX_ERR_INVALID 1
#define CUBLOX_ERR_RUNTIME 2
#define CUBLOX_ERR_CAPACITY 3

typedef void *cublox_stream; /* hipStream_t; NULL selects the null stream */
typedef void *cublox_event;  /* hipEvent_t */

typedef struct {
  float x, y, z;
} cublox_float3;
typedef struct {
  int x, y, z;
} cublox_int3;

/* ------------------------------------------------------------------ */
/* Errors and runtime primitives. The reference host library calls the CUDA
 * runtime directly (OccupancyGrid.cpp:206-211, 319-345, 441-446); host code
 * here must not see HIP headers, so these wrap the HIP runtime.           */
/* ------------------------------------------------------------------ */
const char *cublox_last_error(void);
int cublox_device_count(int *count);
int cublox_set_device(int device);
int cublox_get_device(int *device);
int cublox_device_name(char *buf, int cap);
int cublox_mem_info(size_t *free_bytes, size_t *total_bytes);
int cublox_malloc(void **ptr, size_t bytes);
int cublox_free(void *ptr);
int cublox_host_alloc(void **ptr, size_t bytes); /* pinned host memory */
int cublox_host_free(void *ptr);
/* The device address of pinned host memory from cublox_host_alloc: kernels
 * given it read the host memory over the bus, with no copy to HBM first. */
int cublox_host_device_pointer(void **dptr, void *hptr);
int cublox_memset_async(void *ptr, int value, size_t bytes, cublox_stream s);
int cublox_memcpy_h2d_async(void *dst, const void *src, size_t bytes,
                            cublox_stream s);
int cublox_memcpy_d2h_async(void *dst, const void *src, size_t bytes,
                            cublox_stream s);
int cublox_memcpy_d2d_async(void *dst, const void *src, size_t bytes,
                            cublox_stream s);
int cublox_stream_create(cublox_stream *s);
/* A stream on a hardware queue of its own (no reference counterpart): a
 * CU-masked stream (hipExtStreamCreateWithCUMask) with every CU enabled.
 * Plain streams share the runtime's pool of hardware queues, and two
 * streams on one queue run one after the other; GridBatch's overlapped
 * ticks take these. */
int cublox_stream_create_dedicated(cublox_stream *s);
int cublox_stream_destroy(cublox_stream s);
int cublox_stream_synchronize(cublox_stream s);
int cublox_device_synchronize(void);
int cublox_event_create(cublox_event *e);
int cublox_event_destroy(cublox_event e);
int cublox_event_record(cublox_event e, cublox_stream s);
int cublox_event_synchronize(cublox_event e);
int cublox_event_elapsed_ms(cublox_event start, cublox_event stop, float *ms);
/* Work enqueued on s after this call waits for the work before the last
 * cublox_event_record of e (hipStreamWaitEvent). */
int cublox_stream_wait_event(cublox_stream s, cublox_event e);
/* Zero `count` floats/ints with a fill kernel (safe inside graph capture,
 * unlike a memset node). */
int cublox_fill_zero_async(void *ptr, size_t count32, cublox_stream s);

/* HIP graphs: capture the work enqueued on a stream, replay it. */
typedef void *cublox_graph; /* hipGraphExec_t */
int cublox_graph_begin_capture(cublox_stream s);
int cublox_graph_end_capture(cublox_stream s, cublox_graph *exec);
int cublox_graph_launch(cublox_graph exec, cublox_stream s);
int cublox_graph_destroy(cublox_graph exec);

/* ------------------------------------------------------------------ */
/* Ray cast. Replaces RayCastCfg (kernels.hpp:33-44).                   */
/* ------------------------------------------------------------------ */
typedef struct {
  cublox_float3 origin;         /* sensor position, world frame */
  cublox_int3 origin_i;         /* sliding-window origin, global voxel */
  cublox_int3 map_size_i;       /* voxels per axis (2*half+1) */
  cublox_int3 half_map_size_i;
  float resolution;
  float inv_resolution;
  float max_range;              /* longer rays are clipped, free-space only */
  int map_vox_num;
  int origin_at_center;         /* 0 corner binning, 1 centre binning */
} cublox_raycast_cfg;

/* launchRayCastUpdate (kernels.hpp:54-57): world-frame SoA endpoints. One
 * Amanatides-Woo walk per point; op_cnt[h] += 1 per traversed in-window
 * voxel, endpoint op += 1 and hit += 1 unless range-clipped. */
int cublox_launch_ray_cast_update(const cublox_raycast_cfg *cfg,
                                  const float *d_x, const float *d_y,
                                  const float *d_z, int n, cublox_stream s,
                                  int *d_op_cnt, int *d_hit_cnt);

/* Scan transform fused into the ray cast (reference: main.cpp:640-693 builds
 * the world-frame cloud on the host, then kernels.hpp:54-57). Points are in
 * the sensor frame; T is row-major 3x4 [R | t] sensor->world, evaluated as
 * ((R0*x + R1*y) + R2*z) + t per row. Non-finite points are skipped, as the
 * driver drops them (main.cpp:672-674). cfg->origin must equal t. */
int cublox_launch_ray_cast_update_sensor(const cublox_raycast_cfg *cfg,
                                         const float T[12], const float *d_xs,
                                         const float *d_ys, const float *d_zs,
                                         int n, cublox_stream s,
                                         int *d_op_cnt, int *d_hit_cnt);

/* ------------------------------------------------------------------ */
/* Apply. launchApplyUpdate (kernels.hpp:69-76).                         */
/* ------------------------------------------------------------------ */
/* For every voxel with op != 0: stored = clamp(stored + l_unknown
 * + l_miss*(op-hit) + l_hit*hit, l_min, l_max) - l_unknown, then zero both
 * counters. If all three modified_* buffers are given and voxel_num <=
 * modified_capacity, the (hash, new value) of every voxel whose stored value
 * changed is appended (unordered) and *modified_count counts them; the
 * caller zeroes *modified_count first. */
int cublox_launch_apply_update(float *d_occ, int *d_op_cnt, int *d_hit_cnt,
                               int voxel_num, float l_hit, float l_miss,
                               float l_min, float l_max, float l_unknown,
                               cublox_stream s, unsigned int *d_modified_count,
                               int *d_modified_idx, float *d_modified_val,
                               unsigned int modified_capacity);

/* launchClearVoxelsByIndex (kernels.hpp:79-80): occ[idx[i]] = 0 for valid
 * indices. */
int cublox_launch_clear_voxels_by_index(float *d_occ, const int *d_indices,
                                        int n, int voxel_num, cublox_stream s);

/* ------------------------------------------------------------------ */
/* Recenter. RecenterCfg / RecenterResult (kernels.hpp:83-100).          */
/* ------------------------------------------------------------------ */
typedef struct {
  cublox_float3 pos;
  cublox_float3 origin_f;
  cublox_int3 origin_i;
  cublox_int3 map_size_i;
  cublox_int3 half_map_size_i;
  float resolution;
  float inv_resolution;
  float recenter_threshold_sq;
  int origin_at_center;
  int map_vox_num;
} cublox_recenter_cfg;

typedef struct {
  int action; /* 0 none, 1 clear slabs, 2 full reset */
  cublox_int3 new_origin_i;
  int n_slices[3];
  long long zeroed_voxels; /* |union of cleared voxels| (0 for action 0) */
} cublox_recenter_result;

/* The slab plan of planRecenterKernel (kernels.cu:313-398), on the host: it
 * is integer math on the config, so no device round trip is needed. Each
 * slices_* array (may be NULL) receives the raw list for that axis in
 * reference order and must hold 2*max(map_size_i.x, .y, .z) ints: one
 * axis may shift by up to the largest extent minus one before the plan
 * becomes a full reset (kernels.cu:344-347). */
int cublox_plan_recenter(const cublox_recenter_cfg *cfg, int *slices_x,
                         int *slices_y, int *slices_z,
                         cublox_recenter_result *result);

/* launchRecenter (kernels.hpp:104-107): plan on the host, then one
 * coalesced slab-zeroing launch (or a full reset for action 2). No device
 * sync and no device->host copy. Counters are zeroed too when non-NULL. */
int cublox_launch_recenter(float *d_occ, int *d_op_cnt, int *d_hit_cnt,
                           const cublox_recenter_cfg *cfg, cublox_stream s,
                           cublox_recenter_result *result);

/* A recenter's clear staged for later (no reference counterpart): the
 * plan of cublox_launch_recenter made on the host into `plan` (host memory
 * of cublox_recenter_plan_bytes() bytes) without launching anything.
 * cublox_launch_recenter_plan then clears occ (and the counters when
 * non-NULL) on s, as cublox_launch_recenter would have; or the plan rides
 * in the next fused update (cublox_integrate_grid.zero_plan), whose walk
 * launch clears the slabs with workgroups of its own, beside the walk.   */
size_t cublox_recenter_plan_bytes(void);
int cublox_recenter_plan_stage(const cublox_recenter_cfg *cfg, void *plan,
                               cublox_recenter_result *result);
int cublox_launch_recenter_plan(const void *plan, float *d_occ, int *d_op_cnt,
                                int *d_hit_cnt, cublox_stream s);

/* launchClearRecenterSlabsForAxis (kernels.hpp:112-117): zero every voxel
 * whose local index on `axis` is listed in the DEVICE array d_slices. */
/* Batched recenter (no reference counterpart): the plans of many grids,
 * made on the host as cublox_plan_recenter does, zeroed in one launch.
 * Stage (host only; results[k] is grid k's plan outcome, as
 * cublox_launch_recenter reports it), then launch from the staging buffer
 * with d_area (device, cublox_recenter_batch_staging_bytes bytes) as the
 * device copy. The launch shape depends on max_regions only, so it can be
 * captured in a HIP graph. Counters are not cleared (they are zero
 * between updates). */
typedef struct {
  cublox_recenter_cfg cfg;
  float *d_occ;
} cublox_recenter_item;

size_t cublox_recenter_batch_staging_bytes(int max_regions);
int cublox_recenter_batch_stage(const cublox_recenter_item *items, int n,
                                int max_regions, void *h_staging,
                                cublox_recenter_result *results);
int cublox_launch_recenter_batch_staged(const void *h_staging, void *d_area,
                                        int max_regions, cublox_stream s);
/* The two halves of cublox_launch_recenter_batch_staged: the staging copy
 * to the device area, and the zeroing launch that reads it (for callers
 * that move the copy ahead of other work on the stream). */
int cublox_recenter_batch_upload(const void *h_staging, void *d_area,
                                 int max_regions, cublox_stream s);
int cublox_launch_recenter_batch_uploaded(const void *d_area, int max_regions,
                                          cublox_stream s);

int cublox_launch_clear_recenter_slabs_for_axis(
    float *d_occ, int *d_op_cnt, int *d_hit_cnt, cublox_int3 map_size_i,
    cublox_int3 half_map_size_i, const int *d_slices, int n_slices, int axis,
    cublox_stream s);

/* ------------------------------------------------------------------ */
/* Sphere fetch. OccupancyFetchCfg (kernels.hpp:120-136).                */
/* ------------------------------------------------------------------ */
typedef struct {
  cublox_int3 center_g;
  cublox_int3 origin_i;
  cublox_int3 map_size_i;
  cublox_int3 half_map_size_i;
  int gz_min, gz_max;   /* inclusive global z range */
  int disk_radius_vox;  /* ceil(r/res) */
  cublox_float3 center; /* world [m] */
  float radius_sq;
  float resolution;
  float l_free;
  float l_occupied;
  float l_unknown;
  int origin_at_center;
  int state_filter; /* -1 all; 0 unknown; 1 free; 2 occupied */
} cublox_fetch_cfg;

/* Host half of OccupancyGrid::fetchOccupancyAround (OccupancyGrid.cpp:
 * 244-317): clamps the radius to the horizontal window margin, clamps the
 * AABB to the window and fills *cfg. Returns 1 when there is work, 0 when
 * the reference would return an empty result. *est_voxels gets the disk
 * volume; *output_cap the reference's buffer cap (max_results or 500000). */
int cublox_fetch_plan(cublox_int3 origin_i, cublox_int3 map_size_i,
                      cublox_int3 half_map_size_i, float resolution,
                      float inv_resolution, int origin_at_center, float l_free,
                      float l_occupied, float l_unknown, cublox_float3 center,
                      float radius, int state_filter, int max_results,
                      cublox_fetch_cfg *cfg, long long *est_voxels,
                      unsigned int *output_cap);

/* launchFetchOccupancyAround (kernels.hpp:142-147). out_pos holds 3 floats
 * per sample. *d_out_count must be zeroed by the caller. Output order is
 * unspecified (as in the reference). */
int cublox_launch_fetch_occupancy_around(const cublox_fetch_cfg *cfg,
                                         const float *d_occ,
                                         unsigned int *d_out_count,
                                         float *d_out_pos,
                                         unsigned char *d_out_state,
                                         unsigned int output_capacity,
                                         unsigned int max_results,
                                         cublox_stream s);

/* Occupancy viz cloud (the driver's publishOccupancyCloudAsync,
 * main.cpp:740-817, over its OCCUPIED fetch, main.cpp:487-491; no library
 * counterpart): one launch fetches as cublox_launch_fetch_occupancy_around
 * with max_results = max_points (cfg from cublox_fetch_plan with the
 * OCCUPIED filter), drops fetched samples by the global-index subsample
 * (main.cpp:71-84) and writes the rest as (x, y, z, rgb) float records,
 * rgb the driver's distance colour packed into a float (main.cpp:275-304).
 * *d_fetch_count (the fetch's logical count; the driver's sample count is
 * min(it, max_points)) and *d_viz_count (records written, past
 * out_capacity not stored) are set by the launch. Record order is
 * unspecified, as the fetch order is in the reference.                   */
typedef struct {
  cublox_float3 robot_pos;  /* distances are measured from it             */
  float max_viz_radius;     /* colour-map range (values below 1 act as 1) */
  float inv_resolution;     /* grid binning for the subsample's index     */
  int origin_at_center;
  int subsample;            /* keep where (gx ^ gy ^ gz) % subsample == 0 */
  int max_points;           /* the fetch's max_results (<= 0: none)       */
} cublox_viz_cfg;
int cublox_launch_viz_cloud(const cublox_fetch_cfg *cfg, const float *d_occ,
                            const cublox_viz_cfg *viz,
                            unsigned int fetch_capacity,
                            unsigned int *d_fetch_count,
                            unsigned int *d_viz_count, float *d_out,
                            unsigned int out_capacity, cublox_stream s);

/* Batched fetch: n_queries configs in DEVICE memory, one launch. Query q
 * writes at most per_query_cap samples at [q*per_query_cap, ...) and its
 * logical count to d_counts[q] (zeroed by the caller). */
int cublox_launch_fetch_occupancy_batch(const cublox_fetch_cfg *d_cfgs,
                                        int n_queries, int max_disk_diam,
                                        int max_nz, const float *d_occ,
                                        unsigned int *d_counts,
                                        float *d_out_pos,
                                        unsigned char *d_out_state,
                                        unsigned int per_query_cap,
                                        cublox_stream s);

/* Dense gather after a batched fetch: query q's first min(d_counts[q],
 * per_query_cap) samples are copied from its segment to position
 * d_offsets[q] of the dense outputs (offsets computed by the caller, e.g.
 * an exclusive scan of the capped counts). max_count bounds the capped
 * counts (sizes the launch). */
int cublox_launch_fetch_gather(const unsigned int *d_counts,
                               const unsigned int *d_offsets, int n_queries,
                               unsigned int per_query_cap,
                               unsigned int max_count, const float *d_pos_seg,
                               const unsigned char *d_state_seg,
                               float *d_pos_out, unsigned char *d_state_out,
                               cublox_stream s);

/* ------------------------------------------------------------------ */
/* Fused scan update (no reference counterpart as one call; it replaces the
 * pair launchRayCastUpdate + launchApplyUpdate, kernels.hpp:54-76, with the
 * same per-voxel result). Rays are cut into segments at 16x16x32 tiles of
 * the hash-ordered grid; each tile's counts accumulate in LDS and are folded
 * into occ once, so no per-voxel counter arrays are needed. The first
 * near_steps steps of every ray (voxels within that L1 distance of the
 * sensor voxel, near_steps in [1, 12]) accumulate in a per-workgroup LDS
 * cube around the sensor voxel, flushed once per workgroup. The map size
 * must be 2 * half + 1 on every axis. The caller owns the workspace: size it with
 * cublox_integrate_workspace_bytes for at least max_points points, zero it
 * once with cublox_integrate_workspace_init, and reuse it for launches with
 * the same cfg geometry, max_range, max_points and near_steps (the kernels
 * leave it ready for the next launch). T (row-major 3x4 sensor->world) may
 * be NULL for world-frame points with cfg->origin as the sensor. When
 * d_export_op/d_export_hit are given (both zeroed by the caller), the
 * per-voxel counts are written there instead of being folded into occ.   */
/* ------------------------------------------------------------------ */
size_t cublox_integrate_workspace_bytes(const cublox_raycast_cfg *cfg,
                                        int max_points, int near_steps);
int cublox_integrate_workspace_init(void *d_ws, size_t bytes, cublox_stream s);
int cublox_launch_integrate(const cublox_raycast_cfg *cfg, const float *T,
                            const float *d_xs, const float *d_ys,
                            const float *d_zs, int n, float *d_occ,
                            float l_hit, float l_miss, float l_min,
                            float l_max, float l_unknown, void *d_ws,
                            size_t ws_bytes, int max_points, int near_steps,
                            int *d_export_op, int *d_export_hit,
                            cublox_stream s);

/* Fused update that allocates nothing. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It is
 * cublox_launch_integrate with the workspace arguments removed: the
 * workspace is a library-owned device array of CUBLOX_NOALLOC_SCRATCH_BYTES
 * per device, held from the library's load to the end of the process, with
 * at most 64 hot-tile slabs (hotter tiles run unsplit; results are the
 * same). CUBLOX_ERR_CAPACITY (enqueuing nothing) when the workspace for n
 * points on this grid exceeds the array: 256 MiB holds a 65,536-point scan
 * of a 701x701x151 grid at 0.1 m with a 30 m range. Every call on one
 * device shares that device's array, so calls that overlap on one device
 * (on different streams or threads) race; calls in one stream's order do
 * not. The counters the kernels expect at zero are cleared by fill kernels
 * at the start of each call.                                             */
#define CUBLOX_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int cublox_launch_integrate_noalloc(const cublox_raycast_cfg *cfg, const float *T,
                                    const float *d_xs, const float *d_ys,
                                    const float *d_zs, int n, float *d_occ,
                                    float l_hit, float l_miss, float l_min,
                                    float l_max, float l_unknown, int near_steps,
                                    int *d_export_op, int *d_export_hit,
                                    cublox_stream s);

/* Fused update with a changed-voxel stream (the reference's modified-voxel
 * compaction in applyUpdateKernel, kernels.cu:202-209, gated off at
 * kernels.cu:251-262, here enabled for the fused update): for every voxel
 * whose stored value changes, (hash, new log-odds value) is appended to
 * d_delta_idx/d_delta_val in unspecified order, and *d_delta_count counts
 * them all (the launch sets it; the caller need not zero it). Entries at
 * positions >= capacity
 * are not written (count > capacity means the stream is truncated);
 * capacity 0 with NULL arrays counts only. NULL d_delta_count: the same as
 * cublox_launch_integrate. Not with a count export.                      */
int cublox_launch_integrate_delta(
    const cublox_raycast_cfg *cfg, const float *T, const float *d_xs,
    const float *d_ys, const float *d_zs, int n, float *d_occ, float l_hit,
    float l_miss, float l_min, float l_max, float l_unknown, void *d_ws,
    size_t ws_bytes, int max_points, int near_steps, int *d_export_op,
    int *d_export_hit, unsigned int *d_delta_count, int *d_delta_idx,
    float *d_delta_val, unsigned int delta_capacity, cublox_stream s);

/* ------------------------------------------------------------------ */
/* Sensor ingest (the driver's input side, reference src/main.cpp; no
 * counterpart in the library API).                                       */
/* ------------------------------------------------------------------ */
/* Depth image back-projection (depthCallback, main.cpp:558-638): the first
 * min(width*height, max_points) pixels in row-major order become
 * sensor-frame points x = ((u - cx) * z) / fx, y = ((v - cy) * z) / fy, z.
 * encoding 0: 16-bit depth in millimetres (16UC1, mono16; z = raw * 0.001,
 * raw 0 invalid); 1: 32-bit float metres (32FC1; non-finite or <= 0
 * invalid). Invalid pixels give quiet-NaN points, which the update drops.
 * K = {fx, fy, cx, cy}; step_bytes is the image row pitch.               */
int cublox_launch_depth_to_points(const void *d_image, int width, int height,
                                  size_t step_bytes, int encoding,
                                  const float K[4], int max_points,
                                  float *d_xs, float *d_ys, float *d_zs,
                                  cublox_stream s);
/* Census of the driver's point filter (main.cpp:672-679): *d_count (set by
 * the launch) = points that are finite and, for range_drop_sq > 0, have
 * x*x + y*y + z*z <= range_drop_sq. Points laid out as in
 * cublox_integrate_grid (point_stride 0: 1). buildCloudInOdomFrame returns
 * false, and the driver skips the update, when this is 0 (main.cpp:686-690). */
int cublox_launch_count_kept_points(const float *d_xs, const float *d_ys,
                                    const float *d_zs, int point_stride, int n,
                                    float range_drop_sq, unsigned int *d_count,
                                    cublox_stream s);

/* ------------------------------------------------------------------ */
/* Map snapshots (no reference counterpart: the reference cannot save or
 * restore a grid). The sparse form of an occupancy buffer is the (hash,
 * value) pairs of its voxels with a nonzero bit pattern, ascending hash.  */
/* ------------------------------------------------------------------ */
size_t cublox_sparse_scratch_bytes(int voxel_num);
/* *d_count (set by the launch) = the nonzero voxels; the first
 * min(count, capacity) pairs go to d_ids/d_vals (capacity 0: count only).
 * d_scratch holds cublox_sparse_scratch_bytes(voxel_num) bytes.          */
int cublox_launch_sparse_compact(const float *d_occ, int voxel_num,
                                 void *d_scratch, size_t scratch_bytes,
                                 int *d_ids, float *d_vals,
                                 unsigned int capacity, unsigned int *d_count,
                                 cublox_stream s);
/* occ := 0, then occ[ids[k]] = vals[k] for k < n (ids outside [0, V)
 * skipped).                                                              */
int cublox_launch_sparse_restore(float *d_occ, int voxel_num, const int *d_ids,
                                 const float *d_vals, unsigned int n,
                                 cublox_stream s);

/* Per-ray segment bound and tile count of a grid geometry (for sizing
 * batch limits). */
int cublox_integrate_segment_bound(const cublox_raycast_cfg *cfg);
int cublox_integrate_tile_count(const cublox_raycast_cfg *cfg);

/* ------------------------------------------------------------------ */
/* Batched fused update (no reference counterpart; the reference updates
 * one grid per call, kernels.hpp:54-76). One launch sequence updates a
 * batch of grids, each with its own geometry, window, sensor pose, points,
 * occupancy buffer and log-odds, with the same per-grid result as
 * cublox_launch_integrate on that grid alone. The launch shapes depend on
 * the limits only and the per-grid descriptors travel in one host->device
 * copy at its head, so cublox_launch_integrate_batch_staged can be
 * captured once in a HIP graph and replayed per tick after restaging.
 * The staging buffer must be pinned host memory (cublox_host_alloc) when
 * the launch is captured or asynchronous with respect to restaging.      */
/* ------------------------------------------------------------------ */
typedef struct {
  int max_grids;            /* grids per launch                           */
  int max_points;           /* points per launch, all grids               */
  int max_tiles;            /* sum of cublox_integrate_tile_count         */
  int max_segments_per_ray; /* max of cublox_integrate_segment_bound      */
  int max_near_steps;       /* largest near_steps of any grid, 1..12      */
  int chunk;                /* smallest records per hot-tile work item
                               any launch will use (0: 1024)             */
  int max_slabs;            /* hot-tile scratch slabs, 32 KB each (0:
                               65536); hot tiles past them run unsplit   */
  int delta;                /* nonzero: grids may carry changed-voxel
                               streams (selects the folds that feed them) */
} cublox_batch_limits;

typedef struct {
  cublox_raycast_cfg cfg;   /* geometry, window; origin = sensor if !use_T */
  float T[12];              /* row-major 3x4 sensor->world                 */
  int use_T;                /* points are sensor-frame                     */
  const float *d_xs, *d_ys, *d_zs;
  int n;
  float *d_occ;
  float l_hit, l_miss, l_min, l_max, l_unknown;
  int near_steps;           /* 1..max_near_steps (0: 8)                    */
  int *d_export_op, *d_export_hit; /* optional: counts instead of a fold  */
  /* optional changed-voxel stream, as in cublox_launch_integrate_delta
   * (needs limits.delta)                                                 */
  unsigned int *d_delta_count;
  int *d_delta_idx;
  float *d_delta_val;
  unsigned int delta_capacity;
  /* Point layout (the driver's ingest, main.cpp:640-693): point k is
   * (d_xs[k*s], d_ys[k*s], d_zs[k*s]) with s = point_stride (0: 1, separate
   * columns). A packed PointCloud2 payload of point_step bytes with float32
   * x/y/z at byte offsets ox/oy/oz is d_xs = data + ox (and so on) with
   * point_stride = point_step / 4 (all four multiples of 4).              */
  int point_stride;
  /* > 0 (sensor-frame points only): points with x*x + y*y + z*z above it
   * are dropped before the transform (main.cpp:676-679).                  */
  float range_drop_sq;
  /* Optional (single-grid launches): an event the kernels that write occ
   * (or the count exports) wait for; the walk, binning and tile scan, which
   * read neither, start before it completes. A recenter's slab clear on
   * another stream then overlaps the next update's ray walk. NULL: none. */
  cublox_event occ_ready;
  /* Optional (single-grid launches): a recenter's clear staged with
   * cublox_recenter_plan_stage for this grid and not yet launched. The
   * update clears it first: its walk launch takes extra workgroups that
   * zero the slabs of d_occ (and of d_zero_op / d_zero_hit when non-NULL)
   * while the walk runs; the folds come after both. NULL: none.           */
  const void *zero_plan;
  int *d_zero_op, *d_zero_hit;
} cublox_integrate_grid;

/* Single-grid fused update from a batch item (every option above; no
 * reference counterpart as one call). Workspace as for
 * cublox_launch_integrate (near_steps 0 means 8).                        */
int cublox_launch_integrate_grid(const cublox_integrate_grid *grid, void *d_ws,
                                 size_t ws_bytes, int max_points,
                                 cublox_stream s);

size_t cublox_integrate_batch_workspace_bytes(const cublox_batch_limits *lim);
size_t cublox_integrate_batch_staging_bytes(const cublox_batch_limits *lim);
/* Host only: writes the batch's descriptors into h_staging. chunk > 0
 * sets this launch's records per hot-tile item (>= lim->chunk), else the
 * limits' chunk; results do not depend on it. */
int cublox_integrate_batch_stage(const cublox_integrate_grid *grids,
                                 int n_grids, const cublox_batch_limits *lim,
                                 int chunk, void *h_staging);
/* The per-tick settings of a batch (none changes any result; they travel
 * in the staged header, so a captured launch sequence follows them):
 *   chunk          records per hot-tile work item (0: the limits' chunk);
 *   lanes_per_ray  1 (64 rays per wavefront, a binning pass counts the
 *                  records per tile) or 4 (16 rays per wavefront: each
 *                  ray's quad shares the per-axis work of a span and counts
 *                  its workgroup's records per tile in LDS); 0: 4 up to
 *                  262,144 rays of capacity, else 1;
 *   hot_mode       how the chunks of a hot tile combine their LDS counts:
 *                  0 a slab per chunk, summed by the hot-tile fold; 1 one
 *                  shared op/hit slab pair per tile, added to with global
 *                  atomics for the nonzero counts only.
 * The staged launch takes the walk and hot-tile kernels the header names,
 * so a captured graph holds one (lanes, hot_mode) pair: the call sets
 * tuning->lanes_per_ray to the lanes the launch will use (a caller keeps
 * one graph per pair).                                                   */
typedef struct {
  int chunk;
  int lanes_per_ray;
  int hot_mode;
} cublox_batch_tuning;
int cublox_integrate_batch_stage_tuned(const cublox_integrate_grid *grids,
                                       int n_grids, const cublox_batch_limits *lim,
                                       cublox_batch_tuning *tuning, void *h_staging);
/* Copies h_staging to the workspace and launches the update sequence. */
int cublox_launch_integrate_batch_staged(const void *h_staging, void *d_ws,
                                         size_t ws_bytes,
                                         const cublox_batch_limits *lim,
                                         cublox_stream s);
/* The staged launch in parts (no reference counterpart): part 1 copies
 * h_staging to the workspace and launches the kernels that read and write
 * no occupancy (ray walk, record counting, tile scan, placing); part 2
 * launches the folds into the grids' occupancy (h_staging must still hold
 * the same tick); 3 is both, as cublox_launch_integrate_batch_staged. Part 1
 * of a tick may run on another stream and workspace beside part 2 of the
 * tick before it, which is how a caller overlaps consecutive ticks; every
 * grid's folds must stay in tick order (an event between the streams).   */
int cublox_launch_integrate_batch_staged_part(const void *h_staging, void *d_ws,
                                              size_t ws_bytes,
                                              const cublox_batch_limits *lim,
                                              int part, cublox_stream s);
/* stage + staged launch. */
int cublox_launch_integrate_batch(const cublox_integrate_grid *grids,
                                  int n_grids, const cublox_batch_limits *lim,
                                  void *h_staging, void *d_ws,
                                  size_t ws_bytes, cublox_stream s);

/* Diagnostics (no reference counterpart): after the last update launched on
 * stream s with this workspace, synchronises s and writes to out[0..4] the
 * work items, the tile records (segments and endpoints), the overflow flag
 * (sticky since cublox_integrate_workspace_init: nonzero once a ray needed
 * more records than its slot bound), the split hot tiles and the unsplit
 * hot tiles of that launch. The batch form takes
 * the limits the workspace was sized with; the single-grid form the
 * arguments of cublox_integrate_workspace_bytes.                         */
int cublox_integrate_batch_counters(const void *d_ws, size_t ws_bytes,
                                    const cublox_batch_limits *lim, int *out,
                                    cublox_stream s);
int cublox_integrate_counters(const void *d_ws, size_t ws_bytes,
                              const cublox_raycast_cfg *cfg, int max_points,
                              int near_steps, int *out, cublox_stream s);

/* ------------------------------------------------------------------ */
/* Counter census (no reference counterpart): over n voxels, adds to
 * d_out[0] the number of voxels with op != 0, to d_out[1] the sum of op, to
 * d_out[2] the sum of hit and to d_out[3] the number of voxels with
 * hit != 0. The caller zeroes d_out[0..3]. Used to account the algorithmic
 * work of a ray cast (voxel visits, unique voxels touched).              */
/* ------------------------------------------------------------------ */
int cublox_counter_census(const int *d_op_cnt, const int *d_hit_cnt, int n,
                          unsigned long long *d_out, cublox_stream s);

/* ------------------------------------------------------------------ */
/* Device index chain: the reference's device round trip over the index
 * math of utils.cuh (test/test_utils_index_roundtrip.cu:160-211, which runs
 * localIndexToGlobalIndex, globalIndexToLocalIndex, localIndexToHashId,
 * hashIdToLocalIndex, hashIdToGlobalIndex, localIndexToPos and
 * posToGlobalIndex in a kernel). Here the kernel writes every link of the
 * chain for each case, computed by the same device functions the kernels of
 * this library use, so a caller can compare each value with a host
 * computation rather than only count failures. d_cases and d_out are
 * device arrays of n entries.                                          */
/* ------------------------------------------------------------------ */
typedef struct {
  cublox_int3 local;      /* local index, each axis in [-half, half] */
  cublox_int3 map_size_i;
  cublox_int3 half_map_size_i;
  cublox_int3 origin_i;   /* window origin, global voxel */
  float resolution;
  float inv_resolution;
  int origin_at_center;
} cublox_index_case;

typedef struct {
  cublox_int3 global;            /* local -> global */
  cublox_int3 local_back;        /* global -> local */
  int hash_id;                   /* local -> hash */
  cublox_int3 local_from_hash;   /* hash -> local */
  cublox_int3 global_from_hash;  /* hash -> global */
  cublox_float3 pos;             /* local -> position */
  cublox_int3 global_from_pos;   /* position -> global */
} cublox_index_chain;

int cublox_launch_index_chain(const cublox_index_case *d_cases, int n,
                              cublox_index_chain *d_out, cublox_stream s);

#ifdef __cplusplus
}
#endif

#endif /* CUBLOX_CAPI_H */
