/* Flat C binding of cublox::MapServer (include/cublox/MapServer.hpp).
 *
 * No reference counterpart: the reference serves one grid per ROS node
 * (main.cpp:437-483 runs update + recenter per scan message). The server
 * runs many such grids ("sessions") on one GPU with HBM-budget admission
 * and continuous batching of their scans; each session's grid evolves
 * exactly as OccupancyGrid::updateSensorFrame + recenter +
 * fetchOccupancyAround issued alone in submission order.
 *
 * Status: 0 = ok; on failure the message is in cublox_og_last_error().   */
#ifndef CUBLOX_MS_CAPI_H
#define CUBLOX_MS_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cublox_ms cublox_ms;

/* limits: [0] HBM budget bytes for sessions, [1] max admitted sessions,
 * [2] max points per scan, [3] max points per tick, [4] max grids per
 * tick, [5] max tiles per tick, [6] max segments per ray (sessions whose
 * segment bound is larger are refused), [7] graph replay (0/1), [8] bytes
 * of the device arena holding queued scans' points (0: 4 GiB), [9] evict
 * idle sessions to host snapshots when a session with requests waits for
 * room (0/1), [10] pipelined ticks (0/1: consecutive ticks overlap on two
 * streams and workspaces), [11] records per hot-tile work item of the
 * ticks (0: 4096). NULL on failure. */
cublox_ms *cublox_ms_create(const long long limits[12]);
void cublox_ms_destroy(cublox_ms *h);
/* Device bytes a session of this half size is charged (4 B per voxel plus
 * the fetch output reserve). */
long long cublox_ms_session_bytes(const int half[3]);

/* Opens a session (admitted when it fits, else queued FIFO). log_odds:
 * l_hit, l_miss, l_min, l_max, l_free, l_occupied (NULL: defaults). */
int cublox_ms_open(cublox_ms *h, const int half[3], float resolution,
                   int origin_at_center, int has_threshold, double threshold,
                   float max_range, const float log_odds[6], int *sid);
int cublox_ms_close(cublox_ms *h, int sid);
int cublox_ms_admitted(cublox_ms *h, int sid, int *out);

/* Host columns (copied), sensor frame, T row-major 3x4 sensor->world; the
 * scan's grid recenters toward `base` after the update. */
int cublox_ms_submit_scan(cublox_ms *h, int sid, const float *xs,
                          const float *ys, const float *zs, int n,
                          const float T[12], const float base[3],
                          long long *id);
/* state_filter: -1 all, 0 unknown, 1 free, 2 occupied. */
int cublox_ms_submit_fetch(cublox_ms *h, int sid, const float center[3],
                           float radius, int state_filter, int max_results,
                           long long *id);
int cublox_ms_drain(cublox_ms *h);
/* While paused no new tick is formed; requests keep queueing. */
int cublox_ms_pause(cublox_ms *h);
int cublox_ms_resume(cublox_ms *h);

/* Up to cap completions: id, session, kind (0 scan, 1 fetch), times
 * (3 per entry: submit, tick start, done; ms since the server started),
 * tick index and result (scan: points; fetch: samples). */
int cublox_ms_take_completions(cublox_ms *h, long long *ids, int *sessions,
                               int *kinds, double *times, long long *ticks,
                               long long *results, int cap, int *n_out);
int cublox_ms_keep_fetch_results(cublox_ms *h, int on);
/* Takes a kept fetch result: *n = samples (at most cap written). */
int cublox_ms_fetch_result(cublox_ms *h, long long id, float *pos,
                           unsigned char *state, int cap, int *n);

/* After cublox_ms_drain: the session grid's occupancy (hash order) and
 * window origin (read from its snapshot while the session is evicted). */
int cublox_ms_copy_occupancy(cublox_ms *h, int sid, float *dst);
int cublox_ms_origin(cublox_ms *h, int sid, int origin_i[3]);
/* out: ticks, scans, fetches, points, device ms, resident bytes,
 * workspace bytes, admitted, waiting, ms since start, host ms between a
 * tick's end on the device and the next tick's launch (summed), evictions,
 * restores, host bytes of evicted sessions' snapshots. */
int cublox_ms_stats(cublox_ms *h, double out[14]);
/* MapServer::evictSession: the session's map moves to a host snapshot once
 * nothing of it is in flight, and is restored when it is readmitted. */
int cublox_ms_evict(cublox_ms *h, int sid);

/* Open-loop load (benchmark client): `clients` threads, each owning every
 * clients-th session, each with Poisson scan arrivals at its share of
 * scan_rate_hz, each arrival to a uniformly drawn session of its own (so
 * every session sees a Poisson stream), plus independent Poisson fetch
 * arrivals at fetch_ratio times that rate. Session k replays its template scan (xs[k], ys[k],
 * zs[k], n[k], T[12k..], base[3k..]) with the pose moved by `speed` m
 * along x per scan; fetches ask for a radius uniform in [fetch_radius,
 * fetch_radius_max] (fetch_radius when the max is not larger) around the
 * current sensor position. Returns once the last arrival of duration_s is
 * submitted (the caller drains); *submitted = requests submitted and
 * *max_lag_ms the largest delay of a submission behind its arrival time. */
typedef struct {
  double scan_rate_hz;
  double fetch_ratio;
  double duration_s;
  float fetch_radius;
  float speed;
  unsigned long long seed;
  float fetch_radius_max;
  int clients;
} cublox_ms_load_cfg;

int cublox_ms_load_run(cublox_ms *h, const int *sids, int n_sessions,
                       const float *const *xs, const float *const *ys,
                       const float *const *zs, const int *n, const float *T,
                       const float *base, const cublox_ms_load_cfg *cfg,
                       long long *submitted, double *max_lag_ms);

#ifdef __cplusplus
}
#endif

#endif /* CUBLOX_MS_CAPI_H */
