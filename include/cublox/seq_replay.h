/* Native replay of a driver's scan loop over the drop-in C ABI.
 *
 * The reference driver calls, for every scan, the blocking update with the
 * scan's pose and then recenter at the robot's base (reference
 * main.cpp:437-456 and OccupancyGrid.cpp:181-238, 141-179), from C++. This
 * library runs such a sequence of steps natively, timing each step on the
 * host clock, so that a caller written in another language (the Python
 * bench, the tests) measures the calls themselves and not its own
 * interpreter between them. Each step is exactly cublox_og_update_sensor
 * (host columns) or cublox_og_update_sensor_device (device columns)
 * followed by cublox_og_recenter; nothing else runs between steps. Built as
 * libcublox_seq.so against libcublox.so; it has no counterpart in the
 * reference (whose perf test is such a loop, test/test_occupancy_grid_perf.cpp). */
#ifndef CUBLOX_SEQ_REPLAY_H_
#define CUBLOX_SEQ_REPLAY_H_

#include "cublox/og_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One step: update `grid` with the n points of columns xs, ys, zs (sensor
 * frame; T = sensor->world, row-major 3x4), then recenter it at base. */
typedef struct cublox_seq_step {
  cublox_og *grid;
  const float *xs;
  const float *ys;
  const float *zs;
  int n;
  float T[12];
  float base[3];
} cublox_seq_step;

/* Runs steps[0 .. count) in order. on_host = 0: the columns are device
 * memory; 1: host memory (copied inside each update, as the reference's
 * update copies its cloud). step_ms[i] (may be NULL): the host wall time of
 * step i in milliseconds; moved[i] (may be NULL): 1 when step i's recenter
 * moved the window. Returns 0; CUBLOX_ERR_INVALID (1) for a negative
 * count or NULL steps; else the first failing call's status, with the
 * steps before it done (cublox_og_last_error() has the message). */
int cublox_seq_replay(const cublox_seq_step *steps, int count, int on_host,
                      double *step_ms, int *moved);

#ifdef __cplusplus
}
#endif

#endif /* CUBLOX_SEQ_REPLAY_H_ */
