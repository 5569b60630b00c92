// Native replay of a driver's scan loop (include/cublox/seq_replay.h): per
// step the blocking update and the recenter through the drop-in C ABI,
// timed on the host's steady clock.
#include "cublox/seq_replay.h"

#include <chrono>

#include "cublox/capi.h"

extern "C" int cublox_seq_replay(const cublox_seq_step *steps, int count, int on_host,
                                 double *step_ms, int *moved) {
  using clock = std::chrono::steady_clock;
  if (count < 0 || (count > 0 && !steps)) return CUBLOX_ERR_INVALID;
  for (int i = 0; i < count; ++i) {
    const cublox_seq_step &s = steps[i];
    const clock::time_point a = clock::now();
    int st = on_host ? cublox_og_update_sensor(s.grid, s.xs, s.ys, s.zs, s.n, s.T)
                     : cublox_og_update_sensor_device(s.grid, s.xs, s.ys, s.zs, s.n, s.T);
    if (st != 0) return st;
    int mv = 0;
    st = cublox_og_recenter(s.grid, s.base, &mv);
    if (st != 0) return st;
    const clock::time_point b = clock::now();
    if (step_ms) step_ms[i] = std::chrono::duration<double, std::milli>(b - a).count();
    if (moved) moved[i] = mv;
  }
  return 0;
}
