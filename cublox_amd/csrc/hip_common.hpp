// Shared HIP-side helpers: status codes, the per-thread error message and
// the grid index math used by every kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "cublox/capi.h"

namespace cublox_hip {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);

inline int check(hipError_t e, const char *what) {
  if (e == hipSuccess) return CUBLOX_OK;
  return set_error(CUBLOX_ERR_RUNTIME,
                   std::string("cublox HIP error in `") + what +
                       "`: " + hipGetErrorString(e));
}

// Status of the most recent launch on this thread.
inline int check_launch(const char *what) {
  return check(hipGetLastError(), what);
}

#define CUBLOX_TRY(call)                                                      \
  do {                                                                        \
    const int _st = ::cublox_hip::check((call), #call);                       \
    if (_st != CUBLOX_OK) return _st;                                         \
  } while (0)

inline hipStream_t as_stream(cublox_stream s) {
  return reinterpret_cast<hipStream_t>(s);
}

// ---------------------------------------------------------------------
// Index math of the toroidal window (reference utils.cuh:24-176). Every FP32
// expression keeps the reference's operand order; the library is built with
// -ffp-contract=off, so each step is one IEEE-rounded operation.
// ---------------------------------------------------------------------

// utils.cuh:24-32
__host__ __device__ __forceinline__ int pos_to_global(float p, float inv_res,
                                                      bool center) {
  if (center) {
    const float s = p < 0.0f ? -1.0f : 1.0f;
    return static_cast<int>(p * inv_res + s * 0.5f);
  }
  return static_cast<int>(floorf(p * inv_res));
}

// utils.cuh:44-48
__host__ __device__ __forceinline__ float global_to_pos(int g, float res,
                                                        bool center) {
  return center ? static_cast<float>(g) * res
                : (static_cast<float>(g) + 0.5f) * res;
}

// utils.cuh:60-70: C remainder folded into [-half, half].
__host__ __device__ __forceinline__ int global_to_local(int g, int map,
                                                        int half) {
  int l = g % map;
  if (l > half) {
    l -= map;
  } else if (l < -half) {
    l += map;
  }
  return l;
}

// utils.cuh:131-136 with utils.cuh:160-166: x-major, z fastest.
__host__ __device__ __forceinline__ int global_to_hash(int gx, int gy, int gz,
                                                       int mx, int my, int mz,
                                                       int hx, int hy,
                                                       int hz) {
  const int lx = global_to_local(gx, mx, hx) + hx;
  const int ly = global_to_local(gy, my, hy) + hy;
  const int lz = global_to_local(gz, mz, hz) + hz;
  return lx * my * mz + ly * mz + lz;
}

// kernels.cu:23-28
__host__ __device__ __forceinline__ bool inside_window(int gx, int gy, int gz,
                                                       int ox, int oy, int oz,
                                                       int hx, int hy,
                                                       int hz) {
  return abs(gx - ox) <= hx && abs(gy - oy) <= hy && abs(gz - oz) <= hz;
}

} // namespace cublox_hip
