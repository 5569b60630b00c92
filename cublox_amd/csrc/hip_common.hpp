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

// Global-memory view of a device pointer that a kernel reads from memory (a
// descriptor field) rather than receives as an argument. Without it the
// compiler cannot prove the address space and emits flat_* accesses, which
// also count against the LDS wait counter and so serialise with LDS work.
#define CUBLOX_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ CUBLOX_GLOBAL T *as_global(T *p) {
  return (CUBLOX_GLOBAL T *)p;
}

// Memory that no thread writes while a kernel runs (launch descriptors) is
// read through the constant address space: uniform addresses then compile to
// scalar loads.
#define CUBLOX_CONST __attribute__((address_space(4)))

// Copy of *p read word by word through the constant address space; pointer
// fields of the copy are then known to address global memory.
template <typename T>
__device__ __forceinline__ T load_const(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "load_const: whole 32-bit words");
  T v;
  int *o = reinterpret_cast<int *>(&v);
#pragma unroll
  for (int i = 0; i < static_cast<int>(sizeof(T) / 4); ++i) {
    o[i] = ((const CUBLOX_CONST int *)p)[i];
  }
  return v;
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

// utils.cuh:82-92: the global coordinate of local index l in the window
// centred on o, i.e. the one g in [o - h, o + h] with g = l (mod map).
__host__ __device__ __forceinline__ int local_to_global(int l, int map, int half,
                                                        int o) {
  const int lo = o - half;
  int d = (l - lo) % map;
  if (d < 0) d += map;
  return lo + d;
}

// utils.cuh:120-129: hash id -> local index (each axis in [-half, half]).
__host__ __device__ __forceinline__ void hash_to_local(int hid, int my, int mz,
                                                       int hx, int hy, int hz,
                                                       int &lx, int &ly,
                                                       int &lz) {
  const int yz = my * mz;
  const int x = hid / yz;
  const int r = hid - x * yz;
  const int y = r / mz;
  lx = x - hx;
  ly = y - hy;
  lz = (r - y * mz) - hz;
}

// kernels.cu:23-28
__host__ __device__ __forceinline__ bool inside_window(int gx, int gy, int gz,
                                                       int ox, int oy, int oz,
                                                       int hx, int hy,
                                                       int hz) {
  return abs(gx - ox) <= hx && abs(gy - oy) <= hy && abs(gz - oz) <= hz;
}

// ---------------------------------------------------------------------
// Division by a divisor fixed for a launch (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994,
// Fig. 4.1): one high multiply, exact for every 32-bit dividend.
// ---------------------------------------------------------------------
struct FastDiv {
  unsigned m, s1, s2;
};

__host__ __device__ inline FastDiv make_fastdiv(unsigned d) {
  unsigned l = 0;  // ceil(log2(d))
  while (l < 32 && (1ull << l) < d) ++l;
  const unsigned long long m =
      ((1ull << 32) * ((1ull << l) - d)) / d + 1ull;
  return FastDiv{static_cast<unsigned>(m), l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}

__device__ __forceinline__ unsigned fast_div(unsigned n, const FastDiv &f) {
  const unsigned t = __umulhi(n, f.m);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// ---------------------------------------------------------------------
// Wave and block scans (wave64).
// ---------------------------------------------------------------------
__device__ __forceinline__ int wave_scan_incl(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// Inclusive block scan of one int per thread: within waves by shuffles,
// then across the wave totals (s_wave: one int per wave).
__device__ __forceinline__ int block_scan_incl(int *s_wave, int v) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  v = wave_scan_incl(v);
  if (lane == 63) s_wave[wv] = v;
  __syncthreads();
  if (wv == 0) {
    const int x = wave_scan_incl(lane < nw ? s_wave[lane] : 0);
    if (lane < nw) s_wave[lane] = x;
  }
  __syncthreads();
  const int r = v + (wv > 0 ? s_wave[wv - 1] : 0);
  __syncthreads();
  return r;
}

} // namespace cublox_hip
