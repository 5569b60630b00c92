// Sensor ingest on the device: the driver's input side of the update path
// (reference src/main.cpp). A depth image becomes the sensor-frame cloud
// the driver would publish (depthCallback, main.cpp:558-638); the driver's
// point filter (buildCloudInOdomFrame, main.cpp:640-693: finite, within the
// max raycast range) runs inside the fused update's ray set-up, and the
// census below counts the points it keeps, which the driver needs to know
// whether a cloud updates the grid at all (main.cpp:686-690).
#include <algorithm>

#include "hip_common.hpp"
#include "raycast_common.hpp"

namespace cublox_hip {
namespace {

// main.cpp:600-635, one pixel per thread in row-major order (u fastest),
// the first n pixels: 16-bit millimetres (0 invalid) or 32-bit float metres
// (non-finite or <= 0 invalid). Invalid pixels give a quiet NaN point.
template <bool kU16>
__global__ void __launch_bounds__(256)
    depth_to_points_kernel(const unsigned char *__restrict__ img, int width,
                           int n, size_t step, float fx, float fy, float cx,
                           float cy, float *__restrict__ xs,
                           float *__restrict__ ys, float *__restrict__ zs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = i / width;
  const int u = i - v * width;
  const unsigned char *row = img + static_cast<size_t>(v) * step;
  const float nan = __int_as_float(0x7fc00000);
  float z;
  bool ok;
  if (kU16) {
    const unsigned short raw = reinterpret_cast<const unsigned short *>(row)[u];
    ok = raw != 0;
    z = static_cast<float>(raw) * 0.001f;
  } else {
    z = reinterpret_cast<const float *>(row)[u];
    ok = isfinite(z) && z > 0.0f;
  }
  if (!ok) {
    xs[i] = nan;
    ys[i] = nan;
    zs[i] = nan;
    return;
  }
  xs[i] = ((static_cast<float>(u) - cx) * z) / fx;
  ys[i] = ((static_cast<float>(v) - cy) * z) / fy;
  zs[i] = z;
}

// main.cpp:672-679: a point is kept when finite and, for drop_sq > 0, its
// squared norm ((x*x + y*y) + z*z) is at most drop_sq. One atomic per
// workgroup (a counter every workgroup hits serialises at the L2, so the
// launch is kept to a few workgroups).
constexpr int kCountBlocks = 64;

__global__ void __launch_bounds__(256)
    count_kept_kernel(const float *__restrict__ xs, const float *__restrict__ ys,
                      const float *__restrict__ zs, int stride, int n,
                      float drop_sq, unsigned *__restrict__ count) {
  __shared__ int s_wave[4];
  int kept = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    const size_t at = static_cast<size_t>(i) * static_cast<size_t>(stride);
    const float x = xs[at], y = ys[at], z = zs[at];
    bool keep = isfinite(x) && isfinite(y) && isfinite(z);
    if (keep && drop_sq > 0.0f) keep = !((x * x + y * y) + z * z > drop_sq);
    kept += keep ? 1 : 0;
  }
  const int total = block_scan_incl(s_wave, kept);
  if (threadIdx.x == blockDim.x - 1 && total != 0) {
    atomic_add_u32_ret(count, static_cast<unsigned>(total));
  }
}

}  // namespace
}  // namespace cublox_hip

using namespace cublox_hip;

extern "C" {

int cublox_launch_depth_to_points(const void *d_image, int width, int height,
                                  size_t step_bytes, int encoding,
                                  const float K[4], int max_points,
                                  float *d_xs, float *d_ys, float *d_zs,
                                  cublox_stream s) {
  if (width <= 0 || height <= 0 || max_points <= 0) return CUBLOX_OK;
  if (!d_image || !K || !d_xs || !d_ys || !d_zs) {
    return set_error(CUBLOX_ERR_INVALID, "depth_to_points: null argument");
  }
  if (encoding != 0 && encoding != 1) {
    return set_error(CUBLOX_ERR_INVALID,
                     "depth_to_points: encoding must be 0 (16-bit mm) or 1 (32-bit float m)");
  }
  const size_t elem = encoding == 0 ? 2 : 4;
  if (step_bytes < static_cast<size_t>(width) * elem || step_bytes % elem != 0) {
    return set_error(CUBLOX_ERR_INVALID,
                     "depth_to_points: row step below width or not a multiple of the pixel size");
  }
  const long long pixels = static_cast<long long>(width) * height;
  const int n = static_cast<int>(std::min<long long>(pixels, max_points));
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  const unsigned char *img = static_cast<const unsigned char *>(d_image);
  if (encoding == 0) {
    depth_to_points_kernel<true><<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
        img, width, n, step_bytes, K[0], K[1], K[2], K[3], d_xs, d_ys, d_zs);
  } else {
    depth_to_points_kernel<false><<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
        img, width, n, step_bytes, K[0], K[1], K[2], K[3], d_xs, d_ys, d_zs);
  }
  return check_launch("depth_to_points_kernel");
}

int cublox_launch_count_kept_points(const float *d_xs, const float *d_ys,
                                    const float *d_zs, int point_stride, int n,
                                    float range_drop_sq, unsigned int *d_count,
                                    cublox_stream s) {
  if (!d_count) return set_error(CUBLOX_ERR_INVALID, "count_kept_points: null count");
  int st = cublox_fill_zero_async(d_count, 1, s);
  if (st != CUBLOX_OK || n <= 0) return st;
  if (!d_xs || !d_ys || !d_zs || point_stride < 0) {
    return set_error(CUBLOX_ERR_INVALID, "count_kept_points: bad arguments");
  }
  const int stride = point_stride > 0 ? point_stride : 1;
  const unsigned blocks = static_cast<unsigned>(std::min(kCountBlocks, (n + 255) / 256));
  count_kept_kernel<<<dim3(blocks), dim3(256), 0, as_stream(s)>>>(
      d_xs, d_ys, d_zs, stride, n, range_drop_sq, d_count);
  return check_launch("count_kept_kernel");
}

}  // extern "C"
