// Runtime primitives of the C ABI: errors, device memory, streams, events.
// Host C++ reaches the HIP runtime only through these (no HIP headers in
// the drop-in OccupancyGrid build).
#include "hip_common.hpp"

#include <cstring>

namespace cublox_hip {

namespace {
thread_local std::string g_last_error;
} // namespace

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

// Zeroes n 32-bit words at p (4 B aligned): scalar head up to the first 16 B
// boundary, 16 B stores for the body, scalar tail.
__global__ void __launch_bounds__(256)
    fill_zero_kernel(uint32_t *__restrict__ p, size_t n) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  size_t head = ((16u - (a & 15u)) & 15u) / 4u;
  if (head > n) head = n;
  const size_t n4 = (n - head) / 4;
  const size_t tail0 = head + n4 * 4;
  uint4 *__restrict__ p4 = reinterpret_cast<uint4 *>(p + head);
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (size_t i = t; i < n4; i += stride) p4[i] = z;
  if (t < head) p[t] = 0u;
  if (t < n - tail0) p[tail0 + t] = 0u;
}

} // namespace cublox_hip

using namespace cublox_hip;

extern "C" {

const char *cublox_last_error(void) { return g_last_error.c_str(); }

int cublox_device_count(int *count) {
  if (!count) return set_error(CUBLOX_ERR_INVALID, "null count");
  CUBLOX_TRY(hipGetDeviceCount(count));
  return CUBLOX_OK;
}

int cublox_set_device(int device) {
  CUBLOX_TRY(hipSetDevice(device));
  return CUBLOX_OK;
}

int cublox_get_device(int *device) {
  CUBLOX_TRY(hipGetDevice(device));
  return CUBLOX_OK;
}

int cublox_device_name(char *buf, int cap) {
  int dev = 0;
  CUBLOX_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  CUBLOX_TRY(hipGetDeviceProperties(&prop, dev));
  if (buf && cap > 0) {
    std::snprintf(buf, static_cast<size_t>(cap), "%s (%s, %d CUs)", prop.name,
                  prop.gcnArchName, prop.multiProcessorCount);
  }
  return CUBLOX_OK;
}

int cublox_mem_info(size_t *free_bytes, size_t *total_bytes) {
  CUBLOX_TRY(hipMemGetInfo(free_bytes, total_bytes));
  return CUBLOX_OK;
}

int cublox_malloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(CUBLOX_ERR_INVALID, "null ptr");
  *ptr = nullptr;
  if (bytes == 0) return CUBLOX_OK;
  CUBLOX_TRY(hipMalloc(ptr, bytes));
  return CUBLOX_OK;
}

int cublox_free(void *ptr) {
  if (!ptr) return CUBLOX_OK;
  CUBLOX_TRY(hipFree(ptr));
  return CUBLOX_OK;
}

int cublox_host_alloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(CUBLOX_ERR_INVALID, "null ptr");
  *ptr = nullptr;
  if (bytes == 0) return CUBLOX_OK;
  CUBLOX_TRY(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
  return CUBLOX_OK;
}

int cublox_host_device_pointer(void **dptr, void *hptr) {
  if (!dptr || !hptr) return set_error(CUBLOX_ERR_INVALID, "null ptr");
  *dptr = nullptr;
  CUBLOX_TRY(hipHostGetDevicePointer(dptr, hptr, 0));
  return CUBLOX_OK;
}

int cublox_host_free(void *ptr) {
  if (!ptr) return CUBLOX_OK;
  CUBLOX_TRY(hipHostFree(ptr));
  return CUBLOX_OK;
}

int cublox_memset_async(void *ptr, int value, size_t bytes, cublox_stream s) {
  if (!ptr || bytes == 0) return CUBLOX_OK;
  CUBLOX_TRY(hipMemsetAsync(ptr, value, bytes, as_stream(s)));
  return CUBLOX_OK;
}

int cublox_memcpy_h2d_async(void *dst, const void *src, size_t bytes,
                            cublox_stream s) {
  if (bytes == 0) return CUBLOX_OK;
  CUBLOX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                            as_stream(s)));
  return CUBLOX_OK;
}

int cublox_memcpy_d2h_async(void *dst, const void *src, size_t bytes,
                            cublox_stream s) {
  if (bytes == 0) return CUBLOX_OK;
  CUBLOX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost,
                            as_stream(s)));
  return CUBLOX_OK;
}

int cublox_memcpy_d2d_async(void *dst, const void *src, size_t bytes,
                            cublox_stream s) {
  if (bytes == 0) return CUBLOX_OK;
  CUBLOX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                            as_stream(s)));
  return CUBLOX_OK;
}

int cublox_stream_create(cublox_stream *s) {
  if (!s) return set_error(CUBLOX_ERR_INVALID, "null stream out");
  hipStream_t st = nullptr;
  CUBLOX_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = reinterpret_cast<cublox_stream>(st);
  return CUBLOX_OK;
}

int cublox_stream_create_dedicated(cublox_stream *s) {
  if (!s) return set_error(CUBLOX_ERR_INVALID, "null stream out");
  int dev = 0, cus = 0;
  CUBLOX_TRY(hipGetDevice(&dev));
  CUBLOX_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (cus <= 0) return set_error(CUBLOX_ERR_RUNTIME, "stream_create_dedicated: no CUs");
  // Every CU enabled: the mask only gives the stream a queue of its own.
  std::vector<uint32_t> mask(static_cast<size_t>((cus + 31) / 32), 0xFFFFFFFFu);
  if (cus % 32) mask.back() = (1u << (cus % 32)) - 1u;
  hipStream_t st = nullptr;
  CUBLOX_TRY(hipExtStreamCreateWithCUMask(&st, static_cast<uint32_t>(mask.size()), mask.data()));
  *s = reinterpret_cast<cublox_stream>(st);
  return CUBLOX_OK;
}

int cublox_stream_destroy(cublox_stream s) {
  if (!s) return CUBLOX_OK;
  CUBLOX_TRY(hipStreamDestroy(as_stream(s)));
  return CUBLOX_OK;
}

int cublox_stream_synchronize(cublox_stream s) {
  CUBLOX_TRY(hipStreamSynchronize(as_stream(s)));
  return CUBLOX_OK;
}

int cublox_device_synchronize(void) {
  CUBLOX_TRY(hipDeviceSynchronize());
  return CUBLOX_OK;
}

int cublox_event_create(cublox_event *e) {
  if (!e) return set_error(CUBLOX_ERR_INVALID, "null event out");
  hipEvent_t ev = nullptr;
  CUBLOX_TRY(hipEventCreate(&ev));
  *e = reinterpret_cast<cublox_event>(ev);
  return CUBLOX_OK;
}

int cublox_event_destroy(cublox_event e) {
  if (!e) return CUBLOX_OK;
  CUBLOX_TRY(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)));
  return CUBLOX_OK;
}

int cublox_event_record(cublox_event e, cublox_stream s) {
  CUBLOX_TRY(hipEventRecord(reinterpret_cast<hipEvent_t>(e), as_stream(s)));
  return CUBLOX_OK;
}

int cublox_event_synchronize(cublox_event e) {
  CUBLOX_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)));
  return CUBLOX_OK;
}

int cublox_event_elapsed_ms(cublox_event a, cublox_event b, float *ms) {
  CUBLOX_TRY(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(a),
                                 reinterpret_cast<hipEvent_t>(b)));
  return CUBLOX_OK;
}

int cublox_stream_wait_event(cublox_stream s, cublox_event e) {
  if (!e) return CUBLOX_OK;
  CUBLOX_TRY(hipStreamWaitEvent(as_stream(s), reinterpret_cast<hipEvent_t>(e), 0));
  return CUBLOX_OK;
}

int cublox_graph_begin_capture(cublox_stream s) {
  CUBLOX_TRY(hipStreamBeginCapture(as_stream(s), hipStreamCaptureModeThreadLocal));
  return CUBLOX_OK;
}

int cublox_graph_end_capture(cublox_stream s, cublox_graph *exec) {
  if (!exec) return set_error(CUBLOX_ERR_INVALID, "graph: null output");
  hipGraph_t g = nullptr;
  CUBLOX_TRY(hipStreamEndCapture(as_stream(s), &g));
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  CUBLOX_TRY(err);
  *exec = e;
  return CUBLOX_OK;
}

int cublox_graph_launch(cublox_graph exec, cublox_stream s) {
  CUBLOX_TRY(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(exec), as_stream(s)));
  return CUBLOX_OK;
}

int cublox_graph_destroy(cublox_graph exec) {
  if (!exec) return CUBLOX_OK;
  CUBLOX_TRY(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(exec)));
  return CUBLOX_OK;
}

int cublox_fill_zero_async(void *ptr, size_t count32, cublox_stream s) {
  if (!ptr || count32 == 0) return CUBLOX_OK;
  if (reinterpret_cast<uintptr_t>(ptr) % 4 != 0) {
    return set_error(CUBLOX_ERR_INVALID, "fill_zero: pointer not 4 B aligned");
  }
  size_t blocks = (count32 / 4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  fill_zero_kernel<<<dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     as_stream(s)>>>(static_cast<uint32_t *>(ptr), count32);
  return check_launch("fill_zero_kernel");
}

} // extern "C"
