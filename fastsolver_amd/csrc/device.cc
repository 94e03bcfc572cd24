// Device plumbing exported with the C-ABI (memory, copies, streams, events),
// so that hosts without their own HIP runtime binding can drive the solver.
// Not part of the reference's interface.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fastsolver_amd.h"

extern "C" {

int hodlr_device_count(int32_t* n) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return HODLR_ERR_HIP;
  if (n) *n = c;
  return HODLR_OK;
}

int hodlr_device_name(char* buf, int32_t len) {
  hipDeviceProp_t p;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return HODLR_ERR_HIP;
  int i = 0;
  for (; i + 1 < len && p.gcnArchName[i]; ++i) buf[i] = p.gcnArchName[i];
  if (len > 0) buf[i] = 0;
  return HODLR_OK;
}

int hodlr_device_malloc(void** ptr, int64_t bytes) {
  if (!ptr || bytes < 0) return HODLR_ERR_ARG;
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;
  return hipMalloc(ptr, size_t(bytes)) == hipSuccess ? HODLR_OK : HODLR_ERR_NOMEM;
}

int hodlr_device_free(void* ptr) { return hipFree(ptr) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP; }

// kind: 0 host->device, 1 device->host, 2 device->device.  Synchronous with
// respect to all streams of the device (plans run on non-blocking streams).
int hodlr_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind) {
  if (bytes <= 0) return HODLR_OK;
  if (hipDeviceSynchronize() != hipSuccess) return HODLR_ERR_HIP;
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  return hipMemcpy(dst, src, size_t(bytes), k) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP;
}

int hodlr_stream_create(void** s) {
  hipStream_t st;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return HODLR_ERR_HIP;
  *s = st;
  return HODLR_OK;
}

int hodlr_stream_destroy(void* s) { return hipStreamDestroy((hipStream_t)s) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP; }

int hodlr_stream_synchronize(void* s) {
  return hipStreamSynchronize((hipStream_t)s) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP;
}

int hodlr_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? HODLR_OK : HODLR_ERR_HIP; }

int hodlr_event_create(void** e) {
  hipEvent_t ev;
  if (hipEventCreate(&ev) != hipSuccess) return HODLR_ERR_HIP;
  *e = ev;
  return HODLR_OK;
}

int hodlr_event_destroy(void* e) { return hipEventDestroy((hipEvent_t)e) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP; }

int hodlr_event_record(void* e, void* s) {
  return hipEventRecord((hipEvent_t)e, (hipStream_t)s) == hipSuccess ? HODLR_OK : HODLR_ERR_HIP;
}

int hodlr_event_elapsed_ms(void* a, void* b, double* ms) {
  if (hipEventSynchronize((hipEvent_t)b) != hipSuccess) return HODLR_ERR_HIP;
  float f = 0;
  if (hipEventElapsedTime(&f, (hipEvent_t)a, (hipEvent_t)b) != hipSuccess) return HODLR_ERR_HIP;
  *ms = f;
  return HODLR_OK;
}

}  // extern "C"
