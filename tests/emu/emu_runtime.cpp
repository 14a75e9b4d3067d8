// Host implementations of the emulated HIP runtime (test infrastructure).
#include <stdio.h>

#include <chrono>
#include <vector>

#include "hip/hip_runtime.h"

emu_index threadIdx, blockIdx, blockDim, gridDim;
void *emu_dynamic_lds = nullptr;

struct emu_stream {
  int unused;
};
struct emu_event {
  double t;
};

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void emu_begin_launch(size_t lds_bytes) {
  // poison the buffer so a read before any write differs between runs
  emu_dynamic_lds = malloc(lds_bytes ? lds_bytes : 1);
  memset(emu_dynamic_lds, 0xCD, lds_bytes ? lds_bytes : 1);
}
void emu_end_launch() {
  free(emu_dynamic_lds);
  emu_dynamic_lds = nullptr;
}

struct emu_graph {
  std::vector<std::function<void()>> launches;
};
static emu_graph *g_capture = nullptr;

bool emu_capture(const std::function<void()> &launch) {
  if (!g_capture) return false;
  g_capture->launches.push_back(launch);
  return true;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  if (g_capture) return hipErrorInvalidValue;
  g_capture = new emu_graph();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *graph) {
  if (!g_capture) return hipErrorInvalidValue;
  *graph = g_capture;
  g_capture = nullptr;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t graph, void *, char *, size_t) {
  *exec = new emu_graph(*graph);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
  for (auto &f : exec->launches) f();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
  delete exec;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t graph) {
  delete graph;
  return hipSuccess;
}

hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *prop, int) {
  memset(prop, 0, sizeof *prop);
  strcpy(prop->gcnArchName, "gfx950:sramecc+:xnack-");
  prop->multiProcessorCount = 4;
  prop->sharedMemPerBlock = 65536;
  prop->sharedMemPerBlockOptin = 65536;
  prop->clockRate = 2400000;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
  *s = new emu_stream{0};
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new emu_event{0};
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->t = now_ms();
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = (float)(b->t - a->t);
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes) {
  *free_bytes = size_t(8) << 30;
  *total_bytes = size_t(8) << 30;
  return hipSuccess;
}
hipError_t hipMalloc(void **p, size_t bytes) {
  *p = malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipFree(void *p) {
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
  memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t) {
  memset(dst, value, bytes);
  return hipSuccess;
}
hipError_t hipMemset2DAsync(void *dst, size_t pitch, int value, size_t width, size_t height, hipStream_t) {
  for (size_t r = 0; r < height; r++) memset((char *)dst + r * pitch, value, width);
  return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *blocks, const void *, int, size_t lds) {
  *blocks = (int)std::max<size_t>(1, std::min<size_t>(4, 65536 / (lds + 2048)));
  return hipSuccess;
}
