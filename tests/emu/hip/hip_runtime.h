// Host emulation of the HIP runtime and device built-ins, for a CPU build of
// jxr_amd/csrc (test infrastructure only; the shipped library is built by
// hipcc for gfx950).
//
// Execution model: a launch runs its workgroups one after another, each with
// a single thread (threadIdx.x == 0, blockDim.x == 1). Every stage of the
// reconstruction kernel is a loop `for (t = threadIdx.x; t < n; t +=
// blockDim.x)` followed by a barrier, so one thread runs every item of a
// stage before the next stage starts: the order the barriers impose on the
// device. Dynamic LDS is a heap buffer of the launch's exact size, so
// AddressSanitizer sees any access outside it.
#pragma once
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
#define JXR_GLOBAL

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_index {
  unsigned x, y, z;
};
extern emu_index threadIdx, blockIdx, blockDim, gridDim;
extern void *emu_dynamic_lds;
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu_dynamic_lds;

inline void __syncthreads() {}
struct int4 {
  int x, y, z, w;
};
struct uint2 {
  unsigned x, y;
};
struct int2 {
  int x, y;
};
struct uint4 {
  unsigned x, y, z, w;
};
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return uint4{x, y, z, w}; }
inline unsigned atomicCAS(unsigned *p, unsigned cmp, unsigned val) {
  unsigned old = *p;
  if (old == cmp) *p = val;
  return old;
}
inline unsigned atomicOr(unsigned *p, unsigned v) {
  unsigned old = *p;
  *p = old | v;
  return old;
}
inline unsigned atomicAdd(unsigned *p, unsigned v) {
  unsigned old = *p;
  *p = old + v;
  return old;
}
// one thread runs a workgroup: it is its own first lane
inline unsigned __builtin_amdgcn_readfirstlane(unsigned v) { return v; }
// wave issue priority: one thread per workgroup here, nothing to order
inline void __builtin_amdgcn_s_setprio(short) {}
inline float __uint_as_float(unsigned u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
inline unsigned __float_as_uint(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  return u;
}
inline float __fmul_rn(float a, float b) { return a * b; }
inline int __mul24(int a, int b) { return (int)((unsigned)(((a << 8) >> 8) * ((b << 8) >> 8))); }
inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((unsigned long long)a * b) >> 32); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

// ---- runtime API ------------------------------------------------------------
typedef int hipError_t;
#define hipSuccess 0
#define hipErrorNotReady 600
#define hipErrorOutOfMemory 2
#define hipErrorInvalidValue 1
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
struct hipDeviceProp_t {
  char gcnArchName[256];
  int multiProcessorCount;
  size_t sharedMemPerBlock;
  size_t sharedMemPerBlockOptin;
  int clockRate;  // kHz
};

hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *prop, int device);
hipError_t hipSetDevice(int);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipMalloc(void **p, size_t bytes);
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t hipMemsetAsync(void *dst, int value, size_t bytes, hipStream_t s);
hipError_t hipMemset2DAsync(void *dst, size_t pitch, int value, size_t width, size_t height, hipStream_t s);
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute attr, int value);
// a device symbol's address: the host array itself in the CPU build
inline hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
  *p = const_cast<void *>(symbol);
  return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *blocks, const void *f, int threads, size_t lds);

// Graph capture: while a capture is open, launches (on any stream, as the
// streams joined to the capture by events would be) are recorded and run
// when the instantiated graph is launched.
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph *hipGraphExec_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *graph);
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t graph, void *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t exec);
hipError_t hipGraphDestroy(hipGraph_t graph);

void emu_begin_launch(size_t lds_bytes);
void emu_end_launch();
bool emu_capture(const std::function<void()> &launch);  // true when recorded

template <class... A, class... B>
inline void emu_launch(void (*kernel)(A...), dim3 grid, dim3 block, size_t lds, hipStream_t, B... args) {
  (void)block;
  auto run = [=]() {
    emu_begin_launch(lds);
    gridDim.x = grid.x;
    for (unsigned b = 0; b < grid.x; b++) {
      blockIdx.x = b;
      threadIdx.x = 0;
      blockDim.x = 1;
      kernel(args...);
    }
    emu_end_launch();
  };
  if (!emu_capture(run)) run();
}
#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
  emu_launch(kernel, grid, block, lds, stream, __VA_ARGS__)
