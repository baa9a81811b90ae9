// A host-only stand-in for the subset of the HIP runtime API that the
// library's host code uses (engine.hip, batch.hip, api.hip), for the CPU
// harness in tests/fakehip: streams are host threads that run their
// operations in order after a random delay, device memory is host memory
// tracked in a registry, events and captured graphs behave as HIP documents
// them.  Kernels are replaced by fake_kernels.cpp, which simulates the
// wavefront's bookkeeping and checks every device pointer a kernel would use.
//
// Test infrastructure only: compiled with g++ into the harness, never into
// the library.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#define __host__
#define __device__
#define __global__
#define __shared__
#define __forceinline__ inline
#define __align__(n) __attribute__((aligned(n)))

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotReady = 600,
  hipErrorLaunchFailure = 719,
  hipErrorUnknown = 999,
} hipError_t;

typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4,
} hipMemcpyKind;

typedef enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2,
} hipStreamCaptureMode;

typedef enum hipDeviceAttribute_t {
  hipDeviceAttributeMultiprocessorCount = 63,
} hipDeviceAttribute_t;

#define hipStreamNonBlocking 0x01
#define hipEventDisableTiming 0x02
#define hipHostMallocDefault 0x0

struct FakeStream;
struct FakeEvent;
struct FakeGraph;
struct FakeGraphExec;
typedef FakeStream *hipStream_t;
typedef FakeEvent *hipEvent_t;
typedef FakeGraph *hipGraph_t;
typedef FakeGraphExec *hipGraphExec_t;
typedef void *hipGraphNode_t;

struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
  size_t totalGlobalMem;
  int pciDomainID, pciBusID, pciDeviceID;
};

struct float4 {
  float x, y, z, w;
};
struct uint2 {
  unsigned int x, y;
};
struct uint4 {
  unsigned int x, y, z, w;
};
struct uchar4 {
  unsigned char x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline uchar4 make_uchar4(unsigned char x, unsigned char y, unsigned char z, unsigned char w) {
  return uchar4{x, y, z, w};
}

hipError_t hipGetDeviceCount(int *n);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int d);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);

hipError_t hipMalloc(void **p, size_t bytes);
template <class T> hipError_t hipMalloc(T **p, size_t bytes) {
  return hipMalloc(reinterpret_cast<void **>(p), bytes);
}
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
template <class T> hipError_t hipHostMalloc(T **p, size_t bytes, unsigned flags = 0) {
  return hipHostMalloc(reinterpret_cast<void **>(p), bytes, flags);
}
hipError_t hipHostFree(void *p);

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k,
                          hipStream_t s);
hipError_t hipMemset(void *dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s);

hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);

hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);

// A __device__ variable is a host array here; HIP_SYMBOL also passes its
// size so the fake can register it as device memory.
#define HIP_SYMBOL(X) (&(X)), sizeof(X)
hipError_t hipGetSymbolAddress(void **p, const void *symbol, size_t bytes);

hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t g, hipGraphNode_t *err_node,
                               char *log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);
hipError_t hipGraphDestroy(hipGraph_t g);
