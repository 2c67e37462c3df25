// Host model of the subset of the HIP runtime that mmrecode_amd/csrc/recon_host.cpp calls.
//
// Test infrastructure only: tests/served_replay/build.py compiles the library's real host code
// (recon_host.cpp) with g++ against this header, so that the submission, ticket and serving logic
// runs unchanged on the CPU while "device" memory, streams, events and kernel launches are
// modelled by hip_model.cpp and the kernels by kernel_model.cpp (address models that check every
// buffer a kernel would touch against the live allocations). Nothing in the product includes it.
#pragma once

#include <cstddef>
#include <cstdint>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __align__(n) __attribute__((aligned(n)))

typedef int hipError_t;
enum : int {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorMemoryAllocation = 2,
    hipErrorInvalidResourceHandle = 400,
    hipErrorNotReady = 600,
    hipErrorIllegalState = 401,
};

typedef struct ModelStream* hipStream_t;
typedef struct ModelEvent* hipEvent_t;
typedef struct ModelGraph* hipGraph_t;
typedef struct ModelGraph* hipGraphExec_t;
typedef struct ModelGraphNode* hipGraphNode_t;
typedef struct ModelMemPool* hipMemPool_t;
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold = 4 };

enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeRelaxed = 2 };

constexpr unsigned hipStreamDefault = 0, hipStreamNonBlocking = 1;
constexpr unsigned hipEventDefault = 0, hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;

struct hipDeviceProp_t {
    char name[256];
    int multiProcessorCount;
    size_t totalGlobalMem;
};

#define HIP_SYMBOL(x) (&(x))

hipError_t hipGetDeviceCount(int* count);
hipError_t hipSetDevice(int device);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* prop, int device);
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();

hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t s);
hipError_t hipFreeAsync(void* p, hipStream_t s);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t hipMemsetAsync(void* dst, int value, size_t bytes, hipStream_t s);
hipError_t hipGetSymbolAddress(void** out, const void* symbol);
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* pool, int device);
hipError_t hipMemPoolSetAttribute(hipMemPool_t pool, hipMemPoolAttr attr, void* value);

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipDeviceSynchronize();

hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t g, hipGraphNode_t* err_node, char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t exec);
hipError_t hipGraphDestroy(hipGraph_t g);
