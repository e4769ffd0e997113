// Host-only stand-in for the few HIP runtime calls sse_wilson_dslash_amd/csrc/stager.h
// makes, so its worker pool and chunking can be checked on a CPU (and under
// ThreadSanitizer / AddressSanitizer): pinned memory is malloc'd, copies are
// synchronous memcpy, events are always complete.  Test code only.
#pragma once
#include <stdlib.h>
#include <string.h>

typedef int hipError_t;
typedef void* hipEvent_t;
typedef void* hipStream_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2 };
enum { hipEventDisableTiming = 2 };

inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return (*p = malloc(n)) ? hipSuccess : 2; }
inline hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (void*)e; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
  memcpy(d, s, n);
  return hipSuccess;
}
