// Host runtime of libbipp_amd: error conversion for the C-ABI, growable device and pinned
// buffers, and the context (one HIP stream on the current device).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "bipp_hip.h"
#include "common.hpp"

namespace bippamd {

// message of the last failing C-ABI call on this thread (one instance across the library)
inline thread_local std::string g_last_error;

// Runs f and converts exceptions into BippError codes and the thread-local message.
template <typename F>
BippError guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return BIPP_SUCCESS;
  } catch (const BippException& e) {
    // a failed runtime call leaves its code as HIP's last error: clear it, so that the next
    // launch check does not report it again
    (void)hipGetLastError();
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    (void)hipGetLastError();
    g_last_error = "host allocation failed";
    return BIPP_GENERIC_ERROR;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BIPP_GENERIC_ERROR;
  } catch (...) {
    g_last_error = "unknown error";
    return BIPP_GENERIC_ERROR;
  }
}

inline void require(bool cond, BippError code, const char* msg) {
  if (!cond) fail(code, msg);
}
inline void require(bool cond, BippError code, const std::string& msg) {
  if (!cond) fail(code, msg);
}

// Growable device allocation.  Reallocation frees the old block; hipFree synchronises the
// device, so work still in flight on it has finished first.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void* get(size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > cap) {
      if (p) BIPP_HIP(hipFree(p));
      p = nullptr;
      cap = 0;
      BIPP_HIP(hipMalloc(&p, bytes));
      cap = bytes;
    }
    return p;
  }
  template <typename T>
  T* as(size_t n) {
    return static_cast<T*>(get(n * sizeof(T)));
  }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void* get(size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > cap) {
      if (p) BIPP_HIP(hipHostFree(p));
      p = nullptr;
      cap = 0;
      BIPP_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
      cap = bytes;
    }
    return p;
  }
};

// ---------------------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------------------
struct ContextImpl {
  BippProcessingUnit pu = BIPP_PU_GPU;
  int device = 0;
  int n_cu = 1;
  hipStream_t stream = nullptr;

  explicit ContextImpl(BippProcessingUnit p) {
    require(p != BIPP_PU_CPU, BIPP_INVALID_PARAMETER_ERROR,
            "this library computes on the GPU only (processing unit CPU requested)");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      fail(BIPP_GPU_SUPPORT_ERROR, "no GPU device available");
    BIPP_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    BIPP_HIP(hipGetDeviceProperties(&prop, device));
    n_cu = prop.multiProcessorCount;
    BIPP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    BIPP_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    pu = BIPP_PU_GPU;
  }
  ~ContextImpl() {
    if (ready) (void)hipEventDestroy(ready);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // Waits for the stream by polling an event: a blocking wait that spans a long kernel wakes
  // up about a millisecond late, which leaves the device idle between batches.
  void sync() {
    BIPP_HIP(hipEventRecord(ready, stream));
    for (;;) {
      const hipError_t q = hipEventQuery(ready);
      if (q == hipSuccess) break;
      if (q != hipErrorNotReady) BIPP_HIP(q);
    }
  }

 private:
  hipEvent_t ready = nullptr;
};

}  // namespace bippamd
