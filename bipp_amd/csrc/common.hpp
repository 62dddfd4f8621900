// Shared device/host helpers for the gfx950 Bluebild kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "bipp/enums.h"
#include "bipp/errors.h"

namespace bippamd {

// ---------------------------------------------------------------------------------------
// Errors: every failure inside the library is a BippException carrying a BippError code;
// the C-ABI layer converts it to a return code plus a thread-local message.
// ---------------------------------------------------------------------------------------
struct BippException : std::runtime_error {
  BippError code;
  BippException(BippError c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

[[noreturn]] inline void fail(BippError c, const std::string& msg) { throw BippException(c, msg); }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    fail(BIPP_GPU_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}
#define BIPP_HIP(call) ::bippamd::hip_check((call), #call)

inline void check_launch(const char* what) { hip_check(hipGetLastError(), what); }

}  // namespace bippamd

#include "cplx.hpp"

namespace bippamd {

// ---------------------------------------------------------------------------------------
// Wave64 / workgroup reductions.
// ---------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a workgroup of NT threads; `scratch` holds NT/64 values.  All threads get
// the result.  Contains two barriers.
template <int NT, typename R>
__device__ __forceinline__ R block_sum(R v, R* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  R s = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += scratch[w];
  __syncthreads();
  return s;
}

// Lane moves within rows of 16 lanes (DPP, no LDS traffic).  CTRL: 0xB1 quad_perm [1,0,3,2],
// 0x4E quad_perm [2,3,0,1], 0x141 row_half_mirror, 0x140 row_mirror.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// Sum over each aligned group of 16 lanes, returned in every lane of the group; every lane
// adds the same operands (in commuted order), so all lanes hold bit-identical sums.
template <typename R>
__device__ __forceinline__ R sum16(R v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return v;
}

// Orders the LDS accesses of the lanes of one wavefront around a hand-off through LDS
// (lanes write, other lanes of the same wavefront read afterwards, no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace bippamd
