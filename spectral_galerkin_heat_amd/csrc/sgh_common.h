// Shared definitions for the MI355X (gfx950) kernels of the spectral Galerkin
// ETD1 + Picard step: error reporting across the C ABI, dtype dispatch, launch
// checks and wavefront/workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/sgh_mi355x.h"

namespace sgh {

// Per-thread last-error string returned by sgh_last_error().
void set_error(const char* fmt, ...);

// A HIP error as a return code: device memory exhaustion is SGH_ERR_NOMEM (callers can
// back off), anything else SGH_ERR_HIP.
inline int fail_hip(hipError_t e, const char* where) {
  set_error("%s: %s", where, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? SGH_ERR_NOMEM : SGH_ERR_HIP;
}

#define SGH_CHECK_LAUNCH(where)                                        \
  do {                                                                 \
    hipError_t _e = hipGetLastError();                                 \
    if (_e != hipSuccess) return ::sgh::fail_hip(_e, where);           \
  } while (0)

#define SGH_REQUIRE(cond, code, ...)                                   \
  do {                                                                 \
    if (!(cond)) {                                                     \
      ::sgh::set_error(__VA_ARGS__);                                   \
      return code;                                                     \
    }                                                                  \
  } while (0)

constexpr int kWave = 64;

// Sum over a 64-lane wavefront (all lanes receive the total).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

// Block-wide sum of a double; blockDim.x must be a multiple of 64 and <= 1024.
// Result valid in thread 0. `sh` must hold blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / kWave;
    for (int i = 0; i < nw; ++i) t += sh[i];
  }
  __syncthreads();
  return t;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  v = wave_max(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = -1.0 / 0.0;
  if (threadIdx.x == 0) {
    const int nw = blockDim.x / kWave;
    for (int i = 0; i < nw; ++i) t = sh[i] > t ? sh[i] : t;
  }
  __syncthreads();
  return t;
}

inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

}  // namespace sgh

// dtype dispatch: F is double for SGH_F64 and float for SGH_F32.
#define SGH_DISPATCH(dtype, ...)                                       \
  do {                                                                 \
    if ((dtype) == SGH_F64) {                                          \
      using F = double;                                                \
      __VA_ARGS__;                                                     \
    } else if ((dtype) == SGH_F32) {                                   \
      using F = float;                                                 \
      __VA_ARGS__;                                                     \
    } else {                                                           \
      ::sgh::set_error("unsupported dtype code %d", (int)(dtype));     \
      return SGH_ERR_ARG;                                              \
    }                                                                  \
  } while (0)
