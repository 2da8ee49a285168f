// MFMA and LDS-staging building blocks for gfx950 shared by the transform kernels.
//
// Lane layouts (checked on gfx950 with a one-wave probe):
//   v_mfma_f32_16x16x4_f32 / v_mfma_f64_16x16x4_f64: A[i][k] in lane i + 16k,
//     B[k][j] in lane j + 16k (one element per lane, k = 0..3).
//   v_mfma_f32_16x16x32_bf16: A[i][k] in lane i + 16(k/8), element k%8;
//     B[k][j] in lane j + 16(k/8), element k%8 (eight bf16 per lane, k = 0..31).
//   Accumulator D[i][j] (16×16): FP32 results in lane j + 16(i/4), register i%4;
//     FP64 results in lane j + 16(i%4), register i/4.
#pragma once

#include <hip/hip_runtime.h>

#include "sgh_common.h"

namespace sgh {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

template <typename F>
struct Mfma;
template <>
struct Mfma<float> {
  using V = f4v;
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // output row i of accumulator register v in this lane
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <>
struct Mfma<double> {
  using V = d4v;
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};

__device__ __forceinline__ f4v mma_bf16(bf8v a, bf8v b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Split x = hi + lo with hi = bf16(x) (round to nearest even) and lo = bf16(x − hi):
// hi·hi' + hi·lo' + lo·hi' reproduces a product to about 2^-16 relative.
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// LDS leading dimensions (in elements) for conflict-free MFMA operand reads: an A
// operand (16 rows × 4 k, row-major) wants a row stride ≡ 2 words (mod 32), a B operand
// (4 k-rows × 16 columns) a row stride ≡ 16 words (mod 32); a double is 2 words.
template <typename F>
__host__ __device__ constexpr int pad_to_mod(int n, int target_words) {
  const int wpe = (int)(sizeof(F) / 4);
  const int words = n * wpe;
  const int add = ((target_words - words % 32) % 32 + 32) % 32;
  return n + (add + wpe - 1) / wpe;
}
template <typename F>
__host__ __device__ constexpr int pad_a(int n) { return pad_to_mod<F>(n, 2 * (int)(sizeof(F) / 4)); }
template <typename F>
__host__ __device__ constexpr int pad_b(int n) { return pad_to_mod<F>(n, 16); }

// Device-coherent access (COH): loads and stores with device scope (relaxed atomics,
// sc1 on gfx950), which bypass the CU's L1 — for data that other workgroups write or read
// within the same kernel, ordered by the completion waits of a barrier.
template <typename F, bool COH>
__device__ __forceinline__ F ld_x(const F* p) {
  if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}
template <typename F, bool COH>
__device__ __forceinline__ void st_x(F* p, F v) {
  if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else *p = v;
}
constexpr int kCoherentAux = 16;   // cache-policy operand of an LDS DMA: sc1 (device scope)

// Copy a rows × cols block (source row stride src_ld, starting at src) into LDS
// (row stride dst_ld), zero-filling rows >= valid_rows and columns >= valid_cols.
// Each thread keeps A global loads in flight; indices advance incrementally
// (no per-element division). COH: device-coherent loads.
template <typename F, int NT, int A, bool COH = false>
__device__ __forceinline__ void stage_block(F* __restrict__ dst, int dst_ld, const F* __restrict__ src,
                                            int src_ld, int rows, int cols, int valid_rows,
                                            int valid_cols) {
  const int total = rows * cols;
  const int step_r = NT / cols, step_c = NT - (NT / cols) * cols;
  int r = threadIdx.x / cols, c = threadIdx.x - (threadIdx.x / cols) * cols;
  for (int e0 = threadIdx.x; e0 < total; e0 += A * NT) {
    F tmp[A];
    int rr[A], cc[A];
#pragma unroll
    for (int q = 0; q < A; ++q) {
      rr[q] = r;
      cc[q] = c;
      const bool in = e0 + q * NT < total;
      tmp[q] = (in && r < valid_rows && c < valid_cols) ? ld_x<F, COH>(src + (long long)r * src_ld + c)
                                                         : F(0);
      r += step_r;
      c += step_c;
      if (c >= cols) { c -= cols; ++r; }
    }
#pragma unroll
    for (int q = 0; q < A; ++q)
      if (e0 + q * NT < total) dst[rr[q] * dst_ld + cc[q]] = tmp[q];
  }
}

// Element offset XORed into the columns of odd rows of an unpadded LDS operand so that
// rows kk and kk+1 of an MFMA B operand fall in different halves of the 32 banks; 0
// when the row length already staggers them (or cannot be swizzled within a row).
template <typename F>
__device__ __forceinline__ int swz_of(int ld) {
  constexpr int half = sizeof(F) == 4 ? 16 : 8;  // 16 words = half the banks
  return ld % (2 * half) == 0 ? half : 0;
}

// LDS DMA (gfx950 global_load_lds_dwordx4) of `rows` rows of `ld` elements from global
// `src` (row stride `src_ld`) into LDS `dst` packed as [rows][ld], the columns of odd rows
// XOR-swizzled by `swz` elements: every 16-byte unit of LDS receives the 16-byte unit of
// its unswizzled position. Each wave copies 1 KiB per instruction (lane l -> dst unit
// base + l). The caller waits (s_waitcnt + barrier) before reading. Needs 16-byte aligned
// src rows and dst, ld·sizeof(F) a multiple of 16 and swz a multiple of 16 / sizeof(F).
template <int AUX = 0>
__device__ __forceinline__ void dma16(const void* src, void* lds_base) {
  __builtin_amdgcn_global_load_lds(src, lds_base, 16, 0, AUX);
}

template <typename F, int NT, int AUX = 0>
__device__ __forceinline__ void dma_rows(F* dst, const F* src, int rows, int ld, int src_ld,
                                         int swz) {
  constexpr int E = 16 / (int)sizeof(F);
  const int units = rows * ld / E;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (int u0 = wave * kWave; u0 < units; u0 += NT) {
    const int u = u0 + lane;
    if (u < units) {
      const int q = u * E;
      const int r = q / ld;
      const int cs = q - r * ld;
      const int c = (r & 1) ? (cs ^ swz) : cs;
      dma16<AUX>(src + (long long)r * src_ld + c, dst + (long long)u0 * E);
    }
  }
}

}  // namespace sgh
