// Separable cosine transforms as dense axis contractions.
//
// The ortho DCT-II / DCT-III pair of the reference (scipy.fft.dctn on CPU, reorder +
// cuFFT + recombine on GPU) is, per axis, a multiplication by an n×n cosine matrix.
// On MI355X the matrices are tiny (n <= 256) and stay in L2/LDS, so each transform is
// a batched small GEMM with no transposes:
//   sep2d        : out = scale · L · X · R for a surface, one workgroup per 8-row block,
//                  both products staged through LDS in one launch;
//   axis_contract: one axis of a 3-D volume, an LDS-tiled strided GEMM
//                  (x axis: Out[R, n] = In[R, n]·Mᵀ; y/z axes: Out_o = M·In_o).
#include "sgh_common.h"

namespace sgh {

constexpr int kT = 256;
constexpr int kRB = 8;      // output rows per sep2d workgroup
constexpr int kKC = 16;     // contraction chunk staged in LDS
constexpr int kMaxNx = 256; // widest surface axis handled by sep2d
constexpr int kMaxO = kRB * kMaxNx / kT;

template <typename F>
__global__ void __launch_bounds__(kT)
k_sep2d(int ny, int nx, const F* __restrict__ L, const F* __restrict__ R,
        const F* __restrict__ X, int nparts, F scale, F* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  F* Bs = reinterpret_cast<F*>(smem_raw);   // [kKC][nx]
  F* Ys = Bs + kKC * nx;                     // [kRB][nx]
  F* As = Ys + kRB * nx;                     // [kRB][kKC]
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * kRB;
  const int rows = min(kRB, ny - r0);
  const long long plane = (long long)ny * nx;
  const F* Xb = X + (long long)b * nparts * plane;
  F* Ob = out + (long long)b * plane;
  const int nout = rows * nx;

  int ro[kMaxO], co[kMaxO];
  F acc[kMaxO];
#pragma unroll
  for (int j = 0; j < kMaxO; ++j) {
    const int o = tid + j * kT;
    ro[j] = o / nx;
    co[j] = o - ro[j] * nx;
    acc[j] = F(0);
  }

  // Y = L[r0:r0+rows, :] · X
  for (int k0 = 0; k0 < ny; k0 += kKC) {
    const int kc = min(kKC, ny - k0);
    for (int e = tid; e < kc * nx; e += kT) {
      const int kk = e / nx, c = e - kk * nx;
      const long long off = (long long)(k0 + kk) * nx + c;
      F v = Xb[off];
      for (int q = 1; q < nparts; ++q) v += Xb[q * plane + off];
      Bs[kk * nx + c] = v;
    }
    for (int e = tid; e < rows * kKC; e += kT) {
      const int r = e / kKC, kk = e - r * kKC;
      As[r * kKC + kk] = kk < kc ? L[(long long)(r0 + r) * ny + k0 + kk] : F(0);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kMaxO; ++j) {
      if (tid + j * kT < nout) {
        const F* a = As + ro[j] * kKC;
        const F* bcol = Bs + co[j];
        F s = acc[j];
        for (int kk = 0; kk < kc; ++kk) s += a[kk] * bcol[kk * nx];
        acc[j] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kMaxO; ++j) {
    if (tid + j * kT < nout) Ys[tid + j * kT] = acc[j];
    acc[j] = F(0);
  }
  __syncthreads();

  // out = Y · R
  for (int k0 = 0; k0 < nx; k0 += kKC) {
    const int kc = min(kKC, nx - k0);
    for (int e = tid; e < kc * nx; e += kT) {
      const int kk = e / nx, c = e - kk * nx;
      Bs[kk * nx + c] = R[(long long)(k0 + kk) * nx + c];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kMaxO; ++j) {
      if (tid + j * kT < nout) {
        const F* y = Ys + ro[j] * nx + k0;
        const F* bcol = Bs + co[j];
        F s = acc[j];
        for (int kk = 0; kk < kc; ++kk) s += y[kk] * bcol[kk * nx];
        acc[j] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kMaxO; ++j)
    if (tid + j * kT < nout) Ob[(long long)(r0 + ro[j]) * nx + co[j]] = scale * acc[j];
}

// C[b][r][c] = scale · Σ_k A[b][r][k] · B[b][k][c] with arbitrary strides.
constexpr int kTM = 64, kTN = 64, kTK = 16;

template <typename F>
__global__ void __launch_bounds__(kT)
k_gemm_strided(int M, int N, int K, const F* __restrict__ A, long long sAb, long long sAr,
               long long sAk, const F* __restrict__ B, long long sBb, long long sBk,
               long long sBc, F* __restrict__ C, long long sCb, long long sCr, long long sCc,
               F scale) {
  __shared__ F As[kTK][kTM + 1];
  __shared__ F Bs[kTK][kTN + 1];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int c0 = blockIdx.x * kTN, r0 = blockIdx.y * kTM;
  const long long b = blockIdx.z;
  const F* Ab = A + b * sAb;
  const F* Bb = B + b * sBb;
  F acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = F(0);
  const bool a_k_fast = sAk == 1;
  const bool b_c_fast = sBc == 1;
  for (int k0 = 0; k0 < K; k0 += kTK) {
#pragma unroll
    for (int q = 0; q < (kTM * kTK) / kT; ++q) {
      const int e = tid + q * kT;
      int r, k;
      if (a_k_fast) { k = e % kTK; r = e / kTK; } else { r = e % kTM; k = e / kTM; }
      const int gr = r0 + r, gk = k0 + k;
      As[k][r] = (gr < M && gk < K) ? Ab[gr * sAr + gk * sAk] : F(0);
    }
#pragma unroll
    for (int q = 0; q < (kTN * kTK) / kT; ++q) {
      const int e = tid + q * kT;
      int c, k;
      if (b_c_fast) { c = e % kTN; k = e / kTN; } else { k = e % kTK; c = e / kTK; }
      const int gc = c0 + c, gk = k0 + k;
      Bs[k][c] = (gc < N && gk < K) ? Bb[gk * sBk + gc * sBc] : F(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; ++kk) {
      F a[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * bv[j];
    }
    __syncthreads();
  }
  F* Cb = C + b * sCb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gr = r0 + ty + 16 * i;
    if (gr >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gc = c0 + tx + 16 * j;
      if (gc < N) Cb[gr * sCr + gc * sCc] = scale * acc[i][j];
    }
  }
}

}  // namespace sgh

using namespace sgh;

extern "C" {

int sgh_sep2d(int dtype, int batch, int ny, int nx, const void* L, const void* R, const void* X,
              int nparts, double scale, void* out, void* stream) {
  SGH_REQUIRE(batch > 0 && ny > 0 && nx > 0 && L && R && X && out && nparts >= 1, SGH_ERR_ARG,
              "sgh_sep2d: bad args");
  SGH_REQUIRE(nx <= kMaxNx, SGH_ERR_ARG, "sgh_sep2d: nx=%d exceeds %d", nx, kMaxNx);
  SGH_REQUIRE(X != out, SGH_ERR_ARG, "sgh_sep2d: in-place transform is not supported");
  SGH_REQUIRE(batch <= 65535, SGH_ERR_ARG, "sgh_sep2d: batch too large");
  SGH_DISPATCH(dtype, {
    const size_t shm = (size_t)(kKC * nx + kRB * nx + kRB * kKC) * sizeof(F);
    hipLaunchKernelGGL(k_sep2d<F>, dim3(cdiv(ny, kRB), batch), dim3(kT), shm,
                       reinterpret_cast<hipStream_t>(stream), ny, nx, (const F*)L, (const F*)R,
                       (const F*)X, nparts, (F)scale, (F*)out);
  });
  SGH_CHECK_LAUNCH("sgh_sep2d");
  return SGH_OK;
}

int sgh_axis_contract(int dtype, long long outer, int n, long long inner, const void* M,
                      const void* in, double scale, void* out, void* stream) {
  SGH_REQUIRE(outer > 0 && n > 0 && inner > 0 && M && in && out && in != out, SGH_ERR_ARG,
              "sgh_axis_contract: bad args");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SGH_DISPATCH(dtype, {
    if (inner == 1) {
      // Out[o, k] = Σ_j In[o, j]·M[k, j]
      SGH_REQUIRE(outer <= (1LL << 31) - 1, SGH_ERR_ARG, "sgh_axis_contract: outer too large");
      dim3 grid(cdiv(n, kTN), cdiv(outer, kTM), 1);
      hipLaunchKernelGGL(k_gemm_strided<F>, grid, dim3(kT), 0, st, (int)outer, n, n,
                         (const F*)in, 0LL, (long long)n, 1LL, (const F*)M, 0LL, 1LL,
                         (long long)n, (F*)out, 0LL, (long long)n, 1LL, (F)scale);
    } else {
      // Out_o[k, i] = Σ_j M[k, j]·In_o[j, i]
      SGH_REQUIRE(outer <= 65535 && inner <= (1LL << 31) - 1, SGH_ERR_ARG,
                  "sgh_axis_contract: outer/inner too large");
      dim3 grid(cdiv(inner, kTN), cdiv(n, kTM), (unsigned)outer);
      hipLaunchKernelGGL(k_gemm_strided<F>, grid, dim3(kT), 0, st, n, (int)inner, n,
                         (const F*)M, 0LL, (long long)n, 1LL, (const F*)in, (long long)n * inner,
                         inner, 1LL, (F*)out, (long long)n * inner, inner, 1LL, (F)scale);
    }
  });
  SGH_CHECK_LAUNCH("sgh_axis_contract");
  return SGH_OK;
}

}  // extern "C"
