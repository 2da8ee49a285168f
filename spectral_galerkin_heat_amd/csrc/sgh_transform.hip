// Separable cosine transforms as dense axis contractions.
//
// The ortho DCT-II / DCT-III pair of the reference (scipy.fft.dctn on CPU, reorder +
// cuFFT + recombine on GPU) is, per axis, a multiplication by an n×n cosine matrix.
// On MI355X the matrices are tiny (n <= 256) and stay in L2/LDS, so each transform is
// a batched small GEMM with no transposes:
//   sep2d        : out = scale · L · X · R for a surface on the matrix cores, one
//                  workgroup per 16-row block, both products streamed through LDS in
//                  one launch;
//   axis_contract: one axis of a 3-D volume as a strided batched GEMM on the matrix
//                  cores (x axis: Out[R, n] = In[R, n]·Mᵀ; y/z axes: Out_o = M·In_o),
//                  64×64 output tiles, 32-deep K slices staged in LDS with the next
//                  slice prefetched into registers. Variants: MFMA in the data type
//                  (f32 16x16x4 or f64 16x16x4), split-BF16 (three bf16 16x16x32
//                  products hi·hi + hi·lo + lo·hi, FP32 accumulation, FP32 data), and
//                  the LDS-tiled vector-ALU GEMM.
#include <algorithm>

#include "sgh_common.h"
#include "sgh_mfma.h"
#include "sgh_props.h"

namespace sgh {

constexpr int kT = 256;
constexpr int kSRows = 16;    // surface rows per sep2d workgroup (one MFMA row tile)
constexpr int kSCols = 64;    // output columns per sep2d workgroup
constexpr int kMaxNx = 256;   // widest surface axis handled by sep2d
constexpr int kSChunk = 8192; // elements of the LDS buffer that streams X, then R (at most)
constexpr int kSLds = 160 * 1024;  // gfx950 LDS per workgroup

// out[b] = scale · L · (Σ_p X[b][p]) · R for an ny × nx surface, on the matrix cores.
// Workgroup (row block, b, column block) owns 16 output rows × kSCols output columns:
// stage 1 streams X (summing the nparts partial surfaces on load) through LDS in
// K-chunks of rows and accumulates the full Y = L[rows, :] · X; stage 2 streams only
// its columns of R and accumulates Y · R[:, cols]. 4 waves, 16-column tiles dealt
// round-robin. Callers sum z-chunk partials first (sgh_sum_parts) so that X streams as
// one plane.
template <typename F>
__global__ void __launch_bounds__(kT)
k_sep2d(int ny, int nx, const F* __restrict__ L, const F* __restrict__ R,
        const F* __restrict__ X, int nparts, F scale, F* __restrict__ out, int chunk) {
  using M = Mfma<F>;
  using V = typename M::V;
  constexpr int NW = kT / 64;
  constexpr int TPW = (kMaxNx / 16 + NW - 1) / NW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ny4 = (ny + 3) & ~3, nx4 = (nx + 3) & ~3;
  const int ldl = pad_a<F>(ny4), ldy = pad_a<F>(nx4), ldb = pad_b<F>(nx);
  F* Ls = reinterpret_cast<F*>(smem_raw);   // [16][ldl]
  F* Ys = Ls + kSRows * ldl;                // [16][ldy]
  F* Bs = Ys + kSRows * ldy;                // [kr][ldb]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * kSRows;
  const int rows = min(kSRows, ny - r0);
  const long long plane = (long long)ny * nx;
  const F* Xb = X + (long long)b * nparts * plane;
  F* Ob = out + (long long)b * plane;
  const int ntile = (nx + 15) >> 4;
  const int kr = (chunk / ldb) & ~3;   // K rows per streamed chunk (the host makes it >= 4)

  for (int e = tid; e < kSRows * ny4; e += kT) {
    const int r = e / ny4, c = e - r * ny4;
    Ls[r * ldl + c] = (r < rows && c < ny) ? L[(long long)(r0 + r) * ny + c] : F(0);
  }
  V acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = V{0, 0, 0, 0};
  // stage 1: Y = L[rows, :] · Σ_p X_p
  for (int k0 = 0; k0 < ny4; k0 += kr) {
    const int kn = min(kr, ny4 - k0);
    __syncthreads();
    if (nparts == 1) {
      stage_block<F, kT, 4>(Bs, ldb, Xb + (long long)k0 * nx, nx, kn, nx, ny - k0, nx);
    } else {
      for (int e = tid; e < kn * nx; e += kT) {
        const int kk = e / nx, c = e - kk * nx;
        F v = F(0);
        if (k0 + kk < ny) {
          const long long off = (long long)(k0 + kk) * nx + c;
          v = Xb[off];
          for (int q = 1; q < nparts; ++q) v += Xb[q * plane + off];
        }
        Bs[kk * ldb + c] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int jt = wave + NW * t;
      if (jt >= ntile) break;
      const int col = jt * 16 + li;
      const bool cin = col < nx;
      for (int kk = 0; kk < kn; kk += 4) {
        const F bv = cin ? Bs[(kk + lk) * ldb + col] : F(0);
        acc[t] = M::mma(Ls[li * ldl + k0 + kk + lk], bv, acc[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int jt = wave + NW * t;
    const int col = jt * 16 + li;
    if (jt < ntile && col < nx) {
#pragma unroll
      for (int v = 0; v < 4; ++v) Ys[M::row(lane, v) * ldy + col] = acc[t][v];
    }
    acc[t] = V{0, 0, 0, 0};
  }
  for (int e = tid; e < kSRows * (nx4 - nx); e += kT) {
    const int r = e / (nx4 - nx), c = nx + e % (nx4 - nx);
    Ys[r * ldy + c] = F(0);
  }
  // stage 2: out[:, c0 : c0 + kSCols] = Y · R[:, c0 : ...], R streamed in K-chunks of rows
  const int c0 = blockIdx.z * kSCols, cn = min(kSCols, nx - c0);
  const int ldb2 = pad_b<F>(kSCols);
  const int kr2 = (chunk / ldb2) & ~3;
  V acc2 = V{0, 0, 0, 0};
  const int jt2 = wave;                  // one 16-column tile per wave
  const int col2 = c0 + jt2 * 16 + li;
  for (int k0 = 0; k0 < nx4; k0 += kr2) {
    const int kn = min(kr2, nx4 - k0);
    __syncthreads();
    stage_block<F, kT, 4>(Bs, ldb2, R + (long long)k0 * nx + c0, nx, kn, kSCols, nx - k0, cn);
    __syncthreads();
    if (jt2 * 16 < cn) {
      const bool cin = col2 < nx;
      for (int kk = 0; kk < kn; kk += 4) {
        const F bv = cin ? Bs[(kk + lk) * ldb2 + jt2 * 16 + li] : F(0);
        acc2 = M::mma(Ys[li * ldy + k0 + kk + lk], bv, acc2);
      }
    }
  }
  {
    const int col = col2;
    if (jt2 * 16 >= cn || col >= nx) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = M::row(lane, v);
      if (row < rows) Ob[(long long)(r0 + row) * nx + col] = scale * acc2[v];
    }
  }
}

// Surfaces whose operands fit in LDS (and rows of a multiple of 16 bytes): everything the
// workgroup reads lands in one round -- its 16 rows of L through registers, all of X and
// its kSCols columns of R by LDS DMA (gfx950 global_load_lds_dwordx4, odd rows
// swizzled), one wait -- then both MFMA stages run out of LDS. FORCE: the output is a
// top-surface temperature and the epilogue writes its evaporation and net fluxes
// (q_evap, q_net = out) instead.
template <typename F, bool FORCE = false>
__global__ void __launch_bounds__(kT)
k_sep2d_direct(int ny, int nx, const F* __restrict__ L, const F* __restrict__ R,
               const F* __restrict__ X, F scale, F* __restrict__ out,
               ForcingParams<F> fp = {}, const F* __restrict__ q_las = nullptr,
               F* __restrict__ q_evap = nullptr) {
  using M = Mfma<F>;
  using V = typename M::V;
  constexpr int NW = kT / 64;
  constexpr int TPW = (kMaxNx / 16 + NW - 1) / NW;
  constexpr int kLPT = kSRows * kMaxNx / kT;   // L elements per thread
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int ny4 = (ny + 3) & ~3, nx4 = (nx + 3) & ~3;
  const int ldl = pad_a<F>(ny4), ldy = pad_a<F>(nx4);
  const int c0 = blockIdx.z * kSCols, cn = min(kSCols, nx - c0);
  F* Ls = reinterpret_cast<F*>(smem_raw);    // [16][ldl]
  F* Ys = Ls + kSRows * ldl;                 // [16][ldy]
  F* Xs = Ys + kSRows * ldy;                 // [ny4][nx]
  F* Rs = Xs + (long long)ny4 * nx;          // [nx4][cn]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * kSRows;
  const int rows = min(kSRows, ny - r0);
  const long long plane = (long long)ny * nx;
  const F* Xb = X + (long long)blockIdx.y * plane;
  F* Ob = out + (long long)blockIdx.y * plane;
  const int ntile = (nx + 15) >> 4;
  const int swx = swz_of<F>(nx), swr = swz_of<F>(cn);
  const int ltot = kSRows * ny4;

  F lr[kLPT];
#pragma unroll
  for (int q = 0; q < kLPT; ++q) {
    const int e = tid + q * kT;
    const int r = e / ny4, c = e - r * ny4;
    lr[q] = (e < ltot && r < rows && c < ny) ? L[(long long)(r0 + r) * ny + c] : F(0);
  }
  dma_rows<F, kT>(Xs, Xb, ny, nx, nx, swx);
  dma_rows<F, kT>(Rs, R + c0, nx, cn, nx, swr);
  for (int e = tid; e < (ny4 - ny) * nx; e += kT) Xs[ny * nx + e] = F(0);
  for (int e = tid; e < (nx4 - nx) * cn; e += kT) Rs[nx * cn + e] = F(0);
#pragma unroll
  for (int q = 0; q < kLPT; ++q) {
    const int e = tid + q * kT;
    if (e < ltot) Ls[(e / ny4) * ldl + (e - (e / ny4) * ny4)] = lr[q];
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  // stage 1: Y = L[rows, :] · X
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int jt = wave + NW * t;
    if (jt >= ntile) break;
    const int col = jt * 16 + li;
    const bool cin = col < nx;
    const int colx = (lk & 1) ? (col ^ swx) : col;
    V acc = V{0, 0, 0, 0};
    for (int kk = 0; kk < ny4; kk += 4) {
      const F bv = cin ? Xs[(kk + lk) * nx + colx] : F(0);
      acc = M::mma(Ls[li * ldl + kk + lk], bv, acc);
    }
    if (cin) {
#pragma unroll
      for (int v = 0; v < 4; ++v) Ys[M::row(lane, v) * ldy + col] = acc[v];
    }
  }
  for (int e = tid; e < kSRows * (nx4 - nx); e += kT) {
    const int r = e / (nx4 - nx), c = nx + e % (nx4 - nx);
    Ys[r * ldy + c] = F(0);
  }
  __syncthreads();
  // stage 2: out[:, c0 + 16·wave + (0..15)] = Y · R[:, those columns]
  if (wave * 16 >= cn) return;
  const int cl = wave * 16 + li;
  const bool cin = cl < cn;
  const int clx = (lk & 1) ? (cl ^ swr) : cl;
  V acc2 = V{0, 0, 0, 0};
  for (int kk = 0; kk < nx4; kk += 4) {
    const F bv = cin ? Rs[(kk + lk) * cn + clx] : F(0);
    acc2 = M::mma(Ys[li * ldy + kk + lk], bv, acc2);
  }
  if (!cin) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = M::row(lane, v);
    if (row < rows) {
      const long long o = (long long)(r0 + row) * nx + c0 + cl;
      if constexpr (FORCE) {
        F qe, qn;
        surface_forcing_value<F>(fp, scale * acc2[v], q_las ? q_las[o] : F(0), qe, qn);
        if (q_evap) q_evap[o] = qe;
        Ob[o] = qn;
      } else {
        Ob[o] = scale * acc2[v];
      }
    }
  }
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


// ---------------------------------------------------------------------------------
// Strided batched GEMM on the matrix cores:
//   C[b][r][c] = scale · Σ_k A[b][r][k] · B[b][k][c]
// A and C are row-major with unit stride along k and c; B is column-fast (B_CF,
// unit stride along c: the y/z axes, where B is the volume) or k-fast (the x axis,
// where B = Mᵀ). A 256-thread workgroup computes a 64×64 tile: wave w owns the 32×32
// quadrant (w/2, w%2) as 2×2 MFMA tiles of 16×16. K advances in slices of 32: the
// slice is staged in LDS while the next one is already loading into registers.
// ---------------------------------------------------------------------------------
// Workgroups are dealt to the 8 XCDs round-robin by linear id, each XCD with its own L2.
// Renumber them so that XCD x takes one contiguous range of tiles (x-fastest in the grid):
// the column tiles of one row block (which read the same rows of A) then run on one XCD
// at about the same time, and all but the first read those rows from its L2.
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned n = gx * gy * gridDim.z;
  const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned per = n / 8u, rem = n % 8u, x = lin % 8u, i = lin / 8u;
  const unsigned t = (x < rem ? x * (per + 1u) : rem * (per + 1u) + (x - rem) * per) + i;
  bx = (int)(t % gx);
  by = (int)((t / gx) % gy);
  bz = (int)(t / (gx * gy));
}

constexpr int kGM = 64, kGN = 64, kGK = 32;
constexpr int kGL = kGM * kGK / kT;   // elements of A (and of B) per thread per slice
constexpr int kBfLd = kGK + 8;        // bf16 LDS row: 80 bytes, 16-byte aligned reads

// LAT: the epilogue turns each output T into the latent source latent_value(T, Tp) with
// Tp laid out like C, so a reconstruction's last contraction emits the source directly.
// TM: rows of the output tile, 64 (four waves as 2×2, each 32 rows × 32 columns) or 32
// (four waves side by side, each 32 rows × 16 columns) — the latter doubles the workgroups
// of a contraction whose 64-row grid would leave most CUs idle.
template <typename F, int VAR, bool B_CF, bool LAT = false, int TM = kGM>
__global__ void __launch_bounds__(kT)
k_gemm_mfma(int M, int N, int K, const F* __restrict__ A, long long sAb, long long sAr,
            const F* __restrict__ B, long long sBb, long long sBk, long long sBc,
            F* __restrict__ C, long long sCb, long long sCr, F scale,
            const F* __restrict__ Tp = nullptr, LatentParams<F> lp = {}) {
  static_assert(VAR == 0 || (VAR == 1 && sizeof(F) == 4), "split-BF16 takes FP32 data");
  static_assert(TM == 64 || TM == 32, "output tile rows");
  constexpr int TT = TM == 64 ? 2 : 1;             // 16-column tiles per wave
  constexpr int CW = 16 * TT;                      // columns per wave
  constexpr int kGLA = TM * kGK / kT;              // A elements per thread per slice
  using Acc = typename Mfma<F>::V;
  constexpr int ldA = pad_a<F>(kGK), ldB = pad_b<F>(kGN);
  constexpr int kSmemF = VAR == 0 ? (kGM * ldA + kGK * ldB) * (int)sizeof(F) : 16;
  constexpr int kSmemB = VAR == 1 ? 4 * kGM * kBfLd * 2 : 16;
  __shared__ __align__(16) unsigned char sm_f[kSmemF];
  __shared__ __align__(16) unsigned char sm_b[kSmemB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = TM == 64 ? wave >> 1 : 0, wc = TM == 64 ? wave & 1 : wave;
  int tbx, tby, tbz;
  xcd_tile(tbx, tby, tbz);
  const int r0 = tby * TM, c0 = tbx * kGN;
  const long long bz = tbz;
  const F* Ab = A + bz * sAb;
  const F* Bb = B + bz * sBb;

  // this thread's staging coordinates (slice-relative): A (r, k), B (k, c)
  const int ar = tid / kGK, ak = tid % kGK;
  const int bk = B_CF ? tid / kGN : tid % kGK;
  const int bc = B_CF ? tid % kGN : tid / kGK;
  constexpr int a_step = kT / kGK;                 // rows of A per q
  constexpr int b_step = B_CF ? kT / kGN : kT / kGK;  // k rows (B_CF) or columns per q

  F ra[kGLA], rb[kGL];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < kGLA; ++q) {
      const int r = r0 + ar + q * a_step, k = k0 + ak;
      ra[q] = (r < M && k < K) ? Ab[(long long)r * sAr + k] : F(0);
    }
#pragma unroll
    for (int q = 0; q < kGL; ++q) {
      const int kb = k0 + (B_CF ? bk + q * b_step : bk);
      const int cb = c0 + (B_CF ? bc : bc + q * b_step);
      rb[q] = (kb < K && cb < N) ? Bb[(long long)kb * sBk + (long long)cb * sBc] : F(0);
    }
  };

  Acc acc[2][TT];
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < TT; ++t) acc[m][t] = Acc{0, 0, 0, 0};

  load(0);
  for (int k0 = 0; k0 < K; k0 += kGK) {
    if constexpr (VAR == 0) {
      F* As = reinterpret_cast<F*>(sm_f);
      F* Bs = As + kGM * ldA;
#pragma unroll
      for (int q = 0; q < kGLA; ++q) As[(ar + q * a_step) * ldA + ak] = ra[q];
#pragma unroll
      for (int q = 0; q < kGL; ++q) {
        if (B_CF) Bs[(bk + q * b_step) * ldB + bc] = rb[q];
        else Bs[bk * ldB + bc + q * b_step] = rb[q];
      }
    } else {
      __bf16* Ah = reinterpret_cast<__bf16*>(sm_b);
      __bf16* Al = Ah + kGM * kBfLd;
      __bf16* Bh = Al + kGM * kBfLd;   // column-major: [c][k]
      __bf16* Bl = Bh + kGN * kBfLd;
#pragma unroll
      for (int q = 0; q < kGLA; ++q) {
        const int ia = (ar + q * a_step) * kBfLd + ak;
        split_bf16((float)ra[q], Ah[ia], Al[ia]);
      }
#pragma unroll
      for (int q = 0; q < kGL; ++q) {
        const int ib = B_CF ? bc * kBfLd + bk + q * b_step : (bc + q * b_step) * kBfLd + bk;
        split_bf16((float)rb[q], Bh[ib], Bl[ib]);
      }
    }
    __syncthreads();
    if (k0 + kGK < K) load(k0 + kGK);
    if constexpr (VAR == 0) {
      const F* As = reinterpret_cast<const F*>(sm_f);
      const F* Bs = As + kGM * ldA;
#pragma unroll
      for (int kk = 0; kk < kGK; kk += 4) {
        F av[2], bv[TT];
#pragma unroll
        for (int m = 0; m < 2; ++m) av[m] = As[(wr * 32 + 16 * m + li) * ldA + kk + lk];
#pragma unroll
        for (int t = 0; t < TT; ++t) bv[t] = Bs[(kk + lk) * ldB + wc * CW + 16 * t + li];
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
          for (int t = 0; t < TT; ++t) acc[m][t] = Mfma<F>::mma(av[m], bv[t], acc[m][t]);
      }
    } else {
      const __bf16* Ah = reinterpret_cast<const __bf16*>(sm_b);
      const __bf16* Al = Ah + kGM * kBfLd;
      const __bf16* Bh = Al + kGM * kBfLd;
      const __bf16* Bl = Bh + kGN * kBfLd;
      bf8v ah[2], al[2], bh[TT], bl[TT];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const int o = (wr * 32 + 16 * m + li) * kBfLd + 8 * lk;
        ah[m] = *reinterpret_cast<const bf8v*>(Ah + o);
        al[m] = *reinterpret_cast<const bf8v*>(Al + o);
      }
#pragma unroll
      for (int t = 0; t < TT; ++t) {
        const int o = (wc * CW + 16 * t + li) * kBfLd + 8 * lk;
        bh[t] = *reinterpret_cast<const bf8v*>(Bh + o);
        bl[t] = *reinterpret_cast<const bf8v*>(Bl + o);
      }
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int t = 0; t < TT; ++t) {
          f4v a4 = reinterpret_cast<f4v&>(acc[m][t]);
          a4 = mma_bf16(al[m], bh[t], a4);
          a4 = mma_bf16(ah[m], bl[t], a4);
          a4 = mma_bf16(ah[m], bh[t], a4);
          acc[m][t] = reinterpret_cast<Acc&>(a4);
        }
    }
    __syncthreads();
  }

  F* Cb = C + bz * sCb;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < TT; ++t) {
      const int col = c0 + wc * CW + 16 * t + li;
      if (col >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + wr * 32 + 16 * m + Mfma<F>::row(lane, v);
        if (row < M) {
          const long long o = (long long)row * sCr + col;
          F val = scale * acc[m][t][v];
          if constexpr (LAT) val = latent_value<F>(lp, val, Tp[bz * sCb + o]);
          Cb[o] = val;
        }
      }
    }
}

}  // namespace sgh

using namespace sgh;

namespace {

// M is n_out × n_in with row stride ldm (ldm > n_in: a column window of a wider basis).
// Output tiles of 64 rows, or of 32 when 64-row tiles would leave the CUs mostly idle (fewer
// workgroups than kFillWgs<F>) or the output has no more than 32 rows. FP32 contractions
// gain from 32-row tiles up to 1024 workgroups (box mode 3.43 -> 3.25 ms/step); FP64 ones do
// not (4.03 -> 4.04).
template <typename F> constexpr long long kFillWgs = sizeof(F) == 4 ? 1024 : 256;
template <typename F, int VAR, int TM>
void launch_axis_tm(long long outer, int n_in, int n_out, long long inner, const F* M,
                    long long ldm, const F* in, F scale, F* out, hipStream_t st, const F* Tp,
                    LatentParams<F> lp) {
  if (inner == 1) {
    // Out[o, k] = Σ_j In[o, j]·M[k, j]: A = In (outer × n_in), B[j][k] = M[k][j] (k-fast)
    dim3 grid(cdiv(n_out, kGN), cdiv(outer, TM), 1);
    if (Tp)
      hipLaunchKernelGGL((k_gemm_mfma<F, VAR, false, true, TM>), grid, dim3(kT), 0, st,
                         (int)outer, n_out, n_in, in, 0LL, (long long)n_in, M, 0LL, 1LL, ldm,
                         out, 0LL, (long long)n_out, scale, Tp, lp);
    else
      hipLaunchKernelGGL((k_gemm_mfma<F, VAR, false, false, TM>), grid, dim3(kT), 0, st,
                         (int)outer, n_out, n_in, in, 0LL, (long long)n_in, M, 0LL, 1LL, ldm,
                         out, 0LL, (long long)n_out, scale, nullptr, LatentParams<F>{});
  } else {
    // Out_o[k, i] = Σ_j M[k, j]·In_o[j, i]: A = M, B = In_o (column-fast), batched over o
    dim3 grid(cdiv(inner, kGN), cdiv(n_out, TM), (unsigned)outer);
    hipLaunchKernelGGL((k_gemm_mfma<F, VAR, true, false, TM>), grid, dim3(kT), 0, st, n_out,
                       (int)inner, n_in, M, 0LL, ldm, in, (long long)n_in * inner, inner, 1LL,
                       out, (long long)n_out * inner, inner, scale, nullptr, LatentParams<F>{});
  }
}

template <typename F, int VAR>
void launch_axis_mfma(long long outer, int n_in, int n_out, long long inner, const F* M,
                      long long ldm, const F* in, F scale, F* out, hipStream_t st,
                      const F* Tp = nullptr, LatentParams<F> lp = {}) {
  const long long rows = inner == 1 ? outer : n_out;
  const long long wgs = inner == 1 ? (long long)cdiv(n_out, kGN) * cdiv(outer, kGM)
                                   : (long long)cdiv(inner, kGN) * cdiv(n_out, kGM) * outer;
  if (rows <= 32 || wgs < kFillWgs<F>)
    launch_axis_tm<F, VAR, 32>(outer, n_in, n_out, inner, M, ldm, in, scale, out, st, Tp, lp);
  else
    launch_axis_tm<F, VAR, 64>(outer, n_in, n_out, inner, M, ldm, in, scale, out, st, Tp, lp);
}

}  // namespace

extern "C" {

int sgh_sep2d(int dtype, int batch, int ny, int nx, const void* L, const void* R, const void* X,
              int nparts, double scale, void* out, void* stream) {
  SGH_REQUIRE(batch > 0 && ny > 0 && nx > 0 && L && R && X && out && nparts >= 1, SGH_ERR_ARG,
              "sgh_sep2d: bad args");
  SGH_REQUIRE(nx <= kMaxNx, SGH_ERR_ARG, "sgh_sep2d: nx=%d exceeds %d", nx, kMaxNx);
  SGH_REQUIRE(X != out, SGH_ERR_ARG, "sgh_sep2d: in-place transform is not supported");
  SGH_REQUIRE(batch <= 65535, SGH_ERR_ARG, "sgh_sep2d: batch too large");
  SGH_DISPATCH(dtype, {
    const int ny4 = (ny + 3) & ~3, nx4 = (nx + 3) & ~3;
    const long long fixed = (long long)kSRows * (pad_a<F>(ny4) + pad_a<F>(nx4));
    const long long direct_elems = fixed + (long long)ny4 * nx + (long long)nx4 * std::min(kSCols, nx);
    const bool aligned = ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(R)) & 15) == 0;
    if (nparts == 1 && aligned && (nx * (int)sizeof(F)) % 16 == 0 &&
        direct_elems * (long long)sizeof(F) <= kSLds) {
      static bool direct_raised[2] = {false, false};
      if (!direct_raised[sizeof(F) == 8]) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sep2d_direct<F>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kSLds);
        if (e != hipSuccess) return fail_hip(e, "sgh_sep2d LDS attribute");
        direct_raised[sizeof(F) == 8] = true;
      }
      hipLaunchKernelGGL(k_sep2d_direct<F>, dim3(cdiv(ny, kSRows), batch, cdiv(nx, kSCols)),
                         dim3(kT), (size_t)direct_elems * sizeof(F),
                         reinterpret_cast<hipStream_t>(stream), ny, nx, (const F*)L, (const F*)R,
                         (const F*)X, (F)scale, (F*)out);
      break;   // leaves the dispatch block: the chunked path below is not needed
    }
    const long long chunk = std::min<long long>(kSChunk, kSLds / (long long)sizeof(F) - fixed);
    SGH_REQUIRE(chunk >= 4LL * pad_b<F>(nx), SGH_ERR_ARG,
                "sgh_sep2d: ny=%d is too large for one workgroup's LDS", ny);
    const size_t shm = (size_t)(fixed + chunk) * sizeof(F);
    static bool lds_raised[2] = {false, false};
    if (!lds_raised[sizeof(F) == 8]) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sep2d<F>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, kSLds);
      if (e != hipSuccess) return fail_hip(e, "sgh_sep2d LDS attribute");
      lds_raised[sizeof(F) == 8] = true;
    }
    hipLaunchKernelGGL(k_sep2d<F>, dim3(cdiv(ny, kSRows), batch, cdiv(nx, kSCols)), dim3(kT), shm,
                       reinterpret_cast<hipStream_t>(stream), ny, nx, (const F*)L, (const F*)R,
                       (const F*)X, nparts, (F)scale, (F*)out, (int)chunk);
  });
  SGH_CHECK_LAUNCH("sgh_sep2d");
  return SGH_OK;
}

int sgh_sep2d_forcing(int dtype, int ny, int nx, const void* L, const void* R, const void* X,
                      double scale, const void* q_las, double P0, double T_boil, double dH,
                      double R_v, double T_liq, double h_top, double T0, const void* prog,
                      double kbar, void* T_s, void* q_evap, void* q_net, void* stream) {
  SGH_REQUIRE(ny > 0 && nx > 0 && L && R && X && q_net && T_s && X != q_net, SGH_ERR_ARG,
              "sgh_sep2d_forcing: bad args");
  SGH_REQUIRE(R_v > 0 && T_boil > 0, SGH_ERR_ARG,
              "sgh_sep2d_forcing: R_v and T_boil must be > 0");
  bool fused = false;
  SGH_DISPATCH(dtype, {
    const int ny4 = (ny + 3) & ~3, nx4 = (nx + 3) & ~3;
    const long long direct_elems = (long long)kSRows * (pad_a<F>(ny4) + pad_a<F>(nx4)) +
                                   (long long)ny4 * nx + (long long)nx4 * std::min(kSCols, nx);
    // the property interpreter's static LDS stack sits beside the dynamic tiles
    const long long stack = (long long)kProgMaxStack * kProgThreads * (long long)sizeof(F);
    const bool aligned =
        ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(R)) & 15) == 0;
    // with a property program the per-cell interpreter would run in the few workgroups
    // of the transform; the separate forcing kernel spreads it over the whole surface
    if (prog == nullptr && nx <= kMaxNx && aligned && (nx * (int)sizeof(F)) % 16 == 0 &&
        direct_elems * (long long)sizeof(F) + stack <= kSLds) {
      static bool raised[2] = {false, false};
      if (!raised[sizeof(F) == 8]) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&k_sep2d_direct<F, true>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kSLds - stack));
        if (e != hipSuccess) return fail_hip(e, "sgh_sep2d_forcing LDS attribute");
        raised[sizeof(F) == 8] = true;
      }
      hipLaunchKernelGGL((k_sep2d_direct<F, true>), dim3(cdiv(ny, kSRows), 1, cdiv(nx, kSCols)),
                         dim3(kT), (size_t)direct_elems * sizeof(F),
                         reinterpret_cast<hipStream_t>(stream), ny, nx, (const F*)L,
                         (const F*)R, (const F*)X, (F)scale, (F*)q_net,
                         make_forcing<F>(P0, T_boil, dH, R_v, T_liq, h_top, T0, prog, kbar),
                         (const F*)q_las, (F*)q_evap);
      fused = true;
    }
  });
  if (fused) {
    SGH_CHECK_LAUNCH("sgh_sep2d_forcing");
    return SGH_OK;
  }
  // surfaces the one-round kernel cannot take: transform into T_s, then the forcing
  int rc = sgh_sep2d(dtype, 1, ny, nx, L, R, X, 1, scale, T_s, stream);
  if (rc != SGH_OK) return rc;
  return sgh_surface_forcing(dtype, (long long)nx * ny, T_s, q_las, P0, T_boil, dH, R_v, T_liq,
                             h_top, T0, prog, kbar, q_evap, q_net, stream);
}

int sgh_axis_contract_strided(int dtype, int variant, long long outer, int n_in, int n_out,
                              long long inner, const void* M, long long ldm, const void* in,
                              double scale, void* out, void* stream) {
  SGH_REQUIRE(outer > 0 && n_in > 0 && n_out > 0 && inner > 0 && M && in && out && in != out,
              SGH_ERR_ARG, "sgh_axis_contract: bad args");
  SGH_REQUIRE(ldm >= n_in && ldm <= (1LL << 31) - 1, SGH_ERR_ARG,
              "sgh_axis_contract: ldm=%lld must be in [n_in=%d, 2^31)", ldm, n_in);
  SGH_REQUIRE(variant == SGH_XFORM_MFMA || variant == SGH_XFORM_BF16X3 ||
                  variant == SGH_XFORM_VALU,
              SGH_ERR_ARG, "sgh_axis_contract: unknown variant %d", variant);
  SGH_REQUIRE(!(variant == SGH_XFORM_BF16X3 && dtype != SGH_F32), SGH_ERR_ARG,
              "sgh_axis_contract: split-BF16 takes FP32 data");
  if (inner == 1)
    SGH_REQUIRE(outer <= (1LL << 31) - 1 && cdiv(outer, kGM) <= 65535u, SGH_ERR_ARG,
                "sgh_axis_contract: outer=%lld too large for an x-axis contraction", outer);
  else
    SGH_REQUIRE(outer <= 65535 && inner <= (1LL << 31) - 1 && cdiv(n_out, kGM) <= 65535u,
                SGH_ERR_ARG, "sgh_axis_contract: outer/inner too large");
  // the tile renumbering over the XCDs counts workgroups in 32 bits
  const long long tiles = inner == 1 ? (long long)cdiv(n_out, kGN) * cdiv(outer, kGM)
                                     : (long long)cdiv(inner, kGN) * cdiv(n_out, kGM) * outer;
  SGH_REQUIRE(tiles <= (1LL << 31) - 1, SGH_ERR_ARG,
              "sgh_axis_contract: %lld output tiles exceed 2^31 - 1", tiles);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (variant == SGH_XFORM_BF16X3) {
    launch_axis_mfma<float, 1>(outer, n_in, n_out, inner, (const float*)M, ldm,
                               (const float*)in, (float)scale, (float*)out, st);
  } else if (variant == SGH_XFORM_MFMA) {
    SGH_DISPATCH(dtype, launch_axis_mfma<F, 0>(outer, n_in, n_out, inner, (const F*)M, ldm,
                                               (const F*)in, (F)scale, (F*)out, st));
  } else {
    SGH_DISPATCH(dtype, {
      if (inner == 1) {
        dim3 grid(cdiv(n_out, kTN), cdiv(outer, kTM), 1);
        hipLaunchKernelGGL(k_gemm_strided<F>, grid, dim3(kT), 0, st, (int)outer, n_out, n_in,
                           (const F*)in, 0LL, (long long)n_in, 1LL, (const F*)M, 0LL, 1LL,
                           ldm, (F*)out, 0LL, (long long)n_out, 1LL, (F)scale);
      } else {
        dim3 grid(cdiv(inner, kTN), cdiv(n_out, kTM), (unsigned)outer);
        hipLaunchKernelGGL(k_gemm_strided<F>, grid, dim3(kT), 0, st, n_out, (int)inner, n_in,
                           (const F*)M, 0LL, ldm, 1LL, (const F*)in,
                           (long long)n_in * inner, inner, 1LL, (F*)out,
                           (long long)n_out * inner, inner, 1LL, (F)scale);
      }
    });
  }
  SGH_CHECK_LAUNCH("sgh_axis_contract");
  return SGH_OK;
}

int sgh_axis_contract_rect(int dtype, int variant, long long outer, int n_in, int n_out,
                           long long inner, const void* M, const void* in, double scale,
                           void* out, void* stream) {
  return sgh_axis_contract_strided(dtype, variant, outer, n_in, n_out, inner, M, n_in, in,
                                   scale, out, stream);
}

int sgh_axis_contract(int dtype, int variant, long long outer, int n, long long inner,
                      const void* M, const void* in, double scale, void* out, void* stream) {
  return sgh_axis_contract_rect(dtype, variant, outer, n, n, inner, M, in, scale, out, stream);
}

int sgh_box_latent_source(int dtype, int variant, int nx, int ny, int nz, const void* a,
                          int nxb, int nyb, int nzb, const void* BxT, long long ldx,
                          const void* ByT, long long ldy, const void* BzT, const void* T_prev,
                          double T_S, double T_L, double rho, double L_f, double dt,
                          const void* prog, void* work, void* Q, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nxb > 0 && nyb > 0 && nzb > 0 && a && BxT && ByT &&
                  BzT && T_prev && work && Q && dt > 0,
              SGH_ERR_ARG, "sgh_box_latent_source: bad args");
  SGH_REQUIRE(ldx >= nx && ldy >= ny, SGH_ERR_ARG, "sgh_box_latent_source: ldx/ldy too small");
  SGH_REQUIRE(variant == SGH_XFORM_MFMA || variant == SGH_XFORM_BF16X3 ||
                  variant == SGH_XFORM_VALU,
              SGH_ERR_ARG, "sgh_box_latent_source: unknown variant %d", variant);
  SGH_REQUIRE(!(variant == SGH_XFORM_BF16X3 && dtype != SGH_F32), SGH_ERR_ARG,
              "sgh_box_latent_source: split-BF16 takes FP32 data");
  SGH_REQUIRE(nzb <= 65535 && cdiv((long long)nzb * nyb, kGM) <= 65535u &&
                  (long long)cdiv(nxb, kGN) * cdiv((long long)nzb * nyb, kGM) <= (1LL << 31) - 1,
              SGH_ERR_ARG, "sgh_box_latent_source: box too large");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SGH_DISPATCH(dtype, {
    F* w1 = (F*)work;                          // (nzb, ny, nx)
    F* w2 = w1 + (long long)nzb * ny * nx;     // (nzb, nyb, nx)
    // z, then y, on the coarse x axis (the smallest intermediates), then x with the
    // latent source in the epilogue
    int rc = sgh_axis_contract_strided(dtype, variant, 1, nz, nzb, (long long)ny * nx, BzT, nz,
                                       a, 1.0, w1, stream);
    if (rc != SGH_OK) return rc;
    rc = sgh_axis_contract_strided(dtype, variant, nzb, ny, nyb, nx, ByT, ldy, w1, 1.0, w2,
                                   stream);
    if (rc != SGH_OK) return rc;
    const LatentParams<F> lp = make_latent<F>(T_S, T_L, rho, L_f, dt, prog);
    const long long rows = (long long)nzb * nyb;
    if (variant == SGH_XFORM_VALU) {
      rc = sgh_axis_contract_strided(dtype, variant, rows, nx, nxb, 1, BxT, ldx, w2, 1.0, Q,
                                     stream);
      if (rc != SGH_OK) return rc;
      rc = sgh_latent_source(dtype, rows * nxb, Q, T_prev, T_S, T_L, rho, L_f, dt, prog, Q,
                             stream);
      if (rc != SGH_OK) return rc;
    } else if (variant == SGH_XFORM_BF16X3) {
      launch_axis_mfma<float, 1>(rows, nx, nxb, 1, (const float*)BxT, ldx, (const float*)w2,
                                 1.0f, (float*)Q, st, (const float*)T_prev,
                                 make_latent<float>(T_S, T_L, rho, L_f, dt, prog));
    } else {
      launch_axis_mfma<F, 0>(rows, nx, nxb, 1, (const F*)BxT, ldx, (const F*)w2, (F)1, (F*)Q,
                             st, (const F*)T_prev, lp);
    }
  });
  SGH_CHECK_LAUNCH("sgh_box_latent_source");
  return SGH_OK;
}

}  // extern "C"
