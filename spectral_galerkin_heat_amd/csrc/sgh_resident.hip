// Resident step kernel of the batched engine (see the block comment below); its host
// launcher is declared in sgh_batch.h.
#include <hip/hip_runtime.h>

#include <cmath>

#include "sgh_batch.h"

namespace sgh {

// ---------------------------------------------------------------------------------
// Resident step kernel: one workgroup per cuboid runs the whole surface part of a step
// (laser flux and shifted warm-start flux, warm start, every Picard iteration with its
// decision, the step-end maximum and the metrics) without leaving the workgroup, so a
// step is this launch plus the volume pass instead of 2·cap + 4 launches, and a cuboid's
// arithmetic depends on nothing but its own data. FP32 state and fields; the 2-D
// transforms are four axis contractions per iteration, each on the matrix cores as three
// bf16 products (hi·hi + hi·lo + lo·hi of x = hi + lo, FP32 accumulation), with the
// cosine matrix (split bf16, read from global memory, L2-resident) as the A operand and
// the field (split bf16, in LDS as [column][k]) as the B operand:
//   inverse  P = B̂yᵀ·Xinv  (rows j, cols m)      Tᵀ = B̂xᵀ·Pᵀ (rows i, cols j)
//   forward  Rᵀ = B̂x·qᵀ    (rows m, cols j)      S = B̂y·R    (rows n, cols m)
// Every extent is padded to a multiple of 32 (the MFMA k step) with zero matrix entries and
// zero field entries. Two LDS operand buffers alternate: a stage reads one and writes its
// result (or the field its epilogue forms) into the other, one 16×16 tile at a time, so a
// stage needs one barrier and a wave holds one tile's accumulator. The per-element fields
// (ZE, Ztt, Wtt, U, q_las) are read from global memory (L2), the next tile's while the
// current one multiplies; G lives in registers for the step. Surfaces up to 128 × 128;
// FP64 and wider surfaces use the per-pass launches.
// ---------------------------------------------------------------------------------
constexpr int kResWaves = kResThreads / kWave;
constexpr int kResTPW = 8;               // most 16×16 output tiles of one wave in a stage
constexpr int kResBuf = kResMaxN * kResLd;   // bf16 elements of one hi or lo operand buffer

typedef __bf16 bf4v __attribute__((ext_vector_type(4)));

// A uniform value unchanged, but opaque to the optimizer: addresses derived from it are
// formed where they are used instead of being hoisted out of the Picard loop and kept live
// (only integers pass through: a laundered pointer would lose its address space).
__device__ __forceinline__ int opaque(int v) {
  v = __builtin_amdgcn_readfirstlane(v);
  asm volatile("" : "+s"(v));
  return v;
}
__device__ __forceinline__ long long opaque(long long v) {
  int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffffLL));
  int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  asm volatile("" : "+s"(lo), "+s"(hi));
  return ((long long)hi << 32) | (unsigned int)lo;
}
// The same for a per-lane value (the lane index every LDS and global address derives from).
__device__ __forceinline__ int opaque_lane(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Output tiles of one wave in a stage of RT × CT 16×16 tiles: the waves are grouped per row
// tile (g waves each), so a wave's tiles share their A rows; wave w takes row tile w / g and
// column tiles (w mod g) + g·q, q < nct. rt < 0: no tiles.
struct ResTiles {
  int rt, ct0, g, nct;
  __device__ ResTiles(int RT, int CT, int wave) {
    g = kResWaves / RT > 0 ? kResWaves / RT : 1;
    rt = wave / g;
    ct0 = wave % g;
    nct = rt < RT && ct0 < CT ? (CT - ct0 + g - 1) / g : 0;
    if (nct == 0) rt = -1;
  }
  __device__ int col(int q, int lane) const { return (ct0 + g * q) * 16 + (lane & 15); }
  __device__ int row(int lane, int v) const { return rt * 16 + 4 * (lane >> 4) + v; }
};

// One operand buffer pair (hi, lo) in LDS: bf16 [column][k] with row stride kResLd.
struct ResBuf {
  __bf16* hi;
  __bf16* lo;
};

// Field values of one lane's 4 output rows (r..r+3 of column c) into an operand buffer,
// split into bf16 hi / lo: transposed ([c][r], the next stage contracts over r) or
// straight ([r][c], it contracts over c).
__device__ __forceinline__ void res_store_t(const ResBuf& b, int r, int c, const float (&x)[4]) {
  bf4v h, l;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const __bf16 hv = (__bf16)x[v];
    h[v] = hv;
    l[v] = (__bf16)(x[v] - (float)hv);
  }
  *reinterpret_cast<bf4v*>(b.hi + c * kResLd + r) = h;
  *reinterpret_cast<bf4v*>(b.lo + c * kResLd + r) = l;
}
__device__ __forceinline__ void res_store_n(const ResBuf& b, int r, int c, const float (&x)[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const __bf16 hv = (__bf16)x[v];
    b.hi[(r + v) * kResLd + c] = hv;
    b.lo[(r + v) * kResLd + c] = (__bf16)(x[v] - (float)hv);
  }
}

// A-operand fragments of one wave for a stage: its row tile's rows, every k step (hi, lo).
struct ResA {
  bf8v h[4], l[4];
};

// Load the A fragments of the wave's row tile in tiles t (no-op without tiles): hi / lo
// row-major matrices at m16 + off_h / off_l in global memory, row stride lda, KS k steps of
// 32 (the offsets, not the pointers, are made opaque, so the loads stay global loads).
__device__ __forceinline__ void res_load_a(ResA& a, const ResTiles& t, const __bf16* m16,
                                           long long off_h, long long off_l, int lda, int KS,
                                           int lane) {
  if (t.rt < 0) return;
  const __bf16* Ah = m16 + opaque(off_h);
  const __bf16* Al = m16 + opaque(off_l);
  lane = opaque_lane(lane);
  const long long arow = (long long)(t.rt * 16 + (lane & 15)) * opaque(lda) + 8 * (lane >> 4);
#pragma unroll
  for (int ks = 0; ks < 4; ++ks)
    if (ks < KS) {
      a.h[ks] = *reinterpret_cast<const bf8v*>(Ah + arow + 32 * ks);
      a.l[ks] = *reinterpret_cast<const bf8v*>(Al + arow + 32 * ks);
      // pinned in registers for the stage: without this the register allocator reloads
      // them from memory for every tile (and waits for each reload)
      asm volatile("" : "+v"(a.h[ks]), "+v"(a.l[ks]));
    }
}

// One stage: for each of the wave's tiles, acc = Σ_k A[r][k]·B[k][c] over K = 32·KS (three
// bf16 products, each in an accumulator of its own), then epi(q, acc, lane).
// A: the fragments in `a`, loaded beforehand; after the last tile `next()` reloads them
// for the following stage, so that their latency passes under the barrier. B: the source buffer. pre(q, lane) starts the loads tile q's epilogue needs;
// it is called two tiles ahead. Ends with the barrier after which the epilogue's stores
// are visible.
template <typename Next, typename Pre, typename Epi>
__device__ __forceinline__ void res_stage(const ResTiles& t, ResA& a, int KS, const ResBuf& src,
                                          int lane, Next&& next, Pre&& pre, Epi&& epi) {
  if (t.rt >= 0) {
    lane = opaque_lane(lane);
    pre(0, lane);
    pre(1, lane);
#pragma unroll
    for (int q = 0; q < kResTPW; ++q) {
      if (q >= t.nct) break;
      pre(q + 2, lane);
      const int boff = t.col(q, lane) * kResLd + 8 * (lane >> 4);
      // the three products accumulate separately, so consecutive MFMAs are independent
      f4v ahh = f4v{0.f, 0.f, 0.f, 0.f}, ahl = ahh, alh = ahh;
      // B fragments one k step ahead of the products that use them
      bf8v bh = *reinterpret_cast<const bf8v*>(src.hi + boff);
      bf8v bl = *reinterpret_cast<const bf8v*>(src.lo + boff);
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        if (ks >= KS) break;
        bf8v nh = bh, nl = bl;
        if (ks + 1 < KS) {
          nh = *reinterpret_cast<const bf8v*>(src.hi + boff + 32 * (ks + 1));
          nl = *reinterpret_cast<const bf8v*>(src.lo + boff + 32 * (ks + 1));
        }
        ahh = mma_bf16(a.h[ks], bh, ahh);
        ahl = mma_bf16(a.h[ks], bl, ahl);
        alh = mma_bf16(a.l[ks], bh, alh);
        bh = nh;
        bl = nl;
      }
      epi(q, ahh + (ahl + alh), lane);
    }
  }
  next();
  __syncthreads();
}

// Sum of a double over the workgroup in a fixed order (wave sums, then the 16 wave
// partials summed by one butterfly in every wave); every thread receives it. sh holds
// kResWaves doubles; ends with a barrier.
__device__ __forceinline__ double res_gather(const double* sh) {
  const int lane = threadIdx.x & (kWave - 1);
  return wave_sum(lane < kResWaves ? sh[lane] : 0.0);
}
__device__ __forceinline__ double res_sum(double v, double* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  const double t = res_gather(sh);
  __syncthreads();
  return t;
}

// Surface offsets of one lane's 4 elements of tile q (rows of the tile along x when
// rows_are_x, else along y); −1 outside the surface or beyond the wave's tiles.
__device__ __forceinline__ void res_offsets(const ResTiles& t, int q, int lane, bool rows_are_x,
                                            int nx, int ny, int (&off)[4]) {
  const int c = t.col(q, lane);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = t.row(lane, v);
    const int x = rows_are_x ? r : c, y = rows_are_x ? c : r;
    off[v] = (q < t.nct && x < nx && y < ny) ? y * nx + x : -1;
  }
}

// Four values of `a` at the offsets; element 0 stands in where off < 0 (the loads are
// unconditional, so the wait counts stay exact; the caller ignores those values).
__device__ __forceinline__ void res_load4(const float* a, const int (&off)[4], float (&o)[4]) {
#pragma unroll
  for (int v = 0; v < 4; ++v) o[v] = a[off[v] >= 0 ? off[v] : 0];
}

__global__ void __launch_bounds__(kResThreads)
k_b_resident(BatchPtrs<float> P, const __bf16* __restrict__ m16, const int* __restrict__ list) {
  using F = float;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh[3][kResWaves];
  const int c = list[blockIdx.x];
  if (c < 0) return;
  const CuboidDesc& d = P.desc[c];
  const int step = P.step_idx[c];
  if (!d.live || step >= d.n_steps) {
    if (threadIdx.x == 0) P.done_at[c] = 0x7fffffff;   // the volume pass leaves it alone
    return;
  }
  __bf16* lds = reinterpret_cast<__bf16*>(smem_raw);
  const ResBuf buf0{lds, lds + kResBuf}, buf1{lds + 2 * kResBuf, lds + 3 * kResBuf};
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int nx = d.nx, ny = d.ny;
  const int nxp = pad32(nx), nyp = pad32(ny);
  const int KX = nxp / 32, KY = nyp / 32;
  // split-bf16 tables: B̂ hi, B̂ lo, B̂ᵀ hi, B̂ᵀ lo, each np × np
  const long long X16 = d.bx16_off, Y16 = d.by16_off;   // element offsets into m16
  const long long mx = (long long)nxp * nxp, my = (long long)nyp * nyp;
  // tp: rows i (x), cols j (y): the second inverse stage and the first forward stage;
  // tm: rows n / j (y), cols m (x modes): the second forward and the first inverse stage
  const ResTiles tp(nxp / 16, nyp / 16, wave);
  const ResTiles tm(nyp / 16, nxp / 16, wave);
  const int so = d.s_off;
  F* const q_las = P.q_las + so;
  F* const q_evap = P.q_evap + so;
  const F* const ZE = P.ZE + so;
  const F* const Ztt = P.Ztt + so;
  const F* const Wtt = P.Wtt + so;
  const F* const U = P.U + so;
  auto no_pre = [](int, int) __attribute__((always_inline)) {};

  // ---- step preparation: AA, laser flux, warm-start flux (physical layout, buffer 0) ----
  double aa = 0.0;
  for (int b = tid; b < d.ncolblk; b += kResThreads) aa += P.col_slots[d.col_item0 + b];
  const double AA = res_sum(aa, sh[0]);
  double psum = 0.0;
  {
    const double* L = P.laser + (d.laser_off + step) * 6;
    const double x0 = L[0], y0 = L[1], power = L[2], on = L[3], vx = L[4], vy = L[5];
    const double I0 = on != 0.0 ? d.absorptivity * power / (d.area_factor * d.rx * d.ry) : 0.0;
    const double shy = vy * d.dt / d.dy, shx = vx * d.dt / d.dx;
    const F* qo = P.q_evap_old + so;
    // four elements per pass over the padded surface (zero operand entries outside it):
    // the previous flux's neighbours of all four are loaded before any is used
    const ResBuf& b0 = buf0;
    const int total = nyp * nxp;
#pragma unroll 1
    for (int e0 = tid; e0 < total; e0 += 4 * kResThreads) {
      float nb[4][4];
      int ij[4][2];
      double tw[4][2];
      bool in_s[4], in_r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + u * kResThreads;
        const int j = e / nxp, i = e - j * nxp;
        ij[u][0] = i;
        ij[u][1] = j;
        in_s[u] = e < total && i < nx && j < ny;
        const double py = j - shy, px = i - shx;
        const double fy = floor(py), fx = floor(px);
        tw[u][0] = py - fy;
        tw[u][1] = px - fx;
        in_r[u] = in_s[u] && py >= 0.0 && py <= ny - 1 && px >= 0.0 && px <= nx - 1;
        const int iy0 = in_r[u] ? (int)fy : 0, ix0 = in_r[u] ? (int)fx : 0;
        const int a0 = min(max(iy0, 0), ny - 1), a1 = min(max(iy0 + 1, 0), ny - 1);
        const int c0 = min(max(ix0, 0), nx - 1), c1 = min(max(ix0 + 1, 0), nx - 1);
        nb[u][0] = qo[a0 * nx + c0];
        nb[u][1] = qo[a0 * nx + c1];
        nb[u][2] = qo[a1 * nx + c0];
        nb[u][3] = qo[a1 * nx + c1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = ij[u][0], j = ij[u][1];
        if (j >= nyp) break;
        float qn = 0.f;
        if (in_s[u]) {
          const double x = (i + 0.5) * d.dx, y = (j + 0.5) * d.dy;
          double ql;
          if (d.cell_integrated) {
            const double s2 = sqrt(2.0), cc = sqrt(M_PI / 8.0);
            const double ax = (x - x0 - d.dx / 2.0) * s2 / d.rx, bx = (x - x0 + d.dx / 2.0) * s2 / d.rx;
            const double ay = (y - y0 - d.dy / 2.0) * s2 / d.ry, by = (y - y0 + d.dy / 2.0) * s2 / d.ry;
            ql = I0 * ((d.ry * cc * (erf(by) - erf(ay)) / d.dy) * (d.rx * cc * (erf(bx) - erf(ax)) / d.dx));
          } else {
            const double ux = (x - x0) / d.rx, uy = (y - y0) / d.ry;
            const double r2 = ux * ux + uy * uy;
            ql = I0 * (d.order == 2.0 ? exp(-2.0 * r2) : exp(-2.0 * pow(r2, d.order / 2.0)));
          }
          const F qlf = (F)ql;
          // bilinear shift of the previous evaporation flux, scipy order-1 'constant'
          const double ty = tw[u][0], tx = tw[u][1];
          const double sv = in_r[u] ? (1.0 - ty) * ((1.0 - tx) * (double)nb[u][0] + tx * (double)nb[u][1]) +
                                          ty * ((1.0 - tx) * (double)nb[u][2] + tx * (double)nb[u][3])
                                    : 0.0;
          q_las[j * nx + i] = qlf;
          qn = (F)((double)qlf - (double)(F)sv);
          psum += (double)qlf;
        }
        const __bf16 hv = (__bf16)qn;
        b0.hi[j * kResLd + i] = hv;               // forward operand [j][i]
        b0.lo[j * kResLd + i] = (__bf16)(qn - (float)hv);
      }
    }
  }
  const double P_sum = res_sum(psum, sh[1]);   // its barriers also complete the operand stores

  const double fwd_scale = d.dx * d.dy * d.inv_sqrt_LxLy;
  const double inv_scale = d.inv_sqrt_LxLy;
  const F fwd_scale_f = (F)fwd_scale, inv_scale_f = (F)inv_scale;
  F g_reg[kResTPW][4];

  // A fragments of the four stage kinds; each stage loads the next one's after its last
  // product
  ResA A;
  auto a_fwd1 = [&]() __attribute__((always_inline)) {
    res_load_a(A, tp, m16, X16, X16 + mx, nxp, KX, lane);
  };
  auto a_fwd2 = [&]() __attribute__((always_inline)) {
    res_load_a(A, tm, m16, Y16, Y16 + my, nyp, KY, lane);
  };
  auto a_inv1 = [&]() __attribute__((always_inline)) {
    res_load_a(A, tm, m16, Y16 + 2 * my, Y16 + 3 * my, nyp, KY, lane);
  };
  auto a_inv2 = [&]() __attribute__((always_inline)) {
    res_load_a(A, tp, m16, X16 + 2 * mx, X16 + 3 * mx, nxp, KX, lane);
  };
  auto a_none = []() __attribute__((always_inline)) {};

  // first stage of each transform: the result goes to the other buffer as the second
  // stage's operand
  auto stage_fwd1 = [&]() __attribute__((always_inline)) {    // Rᵀ = B̂x·qᵀ (rows m, cols j)
    res_stage(tp, A, KX, buf0, lane, a_fwd2, no_pre,
              [&](int q, const f4v& a, int lane) __attribute__((always_inline)) {
                const float o[4] = {a[0], a[1], a[2], a[3]};
                res_store_n(buf1, tp.row(lane, 0), tp.col(q, lane), o);   // [m][j]
              });
  };
  auto stage_inv1 = [&]() __attribute__((always_inline)) {    // P = B̂yᵀ·Xinv (rows j, cols m)
    res_stage(tm, A, KY, buf0, lane, a_inv2, no_pre,
              [&](int q, const f4v& a, int lane) __attribute__((always_inline)) {
                const float o[4] = {a[0], a[1], a[2], a[3]};
                res_store_n(buf1, tm.row(lane, 0), tm.col(q, lane), o);   // [j][m]
              });
  };

  // ---- warm start: G = fwd·B̂y q_net B̂xᵀ; Xinv = ZE + G∘Ztt into buffer 0 ----
  a_fwd1();
  stage_fwd1();
  {
    const int nxk = opaque(nx), nyk = opaque(ny);
    const F* ZEk = P.ZE + opaque(so);
    const F* Zttk = P.Ztt + opaque(so);
    float ze[3][4], zt[3][4];
    res_stage(tm, A, KY, buf1, lane, a_inv1,
              [&](int q, int lane) __attribute__((always_inline)) {
                int off[4];
                res_offsets(tm, q, lane, false, nxk, nyk, off);
                res_load4(ZEk, off, ze[q % 3]);
                res_load4(Zttk, off, zt[q % 3]);
              },
              [&](int q, const f4v& a, int lane) __attribute__((always_inline)) {
                int off[4];
                res_offsets(tm, q, lane, false, nxk, nyk, off);
                float xo[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                  const F g = off[v] >= 0 ? a[v] * fwd_scale_f : 0.f;
                  g_reg[q][v] = g;
                  xo[v] = off[v] >= 0 ? ze[q % 3][v] + g * zt[q % 3][v] : 0.f;
                }
                res_store_t(buf0, tm.row(lane, 0), tm.col(q, lane), xo);   // [m][n]
              });
  }

  // ---- Picard loop ----
  int n_it = 0;
  bool conv = false;
  double rel = __builtin_nan("");
  const F omega = (F)d.omega;
  const F T_liq = (F)d.T_liq, f1 = (F)d.f1, f2 = (F)d.f2, T_boil = (F)d.T_boil;
  const int max_it = d.max_it;
  const double tol = d.tol, n_elem = d.n_elem;
#pragma unroll 1
  for (int k = 0; k < max_it; ++k) {
    // inverse + evaporation: q = q_las − q_evap(T_s) into buffer 0
    stage_inv1();
    {
      const int nxk = opaque(nx), nyk = opaque(ny);
      const F* qlk = P.q_las + opaque(so);
      F* qek = P.q_evap + opaque(so);
      float ql[3][4];
      res_stage(tp, A, KX, buf1, lane, a_fwd1,
                [&](int q, int lane) __attribute__((always_inline)) {
                  int off[4];
                  res_offsets(tp, q, lane, true, nxk, nyk, off);
                  res_load4(qlk, off, ql[q % 3]);
                },
                [&](int q, const f4v& a, int lane) __attribute__((always_inline)) {
                  int off[4];
                  res_offsets(tp, q, lane, true, nxk, nyk, off);
                  float qo[4];
#pragma unroll
                  for (int v = 0; v < 4; ++v) {
                    qo[v] = 0.f;
                    if (off[v] >= 0) {
                      const F T = a[v] * inv_scale_f;
                      F qe = F(0);
                      // hardware reciprocal, reciprocal square root and exponential (about
                      // 1 ulp each): the evaporation law is the FP32 hot spot of the loop
                      if (T >= T_liq) {
                        const F rT = __builtin_amdgcn_rcpf(T);
                        qe = f1 * (__builtin_amdgcn_rsqf(T) * __expf(f2 * (F(1) - T_boil * rT)));
                      }
                      qek[off[v]] = qe;
                      qo[v] = ql[q % 3][v] - qe;
                    }
                  }
                  res_store_t(buf0, tp.row(lane, 0), tp.col(q, lane), qo);   // [j][i]
                });
    }
    // forward + relaxation + norm partials; Xinv into buffer 0
    stage_fwd1();
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    {
      const int nxk = opaque(nx), nyk = opaque(ny);
      const F* ZEk = P.ZE + opaque(so);
      const F* Zttk = P.Ztt + opaque(so);
      const F* Wttk = P.Wtt + opaque(so);
      const F* Uk = P.U + opaque(so);
      float ze[3][4], zt[3][4], wt[3][4], uu[3][4];
      res_stage(tm, A, KY, buf1, lane, a_inv1,
                [&](int q, int lane) __attribute__((always_inline)) {
                  int off[4];
                  res_offsets(tm, q, lane, false, nxk, nyk, off);
                  res_load4(ZEk, off, ze[q % 3]);
                  res_load4(Zttk, off, zt[q % 3]);
                  res_load4(Wttk, off, wt[q % 3]);
                  res_load4(Uk, off, uu[q % 3]);
                },
                [&](int q, const f4v& a, int lane) __attribute__((always_inline)) {
                  int off[4];
                  res_offsets(tm, q, lane, false, nxk, nyk, off);
                  float xo[4];
#pragma unroll
                  for (int v = 0; v < 4; ++v) {
                    xo[v] = 0.f;
                    if (off[v] >= 0) {
                      const F Sr = a[v] * fwd_scale_f;
                      const F g = g_reg[q][v];
                      const F dd = Sr - g;
                      const F w = wt[q % 3][v];
                      // FP32 products, FP64 sums
                      v0 += (double)(dd * dd * w);
                      v1 += (double)(g * uu[q % 3][v]);
                      v2 += (double)(g * g * w);
                      const F gn = g + omega * dd;
                      g_reg[q][v] = gn;
                      xo[v] = ze[q % 3][v] + gn * zt[q % 3][v];
                    }
                  }
                  res_store_t(buf0, tm.row(lane, 0), tm.col(q, lane), xo);   // [m][n]
                });
    }
    // the three sums in one exchange, in a fixed order
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    if (lane == 0) {
      sh[0][wave] = v0;
      sh[1][wave] = v1;
      sh[2][wave] = v2;
    }
    __syncthreads();
    const double rr = res_gather(sh[0]);
    const double gu = res_gather(sh[1]);
    const double g2w = res_gather(sh[2]);
    __syncthreads();
    const double a2 = AA + 2.0 * gu + g2w;
    const double rms_diff = sqrt(rr / n_elem);
    const double rms_old = sqrt(a2 / n_elem);
    rel = rms_diff / (rms_old > 1e-9 ? rms_old : 1e-9);
    conv = rel < tol;
    n_it = k + 1;
    if (conv) break;
  }

  // ---- step end: T_s of the final iterate, its maximum; G for the volume pass ----
  stage_inv1();
  double tmax = -1.0 / 0.0;
  res_stage(tp, A, KX, buf1, lane, a_none, no_pre,
            [&](int q, const f4v& a, int lane) __attribute__((always_inline)) {
              int off[4];
              res_offsets(tp, q, lane, true, nx, ny, off);
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (off[v] >= 0) {
                  const double T = (double)(a[v] * inv_scale_f);
                  tmax = T > tmax ? T : tmax;
                }
            });
#pragma unroll
  for (int q = 0; q < kResTPW; ++q) {
    int off[4];
    res_offsets(tm, q, lane, false, nx, ny, off);
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (off[v] >= 0) P.G[so + off[v]] = g_reg[q][v];
  }
  tmax = wave_max(tmax);
  if (lane == 0) sh[0][wave] = tmax;
  __syncthreads();
  if (tid == 0) {
    double tm_all = -1.0 / 0.0;
    for (int w = 0; w < kResWaves; ++w) tm_all = sh[0][w] > tm_all ? sh[0][w] : tm_all;
    double* st = P.status + kStatusStride * c;
    st[0] = n_it;
    st[1] = conv ? 1.0 : 0.0;
    st[2] = rel;
    st[3] = AA;
    if (step < P.metrics_rows) {
      double* mrow = P.metrics + ((long long)c * P.metrics_rows + step) * kMetricStride;
      mrow[0] = P_sum * d.dx * d.dy;
      mrow[1] = tm_all;
      mrow[2] = n_it;
      mrow[3] = conv ? 1.0 : 0.0;
      mrow[4] = rel;
    }
    P.done_at[c] = n_it;
    P.step_idx[c] = step + 1;
  }
}

hipError_t launch_resident(const BatchPtrs<float>& P, const void* m16, const int* list,
                           unsigned nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_b_resident, dim3(nblocks), dim3(kResThreads), kResLdsBytes, st, P,
                     (const __bf16*)m16, list);
  return hipGetLastError();
}

hipError_t resident_allow_lds() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_b_resident),
                             hipFuncAttributeMaxDynamicSharedMemorySize, kResLdsBytes);
}

}  // namespace sgh
