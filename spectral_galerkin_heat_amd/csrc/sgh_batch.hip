// Batched surface-collapsed ETD1 + Picard engine for many independent cuboids.
//
// For a cuboid whose forcing is a top-surface flux only (laser + evaporation; constant
// properties, no latent heat, no convection) every Picard iterate of the reference's
// step (sgh/solvers/spectral.py:320-465) has the form
//     a = aE + KK ∘ (w ⊗ G),   aE = a_n·E,  w[p] = C_p(−1)^p,  G a surface field,
// so the whole iteration reduces to surface-sized work:
//     T_s      = Bᵧᵀ (ZE + G∘Ztt) Bₓ             (ZE = Σ_p w·aE,  Ztt = Σ_p w²·KK)
//     S_raw    = dx·dy · Bᵧ (q_las − q_evap(T_s)) Bₓᵀ
//     D = S_raw − G,   G ← G + ω·D
//     r·r      = Σ D²·Wtt                          (Wtt = Σ_p w²·KK²)
//     aₒ·aₒ    = AA + 2 Σ G·U + Σ G²·Wtt          (AA = Σ aE², U = Σ_p aE·KK·w)
// and one volume pass per step writes a_{n+1} = aE + KK∘(w⊗G) and already gathers
// ZE, U, AA of the next step. E and KK are regenerated in FP64 from 1-D tables.
//
// Work decomposition: every launch covers all cuboids through a packed work list of
// (cuboid, 16-row block) items for surface transforms and (cuboid, 256-column block)
// items for the volume pass. Partial sums go to fixed per-item slots and are summed in
// a fixed order by the next kernel, so a cuboid's arithmetic does not depend on what
// else is in the batch, and no block ever waits on another inside a kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "sgh_batch.h"

static_assert(sizeof(sgh_cuboid_spec) == 176, "sgh_cuboid_spec layout (mirrored in _lib.py)");

namespace sgh {

__device__ __forceinline__ double dexp_(double x) { return exp(x); }

// ---------------------------------------------------------------------------------
// setup: Ztt, Wtt (constant per cuboid)
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double w_top(int p, const CuboidDesc& d) {
  const double c = p == 0 ? d.w0 : d.w1;
  return (p & 1) ? -c : c;
}

// KK = (1 − E)/(αλ)/(ρc), with 1/λ from v_rcp_f64 refined by two Newton steps.
__device__ __forceinline__ double kk_of(double e, double lam, const CuboidDesc& d) {
  if (lam == 0.0) return d.dt / d.rho_cp;
  double r = __builtin_amdgcn_rcp(lam);
  r = fma(r, fma(-lam, r, 1.0), r);
  r = fma(r, fma(-lam, r, 1.0), r);
  return (1.0 - e) * d.inv_arc * r;
}

template <typename F>
__global__ void __launch_bounds__(kBT) k_b_setup(BatchPtrs<F> P) {
  const Item it = P.col_items[blockIdx.x];
  if (it.c < 0) return;
  const CuboidDesc d = P.desc[it.c];
  const int s = it.blk * kBT + threadIdx.x;
  if (s >= d.S) return;
  const int n = s / d.nx, m = s - n * d.nx;
  const double* ex = P.t1 + d.t1_off;
  const double* kx2 = ex + d.nx;
  const double* ey = kx2 + d.nx;
  const double* ky2 = ey + d.ny;
  const double* ez = ky2 + d.ny;
  const double* kz2 = ez + d.nz;
  const double eyx = ey[n] * ex[m];
  const double kyx = ky2[n] + kx2[m];
  double z = 0.0, w2 = 0.0;
  for (int p = 0; p < d.nz; ++p) {
    const double e = ez[p] * eyx;
    const double kk = kk_of(e, kz2[p] + kyx, d);
    const double w = w_top(p, d);
    z += w * w * kk;
    w2 += w * w * kk * kk;
  }
  P.Ztt[d.s_off + s] = (F)z;
  P.Wtt[d.s_off + s] = (F)w2;
  if (P.zwu) {
    P.zwu[4 * (d.s_off + s) + 1] = (float)z;
    P.zwu[4 * (d.s_off + s) + 2] = (float)w2;
  }
}

// ---------------------------------------------------------------------------------
// volume pass: finalize step n (a ← aE + KK·w·G) and gather ZE, U, AA of step n+1
// ---------------------------------------------------------------------------------
// mode 0: gather only (priming a new cuboid); 1: finalize the step of a cuboid whose step
// counter has not advanced yet (per-pass launches); 2: finalize the cuboids the resident
// step kernel ran in this step (done_at < INT_MAX; it has advanced the counter itself).
template <typename F>
__global__ void __launch_bounds__(kBT) k_b_volume(BatchPtrs<F> P, int mode) {
  constexpr int PC = 8;  // z-planes loaded ahead per column
  __shared__ double sh[kBT / kWave];
  const Item it = P.col_items[blockIdx.x];
  if (it.c < 0) return;
  const CuboidDesc d = P.desc[it.c];
  const int s = it.blk * kBT + threadIdx.x;
  const int finalize = mode != 0;
  const bool active_cuboid = mode == 0 ? true
                             : mode == 1 ? P.step_idx[it.c] < d.n_steps
                                         : P.done_at[it.c] != 0x7fffffff;
  double aa = 0.0;
  if (s < d.S && active_cuboid) {
    const int n = s / d.nx, m = s - n * d.nx;
    const double* ex = P.t1 + d.t1_off;
    const double* kx2 = ex + d.nx;
    const double* ey = kx2 + d.nx;
    const double* ky2 = ey + d.ny;
    const double* ez = ky2 + d.ny;
    const double* kz2 = ez + d.nz;
    const double eyx = ey[n] * ex[m];
    const double kyx = ky2[n] + kx2[m];
    const double g = finalize ? (double)P.G[d.s_off + s] : 0.0;
    F* __restrict__ a = P.a + d.a_off + s;
    const long long S = d.S;
    double ze = 0.0, u = 0.0;
    for (int p0 = 0; p0 < d.nz; p0 += PC) {
      F buf[PC];
#pragma unroll
      for (int q = 0; q < PC; ++q)
        buf[q] = p0 + q < d.nz ? a[(long long)(p0 + q) * S] : F(0);
#pragma unroll
      for (int q = 0; q < PC; ++q) {
        const int p = p0 + q;
        if (p >= d.nz) break;
        const double e = ez[p] * eyx;
        const double kk = kk_of(e, kz2[p] + kyx, d);
        const double w = w_top(p, d);
        double av = (double)buf[q];
        if (finalize) {
          const F nv = (F)(av * e + kk * (w * g));
          buf[q] = nv;
          av = (double)nv;
        }
        const double ae = av * e;
        ze += w * ae;
        u += ae * kk * w;
        aa += ae * ae;
      }
      if (finalize) {
#pragma unroll
        for (int q = 0; q < PC; ++q)
          if (p0 + q < d.nz) a[(long long)(p0 + q) * S] = buf[q];
      }
    }
    P.ZE[d.s_off + s] = (F)ze;
    P.U[d.s_off + s] = (F)u;
    if (P.zwu) {
      P.zwu[4 * (d.s_off + s)] = (float)ze;
      P.zwu[4 * (d.s_off + s) + 3] = (float)u;
    }
    if (finalize) P.q_evap_old[d.s_off + s] = P.q_evap[d.s_off + s];
  }
  const double t = block_sum(aa, sh);
  if (threadIdx.x == 0) {
    P.col_slots[d.col_item0 + it.blk] = t;
    // after a resident step: the cuboid's team barrier counter back to zero for the next
    if (mode == 2 && it.blk == 0) P.bar[it.c] = 0u;
  }
}

// ---------------------------------------------------------------------------------
// step preparation: laser flux, shifted warm-start flux, per-step resets
// ---------------------------------------------------------------------------------
template <typename F>
__global__ void __launch_bounds__(kBT) k_b_prep(BatchPtrs<F> P) {
  __shared__ double sh[kBT / kWave];
  const Item it = P.row_items[blockIdx.x];
  if (it.c < 0) return;
  const CuboidDesc d = P.desc[it.c];
  const int step = P.step_idx[it.c];
  const bool live = step < d.n_steps;
  if (it.blk == 0 && threadIdx.x == 0) {
    P.done_at[it.c] = live ? -1 : 0x7fffffff;
    P.bar[it.c] = 0u;
  }
  double psum = 0.0;
  if (live) {
    const double* L = P.laser + (d.laser_off + step) * 6;
    const double x0 = L[0], y0 = L[1], power = L[2], on = L[3], vx = L[4], vy = L[5];
    const double I0 = on != 0.0 ? d.absorptivity * power / (d.area_factor * d.rx * d.ry) : 0.0;
    const double sy = vy * d.dt / d.dy, sx = vx * d.dt / d.dx;
    const int r0 = (it.blk / d.ncb) * d.item_rows;
    const int rows = min(d.item_rows, d.ny - r0);
    const int c_lo = (it.blk % d.ncb) * cb_width(d);
    const int cw = min(cb_width(d), d.nx - c_lo);
    const F* qo = P.q_evap_old + d.s_off;
    for (int e = threadIdx.x; e < rows * cw; e += kBT) {
      const int j = r0 + e / cw, i = c_lo + e % cw;
      const int s = j * d.nx + i;
      const double x = (i + 0.5) * d.dx, y = (j + 0.5) * d.dy;
      double q;
      if (d.cell_integrated) {
        const double s2 = sqrt(2.0), cc = sqrt(M_PI / 8.0);
        const double ax = (x - x0 - d.dx / 2.0) * s2 / d.rx, bx = (x - x0 + d.dx / 2.0) * s2 / d.rx;
        const double ay = (y - y0 - d.dy / 2.0) * s2 / d.ry, by = (y - y0 + d.dy / 2.0) * s2 / d.ry;
        q = I0 * ((d.ry * cc * (erf(by) - erf(ay)) / d.dy) * (d.rx * cc * (erf(bx) - erf(ax)) / d.dx));
      } else {
        const double ux = (x - x0) / d.rx, uy = (y - y0) / d.ry;
        const double r2 = ux * ux + uy * uy;
        q = I0 * (d.order == 2.0 ? exp(-2.0 * r2) : exp(-2.0 * pow(r2, d.order / 2.0)));
      }
      const F ql = (F)q;
      // bilinear shift of the previous evaporation flux, scipy order-1 'constant'
      const double py = j - sy, px = i - sx;
      const double fy = floor(py), fx = floor(px);
      const double ty = py - fy, tx = px - fx;
      const int iy0 = (int)fy, ix0 = (int)fx;
      const bool inr = py >= 0.0 && py <= d.ny - 1 && px >= 0.0 && px <= d.nx - 1;
      double sh_v = 0.0;
      if (inr) {
        const int a0 = min(max(iy0, 0), d.ny - 1), b0 = min(max(iy0 + 1, 0), d.ny - 1);
        const int c0 = min(max(ix0, 0), d.nx - 1), d0 = min(max(ix0 + 1, 0), d.nx - 1);
        sh_v = (1.0 - ty) * ((1.0 - tx) * (double)qo[a0 * d.nx + c0] + tx * (double)qo[a0 * d.nx + d0]) +
               ty * ((1.0 - tx) * (double)qo[b0 * d.nx + c0] + tx * (double)qo[b0 * d.nx + d0]);
      }
      P.q_las[d.s_off + s] = ql;
      P.q_net[d.s_off + s] = (F)((double)ql - (double)(F)sh_v);
      psum += (double)ql;
    }
  }
  const double t = block_sum(psum, sh);
  if (threadIdx.x == 0) P.aux_slots[2 * (d.row_item0 + it.blk)] = t;
  if (it.blk == 0) {
    // AA = Σ aE² of this step (per-column-item partials from the last volume pass),
    // used by every Picard decision of the step
    double aa = 0.0;
    for (int b = threadIdx.x; b < d.ncolblk; b += kBT) aa += P.col_slots[d.col_item0 + b];
    aa = block_sum(aa, sh);
    if (threadIdx.x == 0) P.status[kStatusStride * it.c + 3] = aa;
  }
}

// Picard decision for cuboid c before iteration k (k = 0: always run), evaluated by the
// 64 lanes of wave 0 together: lane b loads row item b's norm partials and the wave sums
// them in a fixed butterfly order; AA = Σ aE² is constant over the step and was summed
// once by k_b_prep into status[3]. Returns true (on every lane) while the cuboid iterates.
template <typename F, bool COH = false>
__device__ __forceinline__ bool picard_active(const BatchPtrs<F>& P, const CuboidDesc& d, int c,
                                              int k, bool record) {
  const int lane = threadIdx.x & (kWave - 1);
  // all loads first (independent of each other), so they are in flight together
  const int done_at = P.done_at[c];
  double rr = 0.0, gu = 0.0, g2w = 0.0, aa0 = 0.0;
  if (k > 0) {
    if (lane < n_items(d)) {
      const double* sl = P.row_slots + 3 * (d.row_item0 + lane);
      rr = ld_x<double, COH>(sl);
      gu = ld_x<double, COH>(sl + 1);
      g2w = ld_x<double, COH>(sl + 2);
    }
    aa0 = P.status[kStatusStride * c + 3];
  }
  if (done_at >= 0) return false;
  bool done = false, conv = false;
  double rel = __builtin_nan("");
  if (k > 0) {
    rr = wave_sum(rr);
    gu = wave_sum(gu);
    g2w = wave_sum(g2w);
    const double aa = aa0 + 2.0 * gu + g2w;
    const double rms_diff = sqrt(rr / d.n_elem);
    const double rms_old = sqrt(aa / d.n_elem);
    rel = rms_diff / (rms_old > 1e-9 ? rms_old : 1e-9);
    conv = rel < d.tol;
    done = conv;
  }
  if (k >= d.max_it) done = true;
  if (done && record && lane == 0) {
    P.done_at[c] = k;
    double* st = P.status + kStatusStride * c;
    st[0] = k;
    st[1] = conv ? 1.0 : 0.0;
    st[2] = rel;
  }
  return !done;
}

// ---------------------------------------------------------------------------------
// surface transforms (one kernel, four roles), 16 output rows per item:
//   MODE 0 (warm start): G = dx·dy·Bᵧ q_net Bₓᵀ / √(LxLy)
//   MODE 1 (iteration, forward half): D = S_raw − G, norm slots, G += ω D
//   MODE 2 (iteration, inverse half): T_s = Bᵧᵀ(ZE + G∘Ztt)Bₓ/√(LxLy) → q_evap, q_net
//   MODE 3 (step end): T_s of the final iterate → T_max slot
// Stage 1, Y = L[rows,:]·X, streams X from L2 straight into MFMA B operands (each
// element is read by exactly one wave of the item); Y goes through LDS; stage 2,
// out = Y·R, streams R the same way. Column tiles of 16 are dealt to the 4 waves.
// ---------------------------------------------------------------------------------
#ifdef SGH_BLOCK_TIMESTAMPS
// Diagnostic build only: per-block wall-clock stamps of the launches with k == 0 (every
// cuboid active), slots 0 entry, 1 staging issued, 2 staging landed, 3 stage 1 done,
// 4 stage 2 begins, 5 end.
__device__ unsigned long long g_block_ts[4 * 8192 * 6];
#define SGH_TS(slot)                                                                  \
  do {                                                                                \
    if (threadIdx.x == 0 && blockIdx.x < 8192 && k == 0)                              \
      g_block_ts[(MODE * 8192 + blockIdx.x) * 6 + (slot)] = wall_clock64();           \
  } while (0)
#else
#define SGH_TS(slot) do {} while (0)
#endif

// acc[m] += Σ_{k < kn} A_m[k] · B[k] on the matrix cores for the lane's 16-column tile,
// k in steps of 4 (kn a multiple of 4). ga(m, k): this lane's A element of row tile m
// at k + (lane / 16); gb(k): its B element at k + (lane / 16).
template <typename F, int KMT, int KS, typename GA, typename GB>
__device__ __forceinline__ void mma_k(typename Mfma<F>::V (&acc)[KMT], int mt, int kn, GA ga,
                                      GB gb) {
  static_assert(KS == 1, "one accumulator chain");
#pragma unroll 4
  for (int k = 0; k < kn; k += 4) {
    const F bv = gb(k);
#pragma unroll
    for (int m = 0; m < KMT; ++m)
      if (m < mt) acc[m] = Mfma<F>::mma(ga(m, k), bv, acc[m]);
  }
}

template <typename F, int MODE, int RB, int NT, bool CS>
__global__ void __launch_bounds__(NT) k_b_surface(BatchPtrs<F> P, int k, int bud, int lds_elems) {
  constexpr bool LOOP = false;
#include "sgh_surface_pass.inc"
}

// Per-cuboid heights (variants with rows 0): each workgroup runs its item with the body
// built for the item's own height, 16 or 32 rows (the 32-row body on a 16-row item idles
// half its row tiles and, under 1024 threads, has no registers to prefetch the forward
// pass's operands), so that all items fit one round of CUs without slowing the 16-row ones.
template <typename F, int MODE, int NT, bool CS>
__global__ void __launch_bounds__(NT) k_b_surface_hts(BatchPtrs<F> P, int k, int bud,
                                                      int lds_elems) {
  constexpr bool LOOP = false;
  if (P.row_descs[blockIdx.x].d.item_rows > 16) {
    constexpr int RB = 32;
#include "sgh_surface_pass.inc"
  } else {
    constexpr int RB = 16;
#include "sgh_surface_pass.inc"
  }
}

// Persistent form of one step's surface work (one launch instead of 2·cap + 2): each
// workgroup runs its item through the warm start, every Picard iteration (inverse pass,
// cuboid barrier, forward pass, cuboid barrier) and the step-end pass, taking the Picard
// decisions between iterations from the same slots and in the same order as the per-pass
// launches, so the iterates are identical. It needs every item resident at once: the host
// launches it only when the work list fits one workgroup per CU (the interleaved list
// gives XCD x the workgroups b with b mod 8 = x), and P.bar[c] = 0 at the start of the step
// (k_b_prep). The passes are the per-pass kernel's body included in place, each in a scope
// of its own with MODE and k set, so that they read the kernel arguments directly (as
// calls taking the argument block by reference they made the compiler copy it to scratch
// memory and reload pointers from there in every pass). Measured no faster than the
// per-pass launches on MI355X (graph-launched kernel boundaries cost about what a
// cross-workgroup barrier costs), so it is opt-in (sgh_batch_set_loop) and the
// autotuner times it beside them.
template <typename F, int RB, int NT, bool CS>
__global__ void __launch_bounds__(NT) k_b_loop(BatchPtrs<F> P, int bud, int lds_elems,
                                               int max_it, unsigned long long ticks) {
  constexpr bool LOOP = true;
  __shared__ int s_go;
  const ItemDesc& item = P.row_descs[blockIdx.x];
  const int cub = item.c, blk0 = item.blk;
  if (cub < 0) return;
  const CuboidDesc dc = item.d;
  if (P.step_idx[cub] >= dc.n_steps) return;   // finished cuboid: every pass would skip
  const unsigned n_it = (unsigned)n_items(dc);
  unsigned arrivals = 0;
  auto barrier = [&]() {
    arrivals += n_it;
    cuboid_barrier(P.bar + cub, arrivals, P.loop_err, ticks);
  };
  {
    constexpr int MODE = 0;
    const int k = 0;
#include "sgh_surface_pass.inc"
  }
  barrier();
  int kit = 0;
  for (; kit < max_it; ++kit) {
    if (threadIdx.x < kWave) {
      const bool act = picard_active<F, true>(P, dc, cub, kit, blk0 == 0);
      if (threadIdx.x == 0) s_go = act ? 1 : 0;
    }
    __syncthreads();
    const bool go = __builtin_amdgcn_readfirstlane(s_go) != 0;
    __syncthreads();
    if (!go) break;
    {
      constexpr int MODE = 2;
      const int k = kit;
#include "sgh_surface_pass.inc"
    }
    barrier();
    {
      constexpr int MODE = 1;
      const int k = kit;
#include "sgh_surface_pass.inc"
    }
    barrier();
  }
  // ran to the iteration cap: the decision that the per-pass form takes in its step-end
  // launch records the final status
  if (kit == max_it && threadIdx.x < kWave)
    picard_active<F, true>(P, dc, cub, max_it, blk0 == 0);
  {
    constexpr int MODE = 3;
    const int k = max_it;
#include "sgh_surface_pass.inc"
  }
}

// Per-step metrics and step-counter advance (single block).
template <typename F>
__global__ void __launch_bounds__(kBT) k_b_metrics(BatchPtrs<F> P, int skip_resident) {
  for (int c = threadIdx.x; c < P.n_cuboids; c += kBT) {
    const CuboidDesc d = P.desc[c];
    if (!d.live) continue;
    // a cuboid the resident step kernel ran has written its row and advanced its counter
    if (skip_resident && d.bx16_off >= 0 && d.by16_off >= 0) continue;
    const int step = P.step_idx[c];
    if (step >= d.n_steps) continue;
    double pl = 0.0, tm = -1.0 / 0.0;
    for (int b = 0; b < n_items(d); ++b) {
      pl += P.aux_slots[2 * (d.row_item0 + b)];
      const double t = P.aux_slots[2 * (d.row_item0 + b) + 1];
      tm = t > tm ? t : tm;
    }
    if (step < P.metrics_rows) {
      double* m = P.metrics + ((long long)c * P.metrics_rows + step) * kMetricStride;
      const double* st = P.status + kStatusStride * c;
      m[0] = pl * d.dx * d.dy;
      m[1] = tm;
      m[2] = st[0];
      m[3] = st[1];
      m[4] = st[2];
    }
    P.step_idx[c] = step + 1;
  }
}

// ---------------------------------------------------------------------------------
// Plan: a pool of cuboid slots over fixed-capacity packed arrays. Cuboids are added and
// removed between steps (continuous batching); work lists and the captured step graph
// are rebuilt when membership changes.
// ---------------------------------------------------------------------------------
constexpr int kMaxRowBlk = kBMaxN / kRowsMin;      // row blocks per cuboid (ny <= 256)
constexpr int kColBlkW = 64;                         // column-block width of wide surfaces
constexpr int kMaxItems = 64;                        // transform items per cuboid (one wave
                                                     // sums their slots in a decision)
static_assert(kMaxRowBlk * ((kBMaxN + kColBlkW - 1) / kColBlkW) <= kMaxItems, "item slots");

// Column blocks of a cuboid's transform items: surfaces wider than 128 are split into
// 64-column blocks (a single large cuboid then spreads over more workgroups); narrower
// ones keep whole rows.
static int column_blocks(int nx) { return nx > 128 ? (nx + kColBlkW - 1) / kColBlkW : 1; }
constexpr int kMaxColBlk = kBMaxN * kBMaxN / kBT;  // column blocks per cuboid (S <= 65536)

static inline long long rup64(long long v) { return (v + 63) / 64 * 64; }

// First-fit allocator over [0, capacity) with coalescing free list.
struct Arena {
  long long cap = 0;
  std::vector<std::pair<long long, long long>> free_;  // (offset, size), sorted
  void init(long long c) { cap = c; free_.assign(1, {0, c}); }
  long long alloc(long long n, long long align) {
    for (size_t i = 0; i < free_.size(); ++i) {
      const long long o = (free_[i].first + align - 1) / align * align;
      const long long end = free_[i].first + free_[i].second;
      if (o + n <= end) {
        const long long head = o - free_[i].first, tail = end - (o + n);
        std::vector<std::pair<long long, long long>> repl;
        if (head > 0) repl.push_back({free_[i].first, head});
        if (tail > 0) repl.push_back({o + n, tail});
        free_.erase(free_.begin() + i);
        free_.insert(free_.begin() + i, repl.begin(), repl.end());
        return o;
      }
    }
    return -1;
  }
  void release(long long o, long long n) {
    free_.push_back({o, n});
    std::sort(free_.begin(), free_.end());
    std::vector<std::pair<long long, long long>> m;
    for (auto& f : free_) {
      if (!m.empty() && m.back().first + m.back().second >= f.first)
        m.back().second = std::max(m.back().second, f.first + f.second - m.back().first);
      else
        m.push_back(f);
    }
    free_ = m;
  }
};

struct SlotAlloc {
  long long a_off = 0, a_len = 0, s_off = 0, s_len = 0, l_off = 0, l_len = 0;
};

struct Plan {
  int dtype = 0;
  int n = 0;         // slot capacity
  int n_live = 0;
  int max_it = 0;
  size_t elem = 8;
  std::vector<CuboidDesc> hdesc;
  std::vector<SlotAlloc> halloc;
  std::vector<Item> hrow, hcol;
  Arena arena_a, arena_s, arena_l;
  // matrix table: one (B̂, B̂ᵀ) pair per distinct mode count in use, reference-counted by
  // the live cuboids' x and y axes and released when the last one leaves
  std::vector<int> sizes;
  std::vector<long long> size_off;
  std::vector<int> size_refs;
  Arena arena_m;
  int max_nz = 0;    // deepest cuboid the 1-D table slots hold (t1_stride)
  // split-bf16 tables of the resident step kernel (mode counts up to kResMaxN), one block
  // of four pad32(n)² matrices per entry of `sizes` (−1: none)
  std::vector<long long> size_off16;
  Arena arena_m16;
  void* mats16 = nullptr;
  void* zwu = nullptr;       // FP32: (ZE, Ztt, Wtt, U) per surface element, 16 bytes
  void* res_work = nullptr;  // FP32: kResWork floats per cuboid slot (resident kernel)
  void* res_xch = nullptr;   // FP32: kResXchBytes per cuboid slot (resident teams)
  long long mats16_cap = 0;
  bool resident = false;     // variant rows = −1 requested
  bool res_ok = false;       // ... and it applies (mixed precision, 31-bit surface offsets)
  bool res_on = false;       // ... and a live cuboid fits it (both surface sides ≤ 128)
  std::vector<int> hres;     // resident work list: res_entry per workgroup (−1 idle)
  std::vector<int> hteam;    // resident team size per slot (0: none) of the last lists
  int team_max = kResTeamMax;  // largest resident team
  int* res_list = nullptr;
  long long res_cap = 0;
  int t1_stride = 0;
  size_t shm = 0;
  int bud = 0;
  int rows = 16;       // surface rows per transform item; 0: per cuboid (16 or 32), see assign_heights
  int threads = 1024;  // threads per surface workgroup (the autotuner's usual pick)
  long long total_N = 0, total_S = 0;
  // device allocations
  void* a = nullptr;
  void* surf = nullptr;  // kSurfArrays arrays of total_S
  void* mats = nullptr;
  long long mats_cap = 0;
  double* t1 = nullptr;
  double* laser = nullptr;
  CuboidDesc* desc = nullptr;
  Item* row_items = nullptr;
  ItemDesc* row_descs = nullptr;
  Item* col_items = nullptr;
  Item* prime_items = nullptr;
  long long row_cap = 0, col_cap = 0;
  double* row_slots = nullptr;
  double* col_slots = nullptr;
  double* aux_slots = nullptr;
  int* done_at = nullptr;
  double* status = nullptr;
  int* step_idx = nullptr;
  double* metrics = nullptr;
  int metrics_rows = 0;
  std::vector<int> pending_prime;  // slots added since the last run
  bool dirty = true;               // work lists / graph out of date
  bool colsplit = false;           // a live cuboid's items are split by columns
  bool loop_pref = false;          // run each step's surface work as one persistent launch
  bool loop_on = false;            // ... and the work list allows it (all items resident)
  unsigned long long barrier_ticks = kBarrierTicks;  // bound of a barrier wait (100 MHz)
  int n_cu = 0;                    // compute units of the device
  unsigned* bar = nullptr;         // [n] barrier counters + 1 error word
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t graph_stream = nullptr;
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double t_surface_ms = 0.0, t_volume_ms = 0.0;
  long long n_surface = 0, n_volume = 0;
  // a step holding both resident and per-pass cuboids runs the resident launch on a side
  // stream beside the per-pass launches (created with the first such work list)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

template <typename F>
BatchPtrs<F> ptrs(Plan* p) {
  BatchPtrs<F> P;
  F* s = (F*)p->surf;
  const long long T = p->total_S;
  P.a = (F*)p->a;
  P.G = s + 0 * T;
  P.ZE = s + 1 * T;
  P.U = s + 2 * T;
  P.Ztt = s + 3 * T;
  P.Wtt = s + 4 * T;
  P.q_las = s + 5 * T;
  P.q_net = s + 6 * T;
  P.q_evap = s + 7 * T;
  P.q_evap_old = s + 8 * T;
  P.Xinv = s + 9 * T;
  P.surf_stride = T;
  P.zwu = (float*)p->zwu;
  P.res_work = (float*)p->res_work;
  P.res_xch = (unsigned char*)p->res_xch;
  P.mats = (const F*)p->mats;
  P.t1 = p->t1;
  P.laser = p->laser;
  P.desc = p->desc;
  P.row_items = p->row_items;
  P.row_descs = p->row_descs;
  P.col_items = p->col_items;
  P.row_slots = p->row_slots;
  P.col_slots = p->col_slots;
  P.aux_slots = p->aux_slots;
  P.done_at = p->done_at;
  P.status = p->status;
  P.step_idx = p->step_idx;
  P.metrics = p->metrics;
  P.metrics_rows = p->metrics_rows;
  P.bar = p->bar;
  P.loop_err = reinterpret_cast<int*>(p->bar + p->n);
  P.n_cuboids = p->n;
  return P;
}

// Surface-transform variants: (rows per item, threads per workgroup). Each instantiates
// the four modes of k_b_surface; the autotuner picks one per batch.
struct SurfVariant {
  int rows, threads;
};
// rows = 0: each cuboid's items hold 16 or 32 rows (assign_heights), on the 32-row kernels.
// rows = −1: the resident step kernel (a team of 512-thread workgroups per cuboid), mixed
// precision and surfaces up to 128 × 128; wider ones on 16-row per-pass items beside it.
constexpr int kResidentRows = -1;
constexpr SurfVariant kSurfVariants[] = {{16, 256}, {16, 512}, {16, 1024}, {32, 512},
                                         {32, 1024}, {0, 512}, {0, 1024},
                                         {kResidentRows, kResThreads}};
constexpr int kNumSurfVariants = sizeof(kSurfVariants) / sizeof(kSurfVariants[0]);

template <typename F>
using SurfFn = void (*)(BatchPtrs<F>, int, int, int);

template <typename F, int RB, int NT, bool CS>
static void surf_fns(SurfFn<F>* f) {
  f[0] = &k_b_surface<F, 0, RB, NT, CS>;
  f[1] = &k_b_surface<F, 1, RB, NT, CS>;
  f[2] = &k_b_surface<F, 2, RB, NT, CS>;
  f[3] = &k_b_surface<F, 3, RB, NT, CS>;
}

template <typename F, int NT, bool CS>
static void surf_fns_hts(SurfFn<F>* f) {
  f[0] = &k_b_surface_hts<F, 0, NT, CS>;
  f[1] = &k_b_surface_hts<F, 1, NT, CS>;
  f[2] = &k_b_surface_hts<F, 2, NT, CS>;
  f[3] = &k_b_surface_hts<F, 3, NT, CS>;
}

template <typename F, bool CS>
static bool surf_lookup_cs(int rows, int threads, SurfFn<F>* f) {
  // per-cuboid heights: one kernel holding the 16- and the 32-row bodies
  if (rows == 0 && threads == 512) surf_fns_hts<F, 512, CS>(f);
  else if (rows == 0 && threads == 1024) surf_fns_hts<F, 1024, CS>(f);
  else if (rows == 16 && threads == 256) surf_fns<F, 16, 256, CS>(f);
  else if (rows == 16 && threads == 512) surf_fns<F, 16, 512, CS>(f);
  else if (rows == 16 && threads == 1024) surf_fns<F, 16, 1024, CS>(f);
  else if (rows == 32 && threads == 512) surf_fns<F, 32, 512, CS>(f);
  else if (rows == 32 && threads == 1024) surf_fns<F, 32, 1024, CS>(f);
  else return false;
  return true;
}

// The kernels of a variant; `colsplit` selects those that handle column-split items
// (needed as soon as one live cuboid is wider than 128).
template <typename F>
static bool surf_lookup(int rows, int threads, bool colsplit, SurfFn<F>* f) {
  return colsplit ? surf_lookup_cs<F, true>(rows, threads, f)
                  : surf_lookup_cs<F, false>(rows, threads, f);
}

template <typename F>
using LoopFn = void (*)(BatchPtrs<F>, int, int, int, unsigned long long);

// Persistent step kernels of the fixed 16-row variants with 512 or 1024 threads.
template <typename F>
static LoopFn<F> loop_lookup(int rows, int threads, bool colsplit) {
  if (rows == 16 && threads == 512)
    return colsplit ? &k_b_loop<F, 16, 512, true> : &k_b_loop<F, 16, 512, false>;
  if (rows == 16 && threads == 1024)
    return colsplit ? &k_b_loop<F, 16, 1024, true> : &k_b_loop<F, 16, 1024, false>;
  return nullptr;
}

template <typename F>
static hipError_t launch_surface(const SurfFn<F>* f, int mode, unsigned nrow, const Plan* p,
                                 BatchPtrs<F>& P, int k, hipStream_t st) {
  int bud = p->bud;
  int lds_elems = (int)(p->shm / p->elem);
  void* args[] = {&P, &k, &bud, &lds_elems};
  return hipLaunchKernel(reinterpret_cast<const void*>(f[mode]), dim3(nrow), dim3(p->threads),
                         args, p->shm, st);
}

// LDS of a surface workgroup: Ls[rows][ldl] (reused for Ys[rows][ldy]) + a staging buffer
// of `bud` elements that holds X (or R) whole when it fits, else in chunks of at least 4
// rows (X) or 16 columns (R).
// Rows of the kernel tiles for a plan row setting (0: per-cuboid heights, up to 32).
static inline int kernel_rows(int rows) { return rows == 0 ? 32 : rows; }

static int surface_budget(int dtype, int rows, long long lds_bytes) {
  const int n4 = (kBMaxN + 3) & ~3;
  const int ld = dtype == SGH_F64 ? pad_a<double>(n4) : pad_a<float>(n4);
  const long long elems = lds_bytes / (dtype == SGH_F64 ? 8 : 4);
  const long long bud = elems - (long long)rows * ld;   // L and Y share their rows
  const long long need1 = 4LL * (dtype == SGH_F64 ? pad_b<double>(kBMaxN) : pad_b<float>(kBMaxN));
  const long long need2 = (long long)n4 * (dtype == SGH_F64 ? pad_b<double>(16) : pad_b<float>(16));
  return (bud < need1 || bud < need2) ? -1 : (int)bud;
}

template <typename F>
static hipError_t allow_large_lds(size_t bytes) {
  hipError_t e = hipSuccess;
  for (const SurfVariant& v : kSurfVariants) {
    for (int cs = 0; cs < 2; ++cs) {
      SurfFn<F> f[5] = {};
      if (!surf_lookup<F>(v.rows, v.threads, cs == 1, f)) continue;
      f[4] = reinterpret_cast<SurfFn<F>>(loop_lookup<F>(v.rows, v.threads, cs == 1));
      for (int m = 0; m < 5; ++m) {
        if (!f[m]) continue;
        const int cap = m == 4 ? std::min((int)bytes, kLoopLdsMax) : (int)bytes;
        hipError_t r = hipFuncSetAttribute(reinterpret_cast<const void*>(f[m]),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, cap);
        if (e == hipSuccess) e = r;
      }
    }
  }
  return e;
}

template <typename F>
int enqueue_step(Plan* p, hipStream_t st) {
  BatchPtrs<F> P = ptrs<F>(p);
  const unsigned nrow = (unsigned)p->hrow.size(), ncol = (unsigned)p->hcol.size();
  const bool res = p->res_on;
  if (nrow == 0 && !res) return SGH_OK;
  const bool tm = p->timing;
  SurfFn<F> f[4];
  if (!surf_lookup<F>(p->rows, p->threads, p->colsplit, f)) {
    set_error("sgh_batch: no surface variant rows=%d threads=%d", p->rows, p->threads);
    return SGH_ERR_STATE;
  }
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
  // timed surface work: with the resident kernel every surface launch of the step, else
  // the Picard iterations' launches (or the persistent launch)
  if (tm && res) hipEventRecord(p->ev[0], st);
  // with per-pass cuboids in the step too (and not timing kernels), the resident launch
  // runs on the side stream: its workgroups and the per-pass ones share the CUs, and each
  // cuboid's kernels (so its results) are those of a step without the others
  const bool fork = res && nrow > 0 && !tm && p->side;
  if (fork) {
    chk(hipEventRecord(p->ev_fork, st));
    chk(hipStreamWaitEvent(p->side, p->ev_fork, 0));
  }
  if constexpr (std::is_same<F, float>::value) {
    // the whole surface part of the step of every cuboid within the resident kernel's
    // tables in one launch (each advances its own step counter and writes its metrics row)
    if (res)
      chk(launch_resident(P, p->mats16, p->res_list, (unsigned)p->hres.size(),
                          p->barrier_ticks, fork ? p->side : st));
  }
  if (fork) chk(hipEventRecord(p->ev_join, p->side));
  if (nrow > 0) hipLaunchKernelGGL(k_b_prep<F>, dim3(nrow), dim3(kBT), 0, st, P);
  if (nrow == 0) {
    // resident cuboids only
  } else if (p->loop_on) {
    // the whole step's surface work in one launch (timed as a whole)
    LoopFn<F> lf = loop_lookup<F>(p->rows, p->threads, p->colsplit);
    int bud = p->bud, lds_elems = (int)(p->shm / p->elem), max_it = p->max_it;
    unsigned long long ticks = p->barrier_ticks;
    void* args[] = {&P, &bud, &lds_elems, &max_it, &ticks};
    if (tm && !res) hipEventRecord(p->ev[0], st);
    chk(hipLaunchKernel(reinterpret_cast<const void*>(lf), dim3(nrow), dim3(p->threads), args,
                        p->shm, st));
    if (tm && !res) hipEventRecord(p->ev[1], st);
  } else {
    chk(launch_surface<F>(f, 0, nrow, p, P, 0, st));
    if (tm && !res) hipEventRecord(p->ev[0], st);
    for (int k = 0; k < p->max_it; ++k) {
      chk(launch_surface<F>(f, 2, nrow, p, P, k, st));
      chk(launch_surface<F>(f, 1, nrow, p, P, k, st));
    }
    if (tm && !res) hipEventRecord(p->ev[1], st);
    chk(launch_surface<F>(f, 3, nrow, p, P, p->max_it, st));
  }
  if (fork) chk(hipStreamWaitEvent(st, p->ev_join, 0));
  if (tm && res) hipEventRecord(p->ev[1], st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch surface launch");
  // mode 2 finalizes every cuboid that stepped (done_at < INT_MAX): the per-pass ones, which
  // k_b_prep marked, and the resident ones; mode 1 needs the step counters not yet advanced
  if (tm) hipEventRecord(p->ev[2], st);
  hipLaunchKernelGGL(k_b_volume<F>, dim3(ncol), dim3(kBT), 0, st, P, res ? 2 : 1);
  if (tm) hipEventRecord(p->ev[3], st);
  if (nrow > 0) hipLaunchKernelGGL(k_b_metrics<F>, dim3(1), dim3(kBT), 0, st, P, res ? 1 : 0);
  SGH_CHECK_LAUNCH("sgh_batch step");
  if (tm) {
    hipError_t e = hipEventSynchronize(p->ev[3]);
    if (e != hipSuccess) return fail_hip(e, "sgh_batch timing");
    float a = 0.f, b = 0.f;
    hipEventElapsedTime(&a, p->ev[0], p->ev[1]);
    hipEventElapsedTime(&b, p->ev[2], p->ev[3]);
    p->t_surface_ms += a;
    p->n_surface += res ? 1LL + (nrow == 0 ? 0LL : 2LL * p->max_it + 2LL)
                        : p->loop_on ? 1LL : 2LL * p->max_it;
    p->t_volume_ms += b;
    p->n_volume += 1;
  }
  return SGH_OK;
}

// Ztt, Wtt and the first step's ZE, U, AA for newly added cuboids only.
template <typename F>
int prime_new(Plan* p, const std::vector<Item>& items, hipStream_t st) {
  if (items.empty()) return SGH_OK;
  hipError_t e = hipMemcpyAsync(p->prime_items, items.data(), items.size() * sizeof(Item),
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch prime upload");
  BatchPtrs<F> P = ptrs<F>(p);
  P.col_items = p->prime_items;
  hipLaunchKernelGGL(k_b_setup<F>, dim3((unsigned)items.size()), dim3(kBT), 0, st, P);
  hipLaunchKernelGGL(k_b_volume<F>, dim3((unsigned)items.size()), dim3(kBT), 0, st, P, 0);
  SGH_CHECK_LAUNCH("sgh_batch prime");
  return SGH_OK;
}

// B̂[k, i] = (k == 0 ? 1 : √2)·cos(πk(2i+1)/(2n)); the basis of a cuboid of extent L is
// B̂/√L, so one matrix per distinct mode count serves every cuboid.
static std::vector<double> unit_basis(int n) {
  std::vector<double> B((size_t)n * n);
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) {
      const long long num = ((long long)k * (2 * i + 1)) % (4LL * n);
      B[(size_t)k * n + i] = (k == 0 ? 1.0 : sqrt(2.0)) * cos(M_PI * (double)num / (2.0 * n));
    }
  return B;
}

// XCD-aware work lists: each live cuboid is pinned to one of the 8 XCDs (greedy balance of
// its transform work) and its items sit at list positions ≡ xcd (mod 8), so all
// workgroups of a cuboid share one L2. Short bins are padded with empty items (c = −1).
// Per-cuboid item heights for rows = 0: start every cuboid at 16 rows and move the
// cheapest ones (narrowest surfaces: the least MFMA work per row) to 32 rows until the
// surface items fit one workgroup per CU (kXcds × kCusPerXcd), so a launch runs in a
// single round. Returns true if any descriptor changed.
static bool assign_heights(Plan* p) {
  std::vector<int> live;
  for (int c = 0; c < p->n; ++c)
    if (p->hdesc[c].live) live.push_back(c);
  std::vector<int> h(p->n, 16);
  long long items = 0;
  for (int c : live) items += (long long)p->hdesc[c].ncb * ((p->hdesc[c].ny + 15) / 16);
  std::vector<int> order = live;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    const CuboidDesc &da = p->hdesc[a], &db = p->hdesc[b];
    return da.nx != db.nx ? da.nx < db.nx : a < b;
  });
  const long long target = (long long)kXcds * (kCusPerXcd - 1);  // one CU per XCD of slack for binning
  for (int c : order) {
    if (items <= target) break;
    const int ny = p->hdesc[c].ny;
    items -= (long long)p->hdesc[c].ncb * ((ny + 15) / 16 - (ny + 31) / 32);
    h[c] = 32;
  }
  bool changed = false;
  for (int c : live) {
    CuboidDesc& d = p->hdesc[c];
    if (d.item_rows != h[c]) {
      d.item_rows = h[c];
      d.nrowblk = (d.ny + h[c] - 1) / h[c];
      changed = true;
    }
  }
  return changed;
}

// A live cuboid's step runs on the resident step kernel when the resident variant applies
// and its split-bf16 tables exist (both surface sides at most kResMaxN); per-pass launches
// run the others in the same step. k_b_metrics applies the same rule on the device.
static bool on_resident(const Plan* p, const CuboidDesc& d) {
  return p->res_ok && d.live && d.bx16_off >= 0 && d.by16_off >= 0;
}

// Work lists: per-pass row items of the cuboids the resident kernel does not run, volume
// column items of every live cuboid, and (teams of one) the resident list; each cuboid's
// items go to the least loaded XCD (list position ≡ XCD mod 8).
static void build_xcd_items(Plan* p) {
  std::vector<double> load(kXcds, 0.0), rload(kXcds, 0.0);
  std::vector<int> count(kXcds, 0), rcount(kXcds, 0);
  std::vector<std::vector<Item>> rb(kXcds), cb(kXcds), res(kXcds);
  std::vector<int> order;
  for (int c = 0; c < p->n; ++c)
    if (p->hdesc[c].live) order.push_back(c);
  auto cost = [&](int c) {
    const CuboidDesc& d = p->hdesc[c];
    return (double)d.nrowblk * d.nx * (d.nx + d.ny) + 0.05 * (double)d.S * d.nz;
  };
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    const double ca = cost(a), cb_ = cost(b);
    return ca != cb_ ? ca > cb_ : a < b;
  });
  p->hteam.assign(p->n, 0);
  p->max_it = 0;
  for (int c : order) {
    // least loaded XCD that still has a CU for each of the cuboid's workgroups (its items,
    // or its one resident workgroup); else the one with the fewest
    const CuboidDesc& d = p->hdesc[c];
    const bool r = on_resident(p, d);
    std::vector<double>& ld = r ? rload : load;
    std::vector<int>& ct = r ? rcount : count;
    const int w = r ? 1 : n_items(d);
    int x = -1;
    for (int y = 0; y < kXcds; ++y)
      if (ct[y] + w <= kCusPerXcd && (x < 0 || ld[y] < ld[x])) x = y;
    if (x < 0) x = (int)(std::min_element(ct.begin(), ct.end()) - ct.begin());
    ld[x] += cost(c);
    ct[x] += w;
    for (int b = 0; b < d.ncolblk; ++b) cb[x].push_back({c, b});
    if (r) {
      res[x].push_back({c, 0});
      p->hteam[c] = 1;
    } else {
      for (int b = 0; b < n_items(d); ++b) rb[x].push_back({c, b});
      p->max_it = std::max(p->max_it, d.max_it);
    }
  }
  auto interleave = [](std::vector<std::vector<Item>>& bins, std::vector<Item>& out) {
    size_t m = 0;
    for (auto& b : bins) m = std::max(m, b.size());
    out.clear();
    for (size_t r = 0; r < m; ++r)
      for (int x = 0; x < kXcds; ++x) out.push_back(r < bins[x].size() ? bins[x][r] : Item{-1, 0});
  };
  interleave(rb, p->hrow);
  interleave(cb, p->hcol);
  std::vector<Item> rl;
  interleave(res, rl);
  p->hres.clear();
  for (const Item& it : rl) p->hres.push_back(it.c < 0 ? -1 : res_entry(it.c, 1, 0));
  p->res_on = !p->hres.empty();
}

// Resident teams (sgh_resident.hip): how many workgroups each cuboid's step runs on. The
// cost of a member's Picard iteration is modelled in thousands of cycles, fitted to
// per-workgroup phase stamps of the kernel on MI355X: alone, 6 + 3·(KX + KY) +
// 0.25·(TA·KY + TB·KX), with KX, KY the k steps of the x and y contractions and TA, TB the
// tiles per wave of the two stage kinds; a member of a team of tp saves 0.88·K per tile
// per wave it no longer holds (K the mean k steps) and pays 10.5 for its two field exchanges
// (each the epilogue's device-coherent stores, a barrier and a copy: about 2 µs). Starting from one
// workgroup per cuboid, the cuboid whose step is the longest moves to the smallest larger
// team that shortens it, while the batch still fits one workgroup per CU (members wait for
// each other; CUs for the per-pass items of the step are kept free) and the teams still fit
// the XCDs (a team's members run on one XCD). The work
// list then places each team on one XCD (list positions ≡ xcd mod 8), largest teams first.
static double team_cost(const CuboidDesc& d, int tp) {
  const int ntb = pad32(d.nx) / 16, nay = pad32(d.ny) / 16;
  const int kx = pad32(d.nx) / 32, ky = pad32(d.ny) / 32;
  const int ga = std::max(1, kWavesPerRes / nay), gb = std::max(1, kWavesPerRes / ntb);
  const int ta = (ntb + ga - 1) / ga, tb = (nay + gb - 1) / gb;
  const double alone = 6.0 + 3.0 * (kx + ky) + 0.25 * (ta * ky + tb * kx);
  if (tp <= 1) return alone;
  const int rows = (nay + tp - 1) / tp;
  const int t1 = (ntb * nay + kWavesPerRes - 1) / kWavesPerRes;
  const int tt = (ntb * rows + kWavesPerRes - 1) / kWavesPerRes;
  return alone - 0.88 * 0.5 * (kx + ky) * (t1 - tt) + 10.5;
}

static bool place_teams(Plan* p, const std::vector<int>& order) {
  std::vector<int> by(order);
  std::stable_sort(by.begin(), by.end(), [&](int a, int b) {
    const int ta = p->hteam[a], tb = p->hteam[b];
    if (ta != tb) return ta > tb;
    return team_cost(p->hdesc[a], ta) > team_cost(p->hdesc[b], tb);
  });
  std::vector<int> count(kXcds, 0);
  std::vector<double> load(kXcds, 0.0);
  std::vector<std::vector<int>> bins(kXcds);
  for (int c : by) {
    const int tp = p->hteam[c];
    int x = -1;
    for (int y = 0; y < kXcds; ++y)
      if (count[y] + tp <= kCusPerXcd && (x < 0 || load[y] < load[x])) x = y;
    if (x < 0) return false;
    count[x] += tp;
    load[x] += tp * team_cost(p->hdesc[c], tp);
    for (int m = 0; m < tp; ++m) bins[x].push_back(res_entry(c, tp, m));
  }
  size_t rows = 0;
  for (auto& b : bins) rows = std::max(rows, b.size());
  p->hres.clear();
  for (size_t r = 0; r < rows; ++r)
    for (int x = 0; x < kXcds; ++x) p->hres.push_back(r < bins[x].size() ? bins[x][r] : -1);
  return true;
}

static void assign_teams(Plan* p) {
  std::vector<int> order;
  for (int c = 0; c < p->n; ++c)
    if (p->hteam[c] > 0) order.push_back(c);
  const long long hw = std::min<long long>(p->n_cu, (long long)kXcds * kCusPerXcd);
  if (p->team_max <= 1 || order.empty() || (long long)order.size() > hw) return;
  // per-pass items of the same step (run beside the resident launch) keep a CU each
  long long items = 0;
  for (const Item& it : p->hrow) items += it.c >= 0 ? 1 : 0;
  const long long slots = std::max<long long>((long long)order.size(), hw - items);
  long long total = (long long)order.size();
  for (;;) {
    int worst = -1;
    double wc = -1.0;
    for (int c : order) {
      const double v = team_cost(p->hdesc[c], p->hteam[c]);
      if (v > wc) { wc = v; worst = c; }
    }
    // the smallest larger team that shortens its step
    const CuboidDesc& d = p->hdesc[worst];
    const int tp = p->hteam[worst], nay = pad32(d.ny) / 16, lim = std::min(p->team_max, nay);
    int nt = tp + 1;
    while (nt <= lim && team_cost(d, nt) >= wc) ++nt;
    if (nt > lim || total + (nt - tp) > slots) break;
    p->hteam[worst] = nt;
    total += nt - tp;
  }
  // a team assignment that does not pack into the XCDs gives up members, largest teams
  // first
  while (!place_teams(p, order)) {
    int big = -1;
    for (int c : order)
      if (big < 0 || p->hteam[c] > p->hteam[big]) big = c;
    if (p->hteam[big] <= 1) break;
    --p->hteam[big];
  }
}

static void drop_graph(Plan* p) {
  if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
  if (p->graph) hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
}

static void free_plan(Plan* p) {
  if (!p) return;
  drop_graph(p);
  for (auto& e : p->ev)
    if (e) hipEventDestroy(e);
  if (p->ev_fork) hipEventDestroy(p->ev_fork);
  if (p->ev_join) hipEventDestroy(p->ev_join);
  if (p->side) hipStreamDestroy(p->side);
  void* bufs[] = {p->a, p->surf, p->mats, p->t1, p->laser, p->desc, p->row_items, p->row_descs,
                  p->col_items,
                  p->prime_items, p->row_slots, p->col_slots, p->aux_slots, p->done_at,
                  p->status, p->step_idx, p->metrics, p->bar, p->mats16, p->res_list,
                  p->zwu, p->res_work, p->res_xch};
  for (void* b : bufs)
    if (b) hipFree(b);
  delete p;
}

template <typename T>
static hipError_t dmalloc(T** ptr, size_t bytes) {
  return hipMalloc((void**)ptr, bytes < 16 ? 16 : bytes);
}

// float -> bf16 bits, round to nearest even (finite values), and back.
static inline uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static inline float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Offset of the (B̂, B̂ᵀ) pair for mode count m, uploading it if no live cuboid uses m yet;
// every successful call takes one reference (dropped by mats_release).
static long long mats_offset(Plan* p, int m, int* rc) {
  for (size_t i = 0; i < p->sizes.size(); ++i)
    if (p->sizes[i] == m) {
      p->size_refs[i] += 1;
      return p->size_off[i];
    }
  const long long mm = rup64((long long)m * m);  // B̂ and B̂ᵀ start 256-byte aligned
  const long long need = 2LL * mm;
  const long long off = p->arena_m.alloc(need, 64);
  if (off < 0) {
    *rc = SGH_ERR_NOMEM;
    set_error("sgh_batch: matrix table full (%d distinct mode counts in use, %lld elements)",
              (int)p->sizes.size(), p->mats_cap);
    return -1;
  }
  std::vector<double> B = unit_basis(m);
  std::vector<unsigned char> buf((size_t)need * p->elem, 0);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      const double v = B[(size_t)r * m + c], vt = B[(size_t)c * m + r];
      const size_t i0 = (size_t)r * m + c, i1 = (size_t)mm + i0;
      if (p->elem == 8) {
        reinterpret_cast<double*>(buf.data())[i0] = v;
        reinterpret_cast<double*>(buf.data())[i1] = vt;
      } else {
        reinterpret_cast<float*>(buf.data())[i0] = (float)v;
        reinterpret_cast<float*>(buf.data())[i1] = (float)vt;
      }
    }
  // synchronous: the host buffer goes out of scope
  hipError_t e = hipMemcpy((char*)p->mats + off * p->elem, buf.data(), buf.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    p->arena_m.release(off, need);
    *rc = fail_hip(e, "sgh_batch matrix upload");
    return -1;
  }
  long long off16 = -1;
  if (m <= kResMaxN) {
    const int np = pad32(m);
    const long long mm16 = (long long)np * np;
    off16 = p->arena_m16.alloc(4 * mm16, 64);
    if (off16 >= 0) {
      std::vector<uint16_t> t((size_t)(4 * mm16), 0);
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) {
          const double v = B[(size_t)r * m + c];
          const uint16_t hi = bf16_bits((float)v);
          const uint16_t lo = bf16_bits((float)(v - (double)bf16_value(hi)));
          t[(size_t)r * np + c] = hi;                  // B̂
          t[(size_t)(mm16 + r * np + c)] = lo;
          t[(size_t)(2 * mm16 + c * np + r)] = hi;     // B̂ᵀ
          t[(size_t)(3 * mm16 + c * np + r)] = lo;
        }
      e = hipMemcpy((uint16_t*)p->mats16 + off16, t.data(), t.size() * 2, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        p->arena_m16.release(off16, 4 * mm16);
        p->arena_m.release(off, need);
        *rc = fail_hip(e, "sgh_batch bf16 matrix upload");
        return -1;
      }
    }
  }
  p->sizes.push_back(m);
  p->size_off.push_back(off);
  p->size_off16.push_back(off16);
  p->size_refs.push_back(1);
  return off;
}

// Offset of the split-bf16 tables of mode count m (−1: none); m must be in the table.
static long long mats16_offset(const Plan* p, int m) {
  for (size_t i = 0; i < p->sizes.size(); ++i)
    if (p->sizes[i] == m) return p->size_off16[i];
  return -1;
}

// Drop one reference to mode count m's matrices; the last one frees their table room.
static void mats_release(Plan* p, int m) {
  for (size_t i = 0; i < p->sizes.size(); ++i)
    if (p->sizes[i] == m) {
      if (--p->size_refs[i] > 0) return;
      p->arena_m.release(p->size_off[i], 2LL * rup64((long long)m * m));
      if (p->size_off16[i] >= 0)
        p->arena_m16.release(p->size_off16[i], 4LL * pad32(m) * pad32(m));
      p->sizes.erase(p->sizes.begin() + i);
      p->size_off.erase(p->size_off.begin() + i);
      p->size_off16.erase(p->size_off16.begin() + i);
      p->size_refs.erase(p->size_refs.begin() + i);
      return;
    }
}

static bool spec_ok(const sgh_cuboid_spec& sp) {
  return !(sp.nx < 2 || sp.ny < 1 || sp.nz < 1 || sp.nx > kBMaxN || sp.ny > kBMaxN ||
           sp.nz > 4096 || sp.Lx <= 0 || sp.Ly <= 0 || sp.Lz <= 0 || sp.k <= 0 || sp.rho <= 0 ||
           sp.cp <= 0 || sp.dt <= 0 || sp.rx <= 0 || sp.ry <= 0 || sp.R_v <= 0 ||
           sp.T_boil <= 0 || sp.n_steps < 0 || sp.max_it < 0 || sp.order <= 0 ||
           (sp.n_steps > 0 && sp.laser == nullptr) || (sp.cell_integrated && sp.order != 2.0));
}

static long long spec_state_elems(const sgh_cuboid_spec& sp) {
  return (long long)sp.nx * sp.ny * sp.nz;
}

// Fill slot c with a cuboid (host tables copied synchronously; the caller's laser table
// may be freed on return).
static int add_cuboid(Plan* p, const sgh_cuboid_spec& sp, int c, hipStream_t st) {
  CuboidDesc d{};
  SlotAlloc al;
  d.nx = sp.nx;
  d.ny = sp.ny;
  d.nz = sp.nz;
  d.S = sp.nx * sp.ny;
  const long long N = (long long)d.S * d.nz;
  al.a_len = N;
  al.s_len = d.S;
  al.l_len = std::max(1, sp.n_steps);
  al.a_off = p->arena_a.alloc(al.a_len, 64);
  al.s_off = al.a_off < 0 ? -1 : p->arena_s.alloc(al.s_len, 64);
  al.l_off = al.s_off < 0 ? -1 : p->arena_l.alloc(al.l_len, 1);
  if (al.a_off < 0 || al.s_off < 0 || al.l_off < 0) {
    if (al.a_off >= 0) p->arena_a.release(al.a_off, al.a_len);
    if (al.s_off >= 0) p->arena_s.release(al.s_off, al.s_len);
    set_error("sgh_batch_add: pool capacity exhausted");
    return SGH_ERR_NOMEM;
  }
  int rc = SGH_OK;
  const long long bx = mats_offset(p, sp.nx, &rc);
  const long long by = rc ? -1 : mats_offset(p, sp.ny, &rc);
  if (rc) {
    if (bx >= 0) mats_release(p, sp.nx);
    p->arena_a.release(al.a_off, al.a_len);
    p->arena_s.release(al.s_off, al.s_len);
    p->arena_l.release(al.l_off, al.l_len);
    return rc;
  }
  d.a_off = al.a_off;
  d.s_off = (int)al.s_off;
  d.bx_off = (int)bx;
  d.bxT_off = (int)(bx + rup64((long long)sp.nx * sp.nx));
  d.by_off = (int)by;
  d.byT_off = (int)(by + rup64((long long)sp.ny * sp.ny));
  d.bx16_off = (int)mats16_offset(p, sp.nx);
  d.by16_off = (int)mats16_offset(p, sp.ny);
  d.t1_off = c * p->t1_stride;
  const double alpha = sp.k / (sp.rho * sp.cp);
  std::vector<double> t1;
  const double Ls[3] = {sp.Lx, sp.Ly, sp.Lz};
  const int ns[3] = {sp.nx, sp.ny, sp.nz};
  for (int ax = 0; ax < 3; ++ax) {
    std::vector<double> e(ns[ax]), k2(ns[ax]);
    for (int m = 0; m < ns[ax]; ++m) {
      const double kk = M_PI * m / Ls[ax];
      k2[m] = kk * kk;
      e[m] = exp(-alpha * k2[m] * sp.dt);
    }
    t1.insert(t1.end(), e.begin(), e.end());
    t1.insert(t1.end(), k2.begin(), k2.end());
  }
  d.row_item0 = c * kMaxItems;
  d.item_rows = p->rows == 0 ? 16 : p->rows;  // per-cuboid heights are set when lists are built
  d.nrowblk = (sp.ny + d.item_rows - 1) / d.item_rows;
  d.ncb = column_blocks(sp.nx);
  d.col_item0 = c * kMaxColBlk;
  d.ncolblk = (d.S + kBT - 1) / kBT;
  d.max_it = sp.max_it;
  d.n_steps = sp.n_steps;
  d.laser_off = al.l_off;
  d.dx = sp.Lx / sp.nx;
  d.dy = sp.Ly / sp.ny;
  d.inv_sqrt_LxLy = 1.0 / sqrt(sp.Lx * sp.Ly);
  d.Lz = sp.Lz;
  d.alpha = alpha;
  d.rho_cp = sp.rho * sp.cp;
  d.w0 = sqrt(1.0 / sp.Lz);
  d.w1 = sqrt(2.0 / sp.Lz);
  d.inv_arc = 1.0 / (alpha * d.rho_cp);
  d.dt = sp.dt;
  d.omega = sp.omega;
  d.tol = sp.tol;
  d.n_elem = (double)N;
  d.f1 = 0.82 * sp.dH * sp.Pa / sqrt(2.0 * M_PI * sp.R_v);
  d.f2 = sp.dH / (sp.R_v * sp.T_boil);
  d.T_boil = sp.T_boil;
  d.T_liq = sp.T_liq;
  d.absorptivity = sp.absorptivity;
  d.rx = sp.rx;
  d.ry = sp.ry;
  d.order = sp.order;
  d.area_factor = M_PI * pow(2.0, 1.0 - 2.0 / sp.order) * tgamma(2.0 / sp.order) / sp.order;
  d.cell_integrated = sp.cell_integrated;
  d.live = 1;

  const int zero = 0;
  const double zstat[kStatusStride] = {0, 0, 0, 0};
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
  if ((long long)t1.size() > p->t1_stride) {   // callers check nz <= max_nz first
    mats_release(p, sp.nx);
    mats_release(p, sp.ny);
    p->arena_a.release(al.a_off, al.a_len);
    p->arena_s.release(al.s_off, al.s_len);
    p->arena_l.release(al.l_off, al.l_len);
    set_error("sgh_batch_add: 1-D tables of %d doubles exceed the slot's %d", (int)t1.size(),
              p->t1_stride);
    return SGH_ERR_ARG;
  }
  chk(hipMemcpy(p->t1 + d.t1_off, t1.data(), t1.size() * sizeof(double), hipMemcpyHostToDevice));
  if (sp.n_steps > 0)
    chk(hipMemcpy(p->laser + 6 * al.l_off, sp.laser, 6 * sizeof(double) * sp.n_steps,
                  hipMemcpyHostToDevice));
  chk(hipMemcpy(p->desc + c, &d, sizeof(CuboidDesc), hipMemcpyHostToDevice));
  chk(hipMemcpy(p->step_idx + c, &zero, sizeof(int), hipMemcpyHostToDevice));
  chk(hipMemcpy(p->done_at + c, &zero, sizeof(int), hipMemcpyHostToDevice));
  chk(hipMemcpy(p->status + kStatusStride * c, zstat, sizeof(zstat), hipMemcpyHostToDevice));
  chk(hipMemsetAsync((char*)p->a + al.a_off * p->elem, 0, al.a_len * p->elem, st));
  for (int arr = 0; arr < kSurfArrays; ++arr)
    chk(hipMemsetAsync((char*)p->surf + (arr * p->total_S + al.s_off) * p->elem, 0,
                       al.s_len * p->elem, st));
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_add upload");
  p->hdesc[c] = d;
  p->halloc[c] = al;
  p->n_live += 1;
  p->pending_prime.push_back(c);
  p->dirty = true;
  return SGH_OK;
}

static int rebuild_lists(Plan* p, hipStream_t st) {
  hipError_t e = hipSuccess;
  if (p->rows == 0 && assign_heights(p)) {
    for (int c = 0; c < p->n && e == hipSuccess; ++c)
      if (p->hdesc[c].live)
        e = hipMemcpy(p->desc + c, &p->hdesc[c], sizeof(CuboidDesc), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "sgh_batch heights");
  }
  // the resident step kernel: mixed precision, the surface arrays addressable with 31-bit
  // byte offsets; it runs the live cuboids within its tables, per-pass launches the rest
  p->res_ok = p->resident && p->dtype == SGH_F32 &&
              (long long)kSurfArrays * p->total_S * 4 < 0x7fffffffLL;
  build_xcd_items(p);
  p->colsplit = false;
  for (int c = 0; c < p->n; ++c)
    if (p->hdesc[c].live && p->hdesc[c].ncb > 1) p->colsplit = true;
  // the persistent step kernel needs every workgroup resident: one per CU (its LDS), and
  // the interleaved list gives XCD x the workgroups b with b mod 8 = x
  const bool has_loop = p->dtype == SGH_F64
                            ? loop_lookup<double>(p->rows, p->threads, p->colsplit) != nullptr
                            : loop_lookup<float>(p->rows, p->threads, p->colsplit) != nullptr;
  p->loop_on = p->loop_pref && has_loop && !p->res_on &&
               (long long)p->hrow.size() <= p->n_cu && p->shm <= (size_t)kLoopLdsMax;
  if (p->res_on) assign_teams(p);
  if (p->res_on && !p->hrow.empty() && !p->side) {
    e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) return fail_hip(e, "sgh_batch side stream");
  }
  if ((long long)p->hres.size() > p->res_cap) {
    set_error("sgh_batch: resident work list exceeds capacity");
    return SGH_ERR_STATE;
  }
  if ((long long)p->hrow.size() > p->row_cap || (long long)p->hcol.size() > p->col_cap) {
    set_error("sgh_batch: work list exceeds capacity");
    return SGH_ERR_STATE;
  }
  std::vector<ItemDesc> hrd(p->hrow.size());
  for (size_t i = 0; i < p->hrow.size(); ++i) {
    const Item& it = p->hrow[i];
    hrd[i].c = it.c;
    hrd[i].blk = it.blk;
    hrd[i].d = it.c >= 0 ? p->hdesc[it.c] : CuboidDesc{};
  }
  if (!p->hrow.empty())
    e = hipMemcpyAsync(p->row_items, p->hrow.data(), p->hrow.size() * sizeof(Item),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !hrd.empty())
    e = hipMemcpyAsync(p->row_descs, hrd.data(), hrd.size() * sizeof(ItemDesc),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !p->hcol.empty())
    e = hipMemcpyAsync(p->col_items, p->hcol.data(), p->hcol.size() * sizeof(Item),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !p->hres.empty())
    e = hipMemcpyAsync(p->res_list, p->hres.data(), p->hres.size() * sizeof(int),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch work lists");
  drop_graph(p);
  p->dirty = false;
  return SGH_OK;
}

}  // namespace sgh

using namespace sgh;

extern "C" {

int sgh_batch_pool_create(int dtype, int max_cuboids, long long state_capacity,
                          long long surface_capacity, long long laser_capacity,
                          int max_nz, int metrics_rows, void** plan_out) {
  SGH_REQUIRE(plan_out && max_cuboids > 0 && state_capacity > 0 && surface_capacity > 0 &&
                  laser_capacity > 0 && max_nz > 0 && max_nz <= 4096 && metrics_rows >= 0,
              SGH_ERR_ARG, "sgh_batch_pool_create: bad args");
  SGH_REQUIRE(dtype == SGH_F64 || dtype == SGH_F32, SGH_ERR_ARG, "sgh_batch_pool_create: dtype");
  *plan_out = nullptr;
  Plan* p = new Plan();
  p->dtype = dtype;
  p->n = max_cuboids;
  p->elem = dtype == SGH_F64 ? 8 : 4;
  p->metrics_rows = metrics_rows;
  p->hdesc.assign(max_cuboids, CuboidDesc{});
  p->halloc.assign(max_cuboids, SlotAlloc{});
  p->arena_a.init(state_capacity);
  p->arena_s.init(surface_capacity);
  p->arena_l.init(laser_capacity);
  p->total_N = state_capacity;
  p->total_S = rup64(surface_capacity);  // each packed surface array starts 256-byte aligned
  p->max_nz = max_nz;
  p->t1_stride = 2 * (kBMaxN + kBMaxN + max_nz);
  p->mats_cap = 2LL * 16 * kBMaxN * kBMaxN;   // room for 16 distinct 256-wide pairs
  p->arena_m.init(p->mats_cap);
  p->mats16_cap = 4LL * 16 * kResMaxN * kResMaxN;   // 16 distinct sizes up to 128
  p->arena_m16.init(p->mats16_cap);
  p->res_cap = (long long)kXcds * std::max(max_cuboids, kCusPerXcd);
  p->bud = surface_budget(dtype, kernel_rows(p->rows), kSurfLds);
  p->shm = (size_t)kSurfLds;
  if (p->bud < 0) {
    free_plan(p);
    set_error("sgh_batch_pool_create: LDS staging buffer too small");
    return SGH_ERR_ARG;
  }
  p->row_cap = (long long)max_cuboids * kMaxItems + 8LL * kMaxItems * kXcds;
  p->col_cap = (long long)max_cuboids * kMaxColBlk + 8LL * kMaxColBlk * kXcds;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
  const size_t E = p->elem;
  chk(dtype == SGH_F64 ? allow_large_lds<double>(kLdsMax) : allow_large_lds<float>(kLdsMax));
  chk(dmalloc(&p->a, state_capacity * E));
  chk(dmalloc(&p->surf, kSurfArrays * p->total_S * E));
  chk(dmalloc(&p->mats, p->mats_cap * E));
  chk(dmalloc(&p->t1, (size_t)max_cuboids * p->t1_stride * sizeof(double)));
  chk(dmalloc(&p->laser, (size_t)laser_capacity * 6 * sizeof(double)));
  chk(dmalloc(&p->desc, (size_t)max_cuboids * sizeof(CuboidDesc)));
  chk(dmalloc(&p->row_items, (size_t)p->row_cap * sizeof(Item)));
  chk(dmalloc(&p->row_descs, (size_t)p->row_cap * sizeof(ItemDesc)));
  chk(dmalloc(&p->col_items, (size_t)p->col_cap * sizeof(Item)));
  chk(dmalloc(&p->prime_items, (size_t)p->col_cap * sizeof(Item)));
  chk(dmalloc(&p->row_slots, 3 * (size_t)max_cuboids * kMaxItems * sizeof(double)));
  chk(dmalloc(&p->col_slots, (size_t)max_cuboids * kMaxColBlk * sizeof(double)));
  chk(dmalloc(&p->aux_slots, 2 * (size_t)max_cuboids * kMaxItems * sizeof(double)));
  chk(dmalloc(&p->done_at, (size_t)max_cuboids * sizeof(int)));
  chk(dmalloc(&p->status, (size_t)kStatusStride * max_cuboids * sizeof(double)));
  chk(dmalloc(&p->step_idx, (size_t)max_cuboids * sizeof(int)));
  chk(dmalloc(&p->metrics, (size_t)metrics_rows * max_cuboids * kMetricStride * sizeof(double)));
  chk(dmalloc(&p->bar, ((size_t)max_cuboids + 1) * sizeof(unsigned)));
  chk(dmalloc(&p->mats16, (size_t)p->mats16_cap * 2));
  chk(dmalloc(&p->res_list, (size_t)p->res_cap * sizeof(int)));
  if (dtype == SGH_F32) {
    chk(dmalloc(&p->zwu, (size_t)p->total_S * 4 * sizeof(float)));
    chk(dmalloc(&p->res_work, (size_t)max_cuboids * kResWork * sizeof(float)));
    chk(dmalloc(&p->res_xch, (size_t)max_cuboids * kResXchBytes));
    chk(resident_allow_lds());
  }
  if (e == hipSuccess) e = hipMemset(p->bar, 0, ((size_t)max_cuboids + 1) * sizeof(unsigned));
  if (e == hipSuccess) {
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  }
  if (e != hipSuccess) {
    free_plan(p);
    set_error("sgh_batch_pool_create: device allocation failed: %s", hipGetErrorString(e));
    return SGH_ERR_NOMEM;
  }
  chk(hipMemcpy(p->desc, p->hdesc.data(), (size_t)max_cuboids * sizeof(CuboidDesc),
                hipMemcpyHostToDevice));
  if (e != hipSuccess) {
    free_plan(p);
    return fail_hip(e, "sgh_batch_pool_create init");
  }
  *plan_out = p;
  return SGH_OK;
}

int sgh_batch_add(void* plan, const sgh_cuboid_spec* spec, int* slot_out, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && spec && slot_out, SGH_ERR_ARG, "sgh_batch_add: bad args");
  SGH_REQUIRE(spec_ok(*spec), SGH_ERR_ARG, "sgh_batch_add: invalid cuboid spec");
  SGH_REQUIRE(spec->nz <= p->max_nz, SGH_ERR_ARG,
              "sgh_batch_add: nz = %d exceeds the pool's max_nz = %d", spec->nz, p->max_nz);
  int c = -1;
  for (int i = 0; i < p->n; ++i)
    if (!p->hdesc[i].live) { c = i; break; }
  SGH_REQUIRE(c >= 0, SGH_ERR_NOMEM, "sgh_batch_add: all %d slots in use", p->n);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipStreamSynchronize(st);  // the slot's previous occupant is finished
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_add");
  const int rc = add_cuboid(p, *spec, c, st);
  if (rc) return rc;
  *slot_out = c;
  return SGH_OK;
}

int sgh_batch_matrix_sizes(void* plan, int* out, int capacity) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && capacity >= 0 && (out || capacity == 0), SGH_ERR_ARG,
              "sgh_batch_matrix_sizes: bad args");
  for (int i = 0; i < (int)p->sizes.size() && i < capacity; ++i) out[i] = p->sizes[i];
  return (int)p->sizes.size();
}

int sgh_batch_remove(void* plan, int c, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && c >= 0 && c < p->n && p->hdesc[c].live, SGH_ERR_ARG,
              "sgh_batch_remove: bad slot");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_remove");
  const SlotAlloc& al = p->halloc[c];
  p->arena_a.release(al.a_off, al.a_len);
  p->arena_s.release(al.s_off, al.s_len);
  p->arena_l.release(al.l_off, al.l_len);
  mats_release(p, p->hdesc[c].nx);
  mats_release(p, p->hdesc[c].ny);
  p->hdesc[c] = CuboidDesc{};
  e = hipMemcpy(p->desc + c, &p->hdesc[c], sizeof(CuboidDesc), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_remove");
  p->pending_prime.erase(std::remove(p->pending_prime.begin(), p->pending_prime.end(), c),
                         p->pending_prime.end());
  p->n_live -= 1;
  p->dirty = true;
  return SGH_OK;
}

int sgh_batch_create(int dtype, int n, const sgh_cuboid_spec* specs, int metrics_rows,
                     void** plan_out, void* stream) {
  SGH_REQUIRE(plan_out && specs && n > 0, SGH_ERR_ARG, "sgh_batch_create: bad args");
  long long cap_a = 0, cap_s = 0, cap_l = 0;
  int max_nz = 1;
  for (int c = 0; c < n; ++c) {
    SGH_REQUIRE(spec_ok(specs[c]), SGH_ERR_ARG, "sgh_batch_create: invalid spec for cuboid %d", c);
    cap_a += (spec_state_elems(specs[c]) + 63) / 64 * 64;
    cap_s += ((long long)specs[c].nx * specs[c].ny + 63) / 64 * 64;
    cap_l += std::max(1, specs[c].n_steps);
    max_nz = std::max(max_nz, specs[c].nz);
  }
  void* pl = nullptr;
  int rc = sgh_batch_pool_create(dtype, n, cap_a, cap_s, cap_l, max_nz, metrics_rows, &pl);
  if (rc) return rc;
  Plan* p = static_cast<Plan*>(pl);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (int c = 0; c < n; ++c) {
    rc = add_cuboid(p, specs[c], c, st);
    if (rc) {
      free_plan(p);
      return rc;
    }
  }
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    free_plan(p);
    return fail_hip(e, "sgh_batch_create");
  }
  *plan_out = p;
  return SGH_OK;
}

int sgh_batch_destroy(void* plan) {
  if (!plan) return SGH_OK;
  hipDeviceSynchronize();
  free_plan(static_cast<Plan*>(plan));
  return SGH_OK;
}

int sgh_batch_state_layout(void* plan, int c, long long* a_offset, int* nx, int* ny, int* nz) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && c >= 0 && c < p->n && p->hdesc[c].live, SGH_ERR_ARG,
              "sgh_batch_state_layout: bad cuboid");
  const CuboidDesc& d = p->hdesc[c];
  if (a_offset) *a_offset = d.a_off;
  if (nx) *nx = d.nx;
  if (ny) *ny = d.ny;
  if (nz) *nz = d.nz;
  return SGH_OK;
}

int sgh_batch_set_state(void* plan, int c, const void* src, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && src && c >= 0 && c < p->n && p->hdesc[c].live, SGH_ERR_ARG,
              "sgh_batch_set_state: bad args");
  const CuboidDesc& d = p->hdesc[c];
  hipError_t e = hipMemcpyAsync((char*)p->a + d.a_off * p->elem, src, (size_t)d.S * d.nz * p->elem,
                                hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_set_state");
  if (std::find(p->pending_prime.begin(), p->pending_prime.end(), c) == p->pending_prime.end())
    p->pending_prime.push_back(c);
  return SGH_OK;
}

int sgh_batch_get_state(void* plan, int c, void* dst, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && dst && c >= 0 && c < p->n && p->hdesc[c].live, SGH_ERR_ARG,
              "sgh_batch_get_state: bad args");
  const CuboidDesc& d = p->hdesc[c];
  hipError_t e = hipMemcpyAsync(dst, (char*)p->a + d.a_off * p->elem, (size_t)d.S * d.nz * p->elem,
                                hipMemcpyDeviceToDevice, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_get_state");
  return SGH_OK;
}

int sgh_batch_run(void* plan, int n_steps, int use_graph, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && n_steps >= 0, SGH_ERR_ARG, "sgh_batch_run: bad args");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int rc = SGH_OK;
  if (!p->pending_prime.empty()) {
    std::vector<Item> items;
    for (int c : p->pending_prime)
      for (int b = 0; b < p->hdesc[c].ncolblk; ++b) items.push_back({c, b});
    rc = p->dtype == SGH_F64 ? prime_new<double>(p, items, st) : prime_new<float>(p, items, st);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(st);  // prime_items is reused by the next prime
    if (e != hipSuccess) return fail_hip(e, "sgh_batch_run prime");
    p->pending_prime.clear();
  }
  if (p->dirty) {
    rc = rebuild_lists(p, st);
    if (rc) return rc;
  }
  if (p->hrow.empty() && p->hres.empty()) return SGH_OK;
  if (p->timing) use_graph = 0;
  if (use_graph) {
    if (p->graph_exec && p->graph_stream != st) drop_graph(p);
    if (!p->graph_exec) {
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return fail_hip(e, "sgh_batch_run capture");
      rc = p->dtype == SGH_F64 ? enqueue_step<double>(p, st) : enqueue_step<float>(p, st);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail_hip(e, "sgh_batch_run end capture");
      e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        hipGraphDestroy(g);
        return fail_hip(e, "sgh_batch_run instantiate");
      }
      p->graph = g;
      p->graph_stream = st;
    }
    for (int s = 0; s < n_steps; ++s) {
      hipError_t e = hipGraphLaunch(p->graph_exec, st);
      if (e != hipSuccess) return fail_hip(e, "sgh_batch_run graph launch");
    }
    return SGH_OK;
  }
  for (int s = 0; s < n_steps; ++s) {
    rc = p->dtype == SGH_F64 ? enqueue_step<double>(p, st) : enqueue_step<float>(p, st);
    if (rc) return rc;
  }
  return SGH_OK;
}

// A persistent step kernel whose cuboid barrier timed out (its workgroups were not all
// resident) leaves an error word; the synchronous readers report it and clear it.
static int check_loop_error(Plan* p, const char* where) {
  int err = 0;
  hipError_t e = hipMemcpy(&err, p->bar + p->n, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail_hip(e, where);
  if (err) {
    const int zero = 0;
    hipMemcpy(p->bar + p->n, &zero, sizeof(int), hipMemcpyHostToDevice);
    set_error("%s: a cuboid barrier of the persistent step kernel or of a resident team "
              "timed out (not every workgroup was resident); the steps since the last check "
              "are invalid. Turn the persistent kernel off with sgh_batch_set_loop(plan, 0, "
              "stream), teams with sgh_batch_set_team_max(plan, 1, stream)", where);
    return SGH_ERR_STATE;
  }
  return SGH_OK;
}

int sgh_batch_read_metrics(void* plan, int c, double* host_out, int rows, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && host_out && c >= 0 && c < p->n && rows >= 0 && rows <= p->metrics_rows,
              SGH_ERR_ARG, "sgh_batch_read_metrics: bad args");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(host_out, p->metrics + (size_t)c * p->metrics_rows * kMetricStride,
                                (size_t)rows * kMetricStride * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_read_metrics");
  return check_loop_error(p, "sgh_batch_read_metrics");
}

int sgh_batch_steps_done(void* plan, int* host_out, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && host_out, SGH_ERR_ARG, "sgh_batch_steps_done: bad args");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipMemcpyAsync(host_out, p->step_idx, (size_t)p->n * sizeof(int),
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_steps_done");
  return check_loop_error(p, "sgh_batch_steps_done");
}

// The persistent step kernel is switched off: a run of the GPU test suite that included it
// (through the autotuner's candidates) ended in a GPU memory fault whose cause is not found.
// Until it is, requests to turn it on are refused and the per-pass launches run.
constexpr bool kLoopEnabled = false;

int sgh_batch_set_loop(void* plan, int enable, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p, SGH_ERR_ARG, "sgh_batch_set_loop: null plan");
  SGH_REQUIRE(kLoopEnabled || enable == 0, SGH_ERR_STATE,
              "sgh_batch_set_loop: the persistent step kernel is switched off (a GPU memory "
              "fault in a run that used it is not yet explained); the per-pass launches run");
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_set_loop");
  if (p->loop_pref != (enable != 0)) {
    p->loop_pref = enable != 0;
    p->dirty = true;  // work lists and the step graph are rebuilt by the next run
  }
  return SGH_OK;
}

int sgh_batch_set_loop_timeout(void* plan, double seconds, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && seconds >= 0.0 && seconds <= 60.0, SGH_ERR_ARG,
              "sgh_batch_set_loop_timeout: seconds must be in [0, 60]");
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_set_loop_timeout");
  p->barrier_ticks = (unsigned long long)(seconds * 1e8);
  p->dirty = true;  // the captured step graph holds the old bound
  return SGH_OK;
}

int sgh_batch_set_team_max(void* plan, int team_max, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && team_max >= 1 && team_max <= kResTeamMax, SGH_ERR_ARG,
              "sgh_batch_set_team_max: team_max must be in [1, %d]", kResTeamMax);
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_set_team_max");
  if (p->team_max != team_max) {
    p->team_max = team_max;
    p->dirty = true;  // work lists and the step graph are rebuilt by the next run
  }
  return SGH_OK;
}

int sgh_batch_team_sizes(void* plan, int* out, int capacity) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && capacity >= 0 && (out || capacity == 0), SGH_ERR_ARG,
              "sgh_batch_team_sizes: bad args");
  for (int c = 0; c < capacity && c < p->n; ++c)
    out[c] = c < (int)p->hteam.size() ? p->hteam[c] : 0;
  return p->n;
}

int sgh_batch_loop_state(void* plan, int* out2) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && out2, SGH_ERR_ARG, "sgh_batch_loop_state: bad args");
  out2[0] = p->loop_pref ? 1 : 0;
  out2[1] = p->loop_on ? 1 : 0;
  return SGH_OK;
}

int sgh_batch_set_timing(void* plan, int enable) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p, SGH_ERR_ARG, "sgh_batch_set_timing: null plan");
  if (enable && !p->ev[0])
    for (auto& e : p->ev) {
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return fail_hip(r, "sgh_batch_set_timing");
    }
  p->timing = enable != 0;
  p->t_surface_ms = p->t_volume_ms = 0.0;
  p->n_surface = p->n_volume = 0;
  return SGH_OK;
}

int sgh_batch_timing(void* plan, double* out4) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && out4, SGH_ERR_ARG, "sgh_batch_timing: bad args");
  out4[0] = p->t_surface_ms;
  out4[1] = (double)p->n_surface;
  out4[2] = p->t_volume_ms;
  out4[3] = (double)p->n_volume;
  return SGH_OK;
}

#ifdef SGH_BLOCK_TIMESTAMPS
int sgh_debug_block_timestamps(unsigned long long* host_out, int n) {
  (void)n;
  return hipMemcpyFromSymbol(host_out, HIP_SYMBOL(g_block_ts), (size_t)4 * 8192 * 6 * 8) == hipSuccess ? 0 : -2;
}
#endif

int sgh_batch_variants(int* out_pairs, int capacity) {
  if (out_pairs)
    for (int i = 0; i < kNumSurfVariants && i < capacity; ++i) {
      out_pairs[2 * i] = kSurfVariants[i].rows;
      out_pairs[2 * i + 1] = kSurfVariants[i].threads;
    }
  return kNumSurfVariants;
}

int sgh_batch_set_variant(void* plan, int rows, int threads, int lds_bytes, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p, SGH_ERR_ARG, "sgh_batch_set_variant: null plan");
  bool known = false;
  for (const SurfVariant& v : kSurfVariants) known |= v.rows == rows && v.threads == threads;
  SGH_REQUIRE(known, SGH_ERR_ARG, "sgh_batch_set_variant: no variant rows=%d threads=%d", rows,
              threads);
  SGH_REQUIRE(lds_bytes >= 4096 && lds_bytes <= kLdsMax, SGH_ERR_ARG,
              "sgh_batch_set_variant: lds_bytes=%d outside [4096, %d]", lds_bytes, kLdsMax);
  const bool res = rows == kResidentRows;
  SGH_REQUIRE(!res || p->dtype == SGH_F32, SGH_ERR_ARG,
              "sgh_batch_set_variant: the resident step kernel runs in mixed precision only");
  // the resident variant keeps a per-pass launch shape for the cuboids it cannot hold (a
  // surface side wider than 128): 16-row items, 1024 threads, the given LDS budget
  if (res) {
    rows = 16;
    threads = 1024;
  }
  const int bud = surface_budget(p->dtype, kernel_rows(rows), lds_bytes);
  SGH_REQUIRE(bud >= 0, SGH_ERR_ARG,
              "sgh_batch_set_variant: %d LDS bytes cannot stage a 256-wide operand", lds_bytes);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipStreamSynchronize(st);  // no step of the old variant is in flight
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_set_variant");
  p->rows = rows;
  p->threads = threads;
  p->bud = bud;
  p->shm = (size_t)lds_bytes;
  p->resident = res;
  for (int c = 0; c < p->n; ++c) {
    CuboidDesc& d = p->hdesc[c];
    if (!d.live) continue;
    d.item_rows = rows == 0 ? 16 : rows;
    d.nrowblk = (d.ny + d.item_rows - 1) / d.item_rows;
    e = hipMemcpy(p->desc + c, &d, sizeof(CuboidDesc), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "sgh_batch_set_variant");
  }
  p->dirty = true;  // work lists and the step graph are rebuilt by the next run
  return SGH_OK;
}

int sgh_batch_rewind(void* plan, void* stream) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p, SGH_ERR_ARG, "sgh_batch_rewind: null plan");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_rewind");
  const int zero = 0;
  const double zstat[kStatusStride] = {0, 0, 0, 0};
  auto chk = [&](hipError_t r) { if (e == hipSuccess) e = r; };
  for (int c = 0; c < p->n; ++c) {
    if (!p->hdesc[c].live) continue;
    const SlotAlloc& al = p->halloc[c];
    chk(hipMemcpy(p->step_idx + c, &zero, sizeof(int), hipMemcpyHostToDevice));
    chk(hipMemcpy(p->done_at + c, &zero, sizeof(int), hipMemcpyHostToDevice));
    chk(hipMemcpy(p->status + kStatusStride * c, zstat, sizeof(zstat), hipMemcpyHostToDevice));
    for (int arr = 0; arr < kSurfArrays; ++arr)
      chk(hipMemsetAsync((char*)p->surf + (arr * p->total_S + al.s_off) * p->elem, 0,
                         al.s_len * p->elem, st));
    if (std::find(p->pending_prime.begin(), p->pending_prime.end(), c) == p->pending_prime.end())
      p->pending_prime.push_back(c);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail_hip(e, "sgh_batch_rewind");
  return SGH_OK;
}

int sgh_batch_get_variant(void* plan, int* out3) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && out3, SGH_ERR_ARG, "sgh_batch_get_variant: bad args");
  out3[0] = p->resident ? kResidentRows : p->rows;
  out3[1] = p->resident ? kResThreads : p->threads;
  out3[2] = (int)p->shm;
  return SGH_OK;
}

int sgh_batch_info(void* plan, long long* out6) {
  Plan* p = static_cast<Plan*>(plan);
  SGH_REQUIRE(p && out6, SGH_ERR_ARG, "sgh_batch_info: bad args");
  long long used_a = 0;
  for (int c = 0; c < p->n; ++c)
    if (p->hdesc[c].live) used_a += p->halloc[c].a_len;
  out6[0] = used_a;
  out6[1] = p->total_S;
  out6[2] = (long long)p->hrow.size();
  out6[3] = (long long)p->hcol.size();
  out6[4] = p->max_it;
  out6[5] = (long long)p->shm;
  return SGH_OK;
}

}  // extern "C"
