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
// bf16 products (hi·hi + hi·lo + lo·hi of x = hi + lo, FP32 accumulation). Every stage
// computes O[a][t] = Σ_k D[a][k]·W[k][t]: the field D (split bf16, in LDS as [a][k]) is the
// A operand, the cosine matrix W (split bf16, read from global memory, L2-resident, held in
// registers for the stage) the B operand, and the accumulator's four consecutive a of one t
// per lane are stored as [t][a], the layout the next stage contracts over a in:
//   inverse  V[m][j] = Σ_n X[m][n]·B̂y[n][j]     T[j][i] = Σ_m V[j][m]·B̂x[m][i]
//   forward  R[i][n] = Σ_j q[i][j]·B̂y[n][j]     S[n][m] = Σ_i R[n][i]·B̂x[m][i]
// (X stored [m][n] and q stored [i][j] in LDS.) The second stage of each transform has t
// along x, so a wave's 16 consecutive lanes hold 16 consecutive x of the surface fields its
// epilogue reads and writes in global memory ([y][x] layout). Every extent is padded
// to a multiple of 32 (the MFMA k step) with zero matrix entries and zero field entries.
// Two LDS operand buffers alternate: a stage reads one and writes its result (or the field
// its epilogue forms) into the other, one 16×16 tile at a time, so a stage needs one
// barrier. The per-element fields (ZE, Ztt, Wtt, U, G, q_las) are read from global memory
// (L2) a tile ahead of their use. Surfaces up to 128 × 128; FP64 and wider surfaces use the
// per-pass launches.
// ---------------------------------------------------------------------------------
constexpr int kResWaves = kResThreads / kWave;

#ifdef SGH_RESIDENT_TIMESTAMPS
// Diagnostic build only: per-workgroup cycle counts (s_memtime) of the last launch. Slot 0
// entry, 1 step preparation done, 2 warm start done, 3 + 7k + {0, ..., 6} for Picard
// iteration k < 20: first inverse stage, second inverse stage + evaporation, q exchange,
// first forward stage, second forward stage + relaxation, X exchange, norm sums done (the
// exchanges of a single-member team take no time); 143 step end done; 144, 145 the cuboid's
// nx and ny, with the team size and the member index in bits 16 and up; 146 + 4s + {0..3}
// inside stage s (0: first inverse, 1: second forward) of iteration 5: stage entry, matrix
// fragments in registers, tiles done, barrier passed (wave 0's view); 154..157 inside the
// step preparation: AA summed, laser factors formed, flux pass done, laser power summed.
constexpr int kResTsSlots = 158, kResTsBlocks = 1024;
__device__ unsigned long long g_res_ts[kResTsBlocks * kResTsSlots];
#define RES_TS(slot)                                                                        \
  do {                                                                                      \
    if (threadIdx.x == 0 && blockIdx.x < kResTsBlocks)                                      \
      g_res_ts[blockIdx.x * kResTsSlots + (slot)] = __builtin_amdgcn_s_memtime();          \
  } while (0)
#else
#define RES_TS(slot) do {} while (0)
#endif
constexpr int kResBuf = kResMaxN * kResLd;   // bf16 elements of one hi or lo operand buffer

typedef __bf16 bf4v __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

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

// Tiles of one wave in a stage producing O[a][t] as NT × NA 16×16 tiles (t: the index the
// stage's matrix introduces, a: the field's free index), the t tiles numbered from toff and
// the a tiles from aoff (a team member's share of a stage). The waves are grouped per t
// tile (g waves each), so a wave's tiles share the matrix fragments it holds; wave w takes
// t tile toff + w / g and a tiles aoff + (w mod g) + g·q, q < nq. tt < 0: no tiles.
struct ResTiles {
  int tt, a0, g, nq;
  __device__ __forceinline__ ResTiles(int NT, int NA, int wave, int toff, int aoff) {
    g = kResWaves / NT > 0 ? kResWaves / NT : 1;
    tt = wave / g;
    a0 = wave % g;
    nq = tt < NT && a0 < NA ? (NA - a0 + g - 1) / g : 0;
    tt = nq == 0 ? -1 : tt + toff;
    a0 += aoff;
  }
  // a tile of the wave's tile q
  __device__ int atile(int q) const { return a0 + g * q; }
  // row a of the field whose k values lane `lane` reads for tile q
  __device__ int arow(int q, int lane) const { return atile(q) * 16 + (lane & 15); }
  // first of the four consecutive a the accumulator of tile q holds in lane `lane`
  __device__ int aout(int q, int lane) const { return atile(q) * 16 + 4 * (lane >> 4); }
  // t of lane `lane`'s outputs
  __device__ int tcol(int lane) const { return tt * 16 + (lane & 15); }
};

// One operand buffer pair (hi, lo) in LDS: bf16 [column][k] with row stride kResLd.
struct ResBuf {
  __bf16* hi;
  __bf16* lo;
};

// One lane's four outputs O[a..a+3][t] into an operand buffer as [t][a] (one 8-byte store
// each for hi and lo), split into bf16 hi / lo.
__device__ __forceinline__ void res_store(const ResBuf& b, int a, int t, const float (&x)[4]) {
  bf4v h, l;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const __bf16 hv = (__bf16)x[v];
    h[v] = hv;
    l[v] = (__bf16)(x[v] - (float)hv);
  }
  *reinterpret_cast<bf4v*>(b.hi + t * kResLd + a) = h;
  *reinterpret_cast<bf4v*>(b.lo + t * kResLd + a) = l;
}

// res_store, and (teams) the same two 8-byte hi / lo values to the exchange field at byte
// offset gb of the exchange area (same layout), device-coherently, so that the band is in
// device memory as soon as the tile is done instead of being copied out of LDS after the
// stage.
__device__ __forceinline__ void res_store_x(const ResBuf& b, __amdgpu_buffer_rsrc_t xr,
                                            int gb, bool team, int a, int t,
                                            const float (&x)[4]) {
  bf4v h, l;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const __bf16 hv = (__bf16)x[v];
    h[v] = hv;
    l[v] = (__bf16)(x[v] - (float)hv);
  }
  const int e = t * kResLd + a;
  *reinterpret_cast<bf4v*>(b.hi + e) = h;
  *reinterpret_cast<bf4v*>(b.lo + e) = l;
  if (team) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, h), xr, gb + 2 * e, 0,
                                          kCoherentAux);
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, l), xr,
                                          gb + 2 * (kResBuf + e), 0, kCoherentAux);
  }
}

// Matrix fragments of one wave for a stage: W[k][t] of its t tile, every k step (hi, lo).
struct ResW {
  bf8v h[4], l[4];
};

// Load slots [k0, k1) of the matrix fragments of the wave's t tile (no-op without tiles) from
// Wᵀ[t][k]: hi / lo row-major matrices at m16 + off_h / off_l in global memory, row stride
// ld, KS k steps of 32 (the offsets, not the pointers, are made opaque, so the loads stay
// global loads).
__device__ __forceinline__ void res_load_w(ResW& w, const ResTiles& t, const __bf16* m16,
                                           long long off_h, long long off_l, int ld, int KS,
                                           int lane, int k0, int k1) {
  if (t.tt < 0) return;
  const __bf16* Wh = m16 + opaque(off_h);
  const __bf16* Wl = m16 + opaque(off_l);
  lane = opaque_lane(lane);
  const long long row = (long long)t.tcol(lane) * opaque(ld) + 8 * (lane >> 4);
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    if (ks < k0 || ks >= k1) continue;
    // k steps beyond KS re-read the last one (unused; unconditional loads keep the wait
    // counts exact)
    const int kk = ks < KS ? ks : KS - 1;
    w.h[ks] = *reinterpret_cast<const bf8v*>(Wh + row + 32 * kk);
    w.l[ks] = *reinterpret_cast<const bf8v*>(Wl + row + 32 * kk);
  }
}

// One stage with KS k steps (K = 32·KS): for each of the wave's tiles q < nq,
// acc = Σ_k D[a][k]·W[k][t] as three bf16 products into one accumulator, then
// epi(q, slot, acc, lane). W: the fragments in `w`, loaded beforehand. D: the source buffer,
// read AH k steps ahead of the products that use it through a ring of R fragment pairs (the
// k steps of all tiles form one sequence; beyond the wave's last tile the reads repeat that
// tile, so every read is unconditional and in bounds). Tile q uses ring slots, accumulator
// and epilogue slot q mod 2: the loop runs two tiles per trip with no branch inside (the
// code stays small enough for the instruction cache, and the wait counts for the ring stay
// exact), the first tile and the last one or two are peeled. The epilogue of tile q runs
// after tile q + 1's products are issued, so its arithmetic overlaps them;
// pre(q, slot, lane) starts the loads that epilogue needs once the epilogue of tile q − 2
// has used the slot. next(k0, k1) loads slots [k0, k1) of the following stage's matrix
// fragments: slots KS..3 at the start (this stage does not use them), the others after the
// last tile's products, so that the loads overlap the last epilogues and the barrier.
// EARLY: the caller has started the loads of the first EARLY tiles (0, 1 or 2) earlier.
template <int KS, int EARLY, typename Next, typename Pre, typename Epi>
__device__ __forceinline__ void res_tiles(const ResTiles& t, ResW& w, const ResBuf& src,
                                          int lane, Next&& next, Pre&& pre, Epi&& epi) {
  constexpr int R = KS == 1 ? 2 : (KS == 3 ? 3 : 4);
  constexpr int AH = R - 1;
  static_assert((2 * KS) % R == 0, "ring slots repeat every two tiles");
  const int nq = t.nq;
  bf8v dh[R], dl[R];
  // (lane is laundered per tile: addresses are formed where they are used instead of being
  // hoisted out of the loop and kept live)
  auto dload = [&](int q, int ks, int slot, int ln) __attribute__((always_inline)) {
    const int qq = q < nq ? q : nq - 1;
    const int off = t.arow(qq, ln) * kResLd + 8 * (ln >> 4) + 32 * ks;
    dh[slot] = *reinterpret_cast<const bf8v*>(src.hi + off);
    dl[slot] = *reinterpret_cast<const bf8v*>(src.lo + off);
  };
  f4v acc[2];
  auto mfma = [&](int q, int par, int ln) __attribute__((always_inline)) {
    f4v& c = acc[par];
    c = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int s = par * KS + ks;   // k step within the two-tile period
      dload(q + (ks + AH) / KS, (ks + AH) % KS, (s + AH) % R, ln);
      c = mma_bf16(dh[s % R], w.h[ks], c);
      c = mma_bf16(dh[s % R], w.l[ks], c);
      c = mma_bf16(dl[s % R], w.h[ks], c);
    }
  };
  next(KS, 4);
  if (EARLY < 1) pre(0, 0, lane);
#pragma unroll
  for (int s = 0; s < AH; ++s) dload(s / KS, s % KS, s % R, lane);
  mfma(0, 0, lane);
  if (EARLY < 2) pre(1, 1, lane);
  int q = 1;
#pragma unroll 1
  for (; q + 1 < nq; q += 2) {
    int ln = opaque_lane(lane);
    mfma(q, 1, ln);
    epi(q - 1, 0, acc[0], ln);
    pre(q + 1, 0, ln);
    ln = opaque_lane(lane);
    mfma(q + 1, 0, ln);
    epi(q, 1, acc[1], ln);
    pre(q + 2, 1, ln);
  }
  const int ln = opaque_lane(lane);
  if (q < nq) {   // nq even: tile q = nq − 1 remains
    mfma(q, 1, ln);
    next(0, KS);
    epi(q - 1, 0, acc[0], ln);
    epi(q, 1, acc[1], ln);
  } else {        // nq odd: tile q − 1 = nq − 1 awaits its epilogue
    next(0, KS);
    epi(q - 1, 0, acc[0], ln);
  }
}

// One stage: the wave's tiles (res_tiles, specialised on the number of k steps; it reloads the
// matrix fragments for the following stage through next), then the barrier after which the
// epilogue's stores are visible. A wave without tiles in this stage reloads them at once.
template <int EARLY = 0, typename Next, typename Pre, typename Epi>
__device__ __forceinline__ void res_stage(const ResTiles& t, ResW& w, int KS, const ResBuf& src,
                                          int lane, Next&& next, Pre&& pre, Epi&& epi,
                                          int ts = -1) {
  (void)ts;   // diagnostic builds: stamps ts .. ts + 3 inside the stage
#ifdef SGH_RESIDENT_TIMESTAMPS
  if (ts >= 0) RES_TS(ts);
#endif
  if (t.tt >= 0) {
    lane = opaque_lane(lane);
    // the fragments are pinned in registers for the stage here, after the barrier that
    // ended the previous stage (their loads were issued before it): without this the
    // register allocator reloads them from memory for every tile
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) asm volatile("" : "+v"(w.h[ks]), "+v"(w.l[ks]));
#ifdef SGH_RESIDENT_TIMESTAMPS
    if (ts >= 0) RES_TS(ts + 1);
#endif
    switch (KS) {
      case 1: res_tiles<1, EARLY>(t, w, src, lane, next, pre, epi); break;
      case 2: res_tiles<2, EARLY>(t, w, src, lane, next, pre, epi); break;
      case 3: res_tiles<3, EARLY>(t, w, src, lane, next, pre, epi); break;
      default: res_tiles<4, EARLY>(t, w, src, lane, next, pre, epi); break;
    }
#ifdef SGH_RESIDENT_TIMESTAMPS
    if (ts >= 0) RES_TS(ts + 2);
#endif
  } else {
    next(0, 4);
  }
  __syncthreads();
#ifdef SGH_RESIDENT_TIMESTAMPS
  if (ts >= 0) RES_TS(ts + 3);
#endif
}

// v plus the value of the lane the DPP control selects (0 where it selects none, and in the
// rows the row mask leaves out), both halves of the double moved by the same control.
template <int Ctrl, int RowMask>
__device__ __forceinline__ double res_dpp_add(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, Ctrl, RowMask, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), Ctrl, RowMask, 0xf, true);
  return v + __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double res_lane(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
// Sum over the wave in a fixed order, with DPP lane moves instead of LDS shuffles: within
// each row of 16 by row shifts 1, 2, 4, 8 (lane 15 then holds the row's sum), then across
// rows by the row broadcasts of lane 15 (rows 1, 3) and lane 31 (rows 2, 3); the total is in
// lane 63 and returned to every lane.
__device__ __forceinline__ double res_wave_sum(double v) {
  v = res_dpp_add<0x111, 0xf>(v);
  v = res_dpp_add<0x112, 0xf>(v);
  v = res_dpp_add<0x114, 0xf>(v);
  v = res_dpp_add<0x118, 0xf>(v);
  v = res_dpp_add<0x142, 0xa>(v);
  v = res_dpp_add<0x143, 0xc>(v);
  return res_lane(v, 63);
}
// Sums of up to three sets of kResWaves wave partials at once: set r in lanes 16r..16r+7 of
// a row, summed by row shifts 1, 2, 4 (lane 16r + 7 then holds it).
static_assert(kResWaves <= 8, "one row of 16 lanes sums the wave partials by shifts 1, 2, 4");
__device__ __forceinline__ double res_sum8(double v) {
  v = res_dpp_add<0x111, 0xf>(v);
  v = res_dpp_add<0x112, 0xf>(v);
  v = res_dpp_add<0x114, 0xf>(v);
  return v;
}

// Sum of a float over each row of 16 lanes by row shifts 1, 2, 4, 8 (lane 16r + 15 then
// holds row r's sum), in FP32.
template <int Ctrl>
__device__ __forceinline__ float res_dpp_addf(float v) {
  return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                                           0, __builtin_bit_cast(int, v), Ctrl, 0xf, 0xf, true));
}
__device__ __forceinline__ float res_row_sum(float v) {
  v = res_dpp_addf<0x111>(v);
  v = res_dpp_addf<0x112>(v);
  v = res_dpp_addf<0x114>(v);
  return res_dpp_addf<0x118>(v);
}

// Sum of a double over the workgroup in a fixed order (wave sums, then the kResWaves wave
// partials); every thread receives it. sh holds kResWaves doubles; ends with a barrier.
__device__ __forceinline__ double res_sum(double v, double* sh) {
  v = res_wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = v;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const double t = res_lane(res_sum8(lane < kResWaves ? sh[lane] : 0.0), kResWaves - 1);
  __syncthreads();
  return t;
}

// The cuboid's surface arrays through buffer resources, with 32-bit byte offsets instead of
// 64-bit addresses: r covers the separate arrays (array k — G 0, q_las 5, q_evap 7 — element
// e at byte 4·(k·T + e) past the cuboid's first element), rz the packed (ZE, Ztt, Wtt, U)
// records (element e at byte 16·e). An element outside the surface gets the offset kResOut,
// beyond either range, where the hardware's range check returns 0 for a load and drops a
// store.
constexpr unsigned kResOut = 0x80000000u;
struct ResFields {
  __amdgpu_buffer_rsrc_t r, rz;
  int fs;   // 4·T
  __device__ float load(int k, unsigned vo) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, k * fs, 0));
  }
  __device__ void store(int k, unsigned vo, float v) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, k * fs, 0);
  }
  __device__ f4v record(unsigned vo16) const {
    return __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rz, (int)vo16, 0, 0));
  }
};

// Surface element indices y·nx + x of one lane's four outputs of tile q in a stage whose t
// index runs along x and whose a index runs along y (the two stages with epilogues): byte
// offsets into the separate arrays (vo) and into the packed records (vo16); kResOut outside
// the surface or beyond the wave's tiles.
__device__ __forceinline__ void res_voffs(const ResTiles& t, int q, int lane, int nx, int ny,
                                          unsigned (&vo)[4], unsigned (&vo16)[4]) {
  const int x = t.tcol(lane), y0 = t.aout(q, lane);
  const bool in = q < t.nq && x < nx;
  const unsigned e = (unsigned)(y0 * nx + x);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const bool ok = in && y0 + v < ny;
    vo[v] = ok ? 4u * (e + v * nx) : kResOut;
    vo16[v] = ok ? 16u * (e + v * nx) : kResOut;
  }
}

// Teams. A cuboid runs on a team of TP workgroups (members mp = 0 .. TP − 1; TP = 1 for most
// cuboids, more for the largest surfaces so that the batch fills the CUs), each owning the
// 16-row tile bands [a_lo, a_hi) along y, in point space (j) and in mode space (n). Of the
// four stages of an iteration, the two whose matrix introduces y (the first stages, t
// along y) compute the member's t tiles and read the whole field; the two with t along x
// (the second stages, with the epilogues) compute the member's a tiles, whose operand rows
// the member's own first stage wrote. So a member needs the others' data twice per
// iteration, the q field before the forward transform and the X field before the inverse
// one: after each of those epilogues every member stores its column band of the field to
// the cuboid's exchange area, the team meets at a barrier (all members are resident: the
// host gives teams only to batches that fit one workgroup per CU, members of a team on one
// XCD), and every member copies the whole field into LDS. The per-tile norm sums go
// through the exchange area as well. Every sum of the step is formed per 16×16 tile and the
// tiles are combined in a fixed order (AA: by every member over the whole surface; the laser
// power per 16-row band, by the band's member, and the bands in order by member 0), so all
// members take the same Picard decision and a cuboid's arithmetic is the same for every team
// size.
__global__ void __launch_bounds__(kResThreads)
k_b_resident(BatchPtrs<float> P, const __bf16* __restrict__ m16, const int* __restrict__ list,
             unsigned long long ticks) {
  using F = float;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double sh[3][kResWaves];
  __shared__ __align__(16) float nsl[3][kResNormSlots];   // norm sums (single-member teams)
  __shared__ double psl[kResMaxN / 16][kResWaves];        // laser power: band × wave partials
  const int ent = list[blockIdx.x];
  if (ent < 0) return;
  const int c = ent >> 6, TP = ((ent >> 3) & 7) + 1, mp = ent & 7;
  const CuboidDesc& d = P.desc[c];
  const int step = P.step_idx[c];
  if (!d.live || step >= d.n_steps) {
    // the volume pass leaves it alone
    if (threadIdx.x == 0 && mp == 0) P.done_at[c] = 0x7fffffff;
    return;
  }
  RES_TS(0);
  __bf16* lds = reinterpret_cast<__bf16*>(smem_raw);
  const ResBuf buf0{lds, lds + kResBuf}, buf1{lds + 2 * kResBuf, lds + 3 * kResBuf};
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int nx = d.nx, ny = d.ny;
  const int nxp = pad32(nx), nyp = pad32(ny);
  const int KX = nxp / 32, KY = nyp / 32;
  const int NTB = nxp / 16, NAY = nyp / 16;
  // the member's y tile bands (the host keeps TP ≤ NAY: at least one each)
  const int a_lo = mp * NAY / TP, a_hi = (mp + 1) * NAY / TP;
  // split-bf16 tables: B̂ hi, B̂ lo, B̂ᵀ hi, B̂ᵀ lo, each np × np
  const long long X16 = d.bx16_off, Y16 = d.by16_off;   // element offsets into m16
  const long long mx = (long long)nxp * nxp, my = (long long)nyp * nyp;
  // tA: t along y (n or j), a along x (i or m): the first stage of each transform, the
  // member's t tiles; tB: t along x (m or i), a along y (n or j): the second stage, with the
  // epilogues, the member's a tiles
  const ResTiles tA(a_hi - a_lo, NTB, wave, a_lo, 0);
  const ResTiles tB(NTB, a_hi - a_lo, wave, 0, a_lo);
  // work-slot index (in 16-byte units) of this lane's outputs of tB tile q (clamped to the
  // wave's last tile)
  auto tidx = [&](int q, int ln) __attribute__((always_inline)) {
    const int qq = q < tB.nq ? q : tB.nq - 1;
    return (tB.atile(qq) * NTB + tB.tt) * 64 + ln;
  };
  // the cuboid's exchange area (teams): X and q fields in the LDS layout (hi, then lo), the
  // per-tile norm sums [3][kResTiles], one T_max per member, the laser power per band
  unsigned char* const xch = P.res_xch + (long long)c * kResXchBytes;
  const __amdgpu_buffer_rsrc_t xr =
      __builtin_amdgcn_make_buffer_rsrc(xch, (short)0, (int)kResXchBytes, 0x00020000);
  __bf16* const xX = reinterpret_cast<__bf16*>(xch);
  __bf16* const xQ = xX + kResXchField;
  float* const xN = reinterpret_cast<float*>(xch + kResXchNorm);
  double* const xT = reinterpret_cast<double*>(xN + 3 * kResNormSlots);
  double* const xP = xT + kResTeamMax;   // laser power per 16-row band
  const int gbX = 0, gbQ = (int)(2 * kResXchField);   // byte offsets of the X and q fields
  unsigned arrivals = 0;
  // (teams) the member's column band [16·a_lo, 16·a_hi) of every row < nxp of operand
  // buffer b → exchange field g (16-byte device-coherent stores); team barrier; rows < nxp
  // of g (the whole field, LDS layout) → b by device-coherent LDS DMA
  auto exchange = [&](const ResBuf& b, const __bf16* g, bool stored, auto&& after_barrier)
                      __attribute__((always_inline)) {
    const int ur = 2 * (a_hi - a_lo);   // 16-byte units of the band in one row
    const int per = nxp * ur;           // ... in the hi (or the lo) field
    const long long gb = (long long)(g - xX) * 2;   // byte offset of g in the exchange area
    // (stored: the stage's epilogues wrote the band to g already)
    for (int u = stored ? 2 * per : tid; u < 2 * per; u += kResThreads) {
      const int h = u >= per ? 1 : 0;
      const int v = u - h * per;
      const int r = v / ur;
      const int off = r * kResLd + 16 * a_lo + 8 * (v - r * ur);
      const bf8v x = *reinterpret_cast<const bf8v*>((h ? b.lo : b.hi) + off);
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4v, x), xr,
                                             (int)(gb + 2LL * (h * kResBuf + off)), 0,
                                             kCoherentAux);
    }
    arrivals += TP;
    cuboid_barrier(P.bar + c, arrivals, P.loop_err, ticks);
    after_barrier();   // loads that complete with the copy's wait
    const int units = nxp * (kResLd / 8);   // 16-byte units of rows < nxp
    for (int h = 0; h < 2; ++h) {
      const __bf16* src = g + h * kResBuf;
      __bf16* const dst = h ? b.lo : b.hi;
      for (int u0 = wave * kWave; u0 < units; u0 += kResThreads) {
        const int u = u0 + lane;
        if (u < units) dma16<kCoherentAux>(src + 8 * u, dst + 8 * u0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  };
  const int so = d.s_off;
  ResFields fld;
  fld.r = __builtin_amdgcn_make_buffer_rsrc(P.G + so, (short)0,
                                            (int)(4 * (kSurfArrays * P.surf_stride - so)),
                                            0x00020000);   // raw 32-bit words, range-checked
  fld.fs = (int)(4 * P.surf_stride);
  fld.rz = __builtin_amdgcn_make_buffer_rsrc(P.zwu + 4LL * so, (short)0,
                                             (int)(16 * (P.surf_stride - so)), 0x00020000);
  F* const q_las = P.q_las + so;
  // the cuboid's work slot: G, q_evap and q_las of the step in the second stages' tile order
  // (tile (a, t) of tB at ((a·NTB + t)·64 + lane)·4 + v), so that a lane's four outputs of a
  // tile are one 16-byte access; written back to the surface arrays at the step end
  float* const wk = P.res_work + (long long)c * kResWork;
  f4v* const Gt = reinterpret_cast<f4v*>(wk);
  f4v* const QEt = reinterpret_cast<f4v*>(wk + kResMaxN * kResMaxN);
  float* const QL = wk + 2 * kResMaxN * kResMaxN;
  const f4v* const QLt = reinterpret_cast<const f4v*>(QL);
  auto no_pre = [](int, int, int) __attribute__((always_inline)) {};

  // ---- step preparation: AA, laser flux, warm-start flux (buffer 0, as [i][j]) ----
  double aa = 0.0;
  for (int b = tid; b < d.ncolblk; b += kResThreads) aa += P.col_slots[d.col_item0 + b];
  const double AA = res_sum(aa, sh[0]);
  RES_TS(154);
  {
    const double* L = P.laser + (d.laser_off + step) * 6;
    const double x0 = L[0], y0 = L[1], power = L[2], on = L[3], vx = L[4], vy = L[5];
    const double I0 = on != 0.0 ? d.absorptivity * power / (d.area_factor * d.rx * d.ry) : 0.0;
    const double shy = vy * d.dt / d.dy, shx = vx * d.dt / d.dx;
    const F* qo = P.q_evap_old + so;
    // the laser flux is separable (cell-integrated or Gaussian profile): its x and y factors
    // are formed once, in buffer 1 (unused until the first transform)
    double* const fac = reinterpret_cast<double*>(buf1.hi);   // x: [0, nx), y: [128, 128 + ny)
    const bool sep = d.cell_integrated || d.order == 2.0;
    if (sep) {
      for (int e = tid; e < nx + ny; e += kResThreads) {
        const bool isx = e < nx;
        const int k = isx ? e : e - nx;
        const double h = isx ? d.dx : d.dy, r = isx ? d.rx : d.ry, c0 = isx ? x0 : y0;
        const double pos = (k + 0.5) * h;
        double f;
        if (d.cell_integrated) {
          const double s2 = sqrt(2.0), cc = sqrt(M_PI / 8.0);
          const double lo = (pos - c0 - h / 2.0) * s2 / r, hi = (pos - c0 + h / 2.0) * s2 / r;
          f = r * cc * (erf(hi) - erf(lo)) / h;
        } else {
          const double u = (pos - c0) / r;
          f = exp(-2.0 * (u * u));
        }
        fac[isx ? k : kResMaxN + k] = f;
      }
      __syncthreads();
    }
    RES_TS(155);
    // the laser flux at surface element (i, j)
    auto laser_at = [&](int i, int j) __attribute__((always_inline)) -> double {
      if (sep) return I0 * (fac[kResMaxN + j] * fac[i]);
      const double x = (i + 0.5) * d.dx, y = (j + 0.5) * d.dy;
      if (d.cell_integrated) {
        const double s2 = sqrt(2.0), cc = sqrt(M_PI / 8.0);
        const double ax = (x - x0 - d.dx / 2.0) * s2 / d.rx, bx = (x - x0 + d.dx / 2.0) * s2 / d.rx;
        const double ay = (y - y0 - d.dy / 2.0) * s2 / d.ry, by = (y - y0 + d.dy / 2.0) * s2 / d.ry;
        return I0 * ((d.ry * cc * (erf(by) - erf(ay)) / d.dy) * (d.rx * cc * (erf(bx) - erf(ax)) / d.dx));
      }
      const double ux = (x - x0) / d.rx, uy = (y - y0) / d.ry;
      const double r2 = ux * ux + uy * uy;
      return I0 * (d.order == 2.0 ? exp(-2.0 * r2) : exp(-2.0 * pow(r2, d.order / 2.0)));
    };
    // four elements per pass over the member's rows of the padded surface (zero operand
    // entries outside it): the previous flux's neighbours of all four are loaded before any
    // is used
    const ResBuf& b0 = buf0;
    const int eb = 16 * nxp;   // elements of one 16-row band (at most 4 per thread)
#pragma unroll 1
    for (int b = a_lo; b < a_hi; ++b) {
      const int e0 = b * eb + tid, e_hi = (b + 1) * eb;
      double pb = 0.0;   // this thread's laser power of the band
      float nb[4][4];
      int ij[4][2];
      double tw[4][2];
      bool in_s[4], in_r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + u * kResThreads;
        const int j = e / nxp, i = e - j * nxp;
        ij[u][0] = i;
        ij[u][1] = e < e_hi ? j : nyp;
        in_s[u] = e < e_hi && i < nx && j < ny;
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
        float qn = 0.f, qlw = 0.f;
        if (in_s[u]) {
          const F qlf = (F)laser_at(i, j);
          // bilinear shift of the previous evaporation flux, scipy order-1 'constant'
          const double ty = tw[u][0], tx = tw[u][1];
          const double sv = in_r[u] ? (1.0 - ty) * ((1.0 - tx) * (double)nb[u][0] + tx * (double)nb[u][1]) +
                                          ty * ((1.0 - tx) * (double)nb[u][2] + tx * (double)nb[u][3])
                                    : 0.0;
          q_las[j * nx + i] = qlf;
          qlw = qlf;
          qn = (F)((double)qlf - (double)(F)sv);
          pb += (double)qlf;
        }
        const __bf16 hv = (__bf16)qn;
        b0.hi[i * kResLd + j] = hv;               // forward operand [i][j]
        b0.lo[i * kResLd + j] = (__bf16)(qn - (float)hv);
        QL[(((j >> 4) * NTB + (i >> 4)) * 64 + (i & 15) + 16 * ((j & 15) >> 2)) * 4 + (j & 3)] = qlw;
      }
      pb = res_wave_sum(pb);
      if (lane == 0) psl[b][wave] = pb;
    }
#ifdef SGH_RESIDENT_TIMESTAMPS
    // (diagnostic build) the flux pass's stores drained here, so that stamp 156 splits the
    // pass from the laser power sum
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#endif
    RES_TS(156);
  }
  // the band partials of the laser power, summed over the waves in order; a team's go to
  // the exchange area for member 0 (the barrier also completes the operand stores)
  __syncthreads();
  if (tid >= a_lo && tid < a_hi) {
    double v = 0.0;
    for (int w = 0; w < kResWaves; ++w) v += psl[tid][w];
    if (TP > 1) st_x<double, true>(xP + tid, v);
    else psl[tid][0] = v;
  }
  RES_TS(157);
  if (TP > 1) exchange(buf0, xQ, false, [] {});
  RES_TS(1);

  const double fwd_scale = d.dx * d.dy * d.inv_sqrt_LxLy;
  const double inv_scale = d.inv_sqrt_LxLy;
  const F fwd_scale_f = (F)fwd_scale, inv_scale_f = (F)inv_scale;

  // matrix fragments of the four stage kinds; each stage loads the next one's after its
  // last product
  ResW W;
  auto w_fwd1 = [&](int k0, int k1) __attribute__((always_inline)) {     // W[j][n] = B̂y[n][j]
    res_load_w(W, tA, m16, Y16, Y16 + my, nyp, KY, lane, k0, k1);
  };
  auto w_fwd2 = [&](int k0, int k1) __attribute__((always_inline)) {     // W[i][m] = B̂x[m][i]
    res_load_w(W, tB, m16, X16, X16 + mx, nxp, KX, lane, k0, k1);
  };
  auto w_inv1 = [&](int k0, int k1) __attribute__((always_inline)) {     // W[n][j] = B̂y[n][j]
    res_load_w(W, tA, m16, Y16 + 2 * my, Y16 + 3 * my, nyp, KY, lane, k0, k1);
  };
  auto w_inv2 = [&](int k0, int k1) __attribute__((always_inline)) {     // W[m][i] = B̂x[m][i]
    res_load_w(W, tB, m16, X16 + 2 * mx, X16 + 3 * mx, nxp, KX, lane, k0, k1);
  };
  auto w_none = [](int, int) __attribute__((always_inline)) {};

  // first stage of each transform: the result goes to the other buffer as the second
  // stage's operand
  auto stage_fwd1 = [&]() __attribute__((always_inline)) {    // R[i][n], stored [n][i]
    res_stage(tA, W, KY, buf0, lane, w_fwd2, no_pre,
              [&](int q, int, const f4v& a, int lane) __attribute__((always_inline)) {
                const float o[4] = {a[0], a[1], a[2], a[3]};
                res_store(buf1, tA.aout(q, lane), tA.tcol(lane), o);
              });
  };
  auto stage_inv1 = [&](int ts = -1) __attribute__((always_inline)) {   // V[m][j], as [j][m]
    res_stage(tA, W, KY, buf0, lane, w_inv2, no_pre,
              [&](int q, int, const f4v& a, int lane) __attribute__((always_inline)) {
                const float o[4] = {a[0], a[1], a[2], a[3]};
                res_store(buf1, tA.aout(q, lane), tA.tcol(lane), o);
              }, ts);
  };

  // ---- warm start: G = fwd·B̂y q_net B̂xᵀ; Xinv = ZE + G∘Ztt into buffer 0 ----
  // (outside the surface every epilogue input is 0: the matrices' padding is zero, and
  // loads beyond the surface return 0, so the operand padding stays 0 without selects)
  w_fwd1(0, 4);
  stage_fwd1();
  {
    const int nxk = opaque(nx), nyk = opaque(ny);
    f4v zw[2][4];   // (ZE, Ztt, Wtt, U) of the tile's elements
    res_stage(tB, W, KX, buf1, lane, w_inv1,
              [&](int q, int sl, int lane) __attribute__((always_inline)) {
                unsigned vo[4], vo16[4];
                res_voffs(tB, q, lane, nxk, nyk, vo, vo16);
#pragma unroll
                for (int v = 0; v < 4; ++v) zw[sl][v] = fld.record(vo16[v]);
              },
              [&](int q, int sl, const f4v& a, int lane) __attribute__((always_inline)) {
                unsigned vo[4], vo16[4];
                res_voffs(tB, q, lane, nxk, nyk, vo, vo16);
                float xo[4];
                f4v g4;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                  const F g = a[v] * fwd_scale_f;
                  g4[v] = g;
                  xo[v] = zw[sl][v][0] + g * zw[sl][v][1];
                }
                Gt[tidx(q, lane)] = g4;
                res_store(buf0, tB.aout(q, lane), tB.tcol(lane), xo);   // [m][n]
              });
  }
  if (TP > 1) exchange(buf0, xX, false, [] {});

  RES_TS(2);
  // ---- Picard loop ----
  int n_it = 0;
  bool conv = false;
  double rel = __builtin_nan("");
  const F omega = (F)d.omega;
  const F T_liq = (F)d.T_liq, f1 = (F)d.f1, f2 = (F)d.f2, T_boil = (F)d.T_boil;
  // below T_cut the FP32 evaporation law is exactly 0: exp's argument, in powers of two,
  // is under −160, past the smallest subnormal, so the transcendentals are skipped there
  // (f2 > 0; otherwise there is no such temperature)
  const F T_cut = d.f2 > 0.0 ? (F)(d.T_boil / (1.0 + 160.0 / (d.f2 * M_LOG2E))) : -INFINITY;
  const F T_hot = T_liq > T_cut ? T_liq : T_cut;
  const int max_it = d.max_it;
  const double tol = d.tol, n_elem = d.n_elem;
  // epilogue inputs of the two stages with epilogues (tB tiles, slot = tile parity): q_las
  // of the evaporation stage; (ZE, Ztt, Wtt, U) and G of the relaxation stage. The loads of
  // the first tiles (two of the evaporation stage, one of the relaxation stage, as registers
  // allow) start before the stage that precedes each of them, so that a wave with one or two
  // tiles (the members of a team) does not wait for them.
  f4v ql[2];
  f4v zw[2][4];
  f4v gg[2];
  auto pre_evap = [&](int q, int sl, int lane) __attribute__((always_inline)) {
    ql[sl] = QLt[tidx(q, lane)];
  };
  auto pre_relax = [&](int q, int sl, int lane) __attribute__((always_inline)) {
    unsigned vo[4], vo16[4];
    res_voffs(tB, q, lane, opaque(nx), opaque(ny), vo, vo16);
#pragma unroll
    for (int v = 0; v < 4; ++v) zw[sl][v] = fld.record(vo16[v]);
    gg[sl] = Gt[tidx(q, lane)];
  };
#pragma unroll 1
  for (int k = 0; k < max_it; ++k) {
    // inverse + evaporation: q = q_las − q_evap(T_s) into buffer 0
    if (tB.tt >= 0) {
      pre_evap(0, 0, lane);
      pre_evap(1, 1, lane);
    }
    stage_inv1(k == 5 ? 146 : -1);
    if (k < 20) RES_TS(3 + 7 * k);
    {
      const int nxk = opaque(nx), nyk = opaque(ny);
      res_stage<2>(tB, W, KX, buf1, lane, w_fwd1, pre_evap,
                [&](int q, int sl, const f4v& a, int lane) __attribute__((always_inline)) {
                  unsigned vo[4], vo16[4];
                  res_voffs(tB, q, lane, nxk, nyk, vo, vo16);
                  F T[4], qe[4] = {F(0), F(0), F(0), F(0)};
                  bool ev[4], hot = false;
#pragma unroll
                  for (int v = 0; v < 4; ++v) {
                    T[v] = a[v] * inv_scale_f;
                    // (the surface's padding, T = 0, never evaporates, even when T_liq ≤ 0)
                    ev[v] = vo[v] != kResOut && T[v] >= T_hot;
                    hot = hot || ev[v];
                  }
                  // hardware reciprocal, reciprocal square root and exponential (about 1 ulp
                  // each), only where a lane holds an evaporating element
                  if (hot) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                      const F rT = __builtin_amdgcn_rcpf(T[v]);
                      const F e = f1 * (__builtin_amdgcn_rsqf(T[v]) *
                                        __expf(f2 * (F(1) - T_boil * rT)));
                      qe[v] = ev[v] ? e : F(0);
                    }
                  }
                  float qo[4];
#pragma unroll
                  for (int v = 0; v < 4; ++v) qo[v] = ql[sl][v] - qe[v];
                  QEt[tidx(q, lane)] = f4v{qe[0], qe[1], qe[2], qe[3]};
                  res_store_x(buf0, xr, gbQ, TP > 1, tB.aout(q, lane), tB.tcol(lane),
                              qo);   // [i][j]
                });
    }
    if (k < 20) RES_TS(4 + 7 * k);
    if (TP > 1) exchange(buf0, xQ, true, [] {});
    if (k < 20) RES_TS(5 + 7 * k);
    // forward + relaxation + norm sums: per lane in FP32 over its four elements, per row
    // group of 16 lanes (4 rows × 16 columns of the tile) by DPP row shifts in FP32, to slot
    // (a·8 + t)·4 + group of tile (a, t); Xinv into buffer 0
    if (tB.tt >= 0) pre_relax(0, 0, lane);
    stage_fwd1();
    if (k < 20) RES_TS(6 + 7 * k);
    {
      res_stage<1>(tB, W, KX, buf1, lane, w_inv1, pre_relax,
                [&](int q, int sl, const f4v& a, int lane) __attribute__((always_inline)) {
                  float xo[4], p0 = 0.f, p1 = 0.f, p2 = 0.f;
                  f4v g4;
#pragma unroll
                  for (int v = 0; v < 4; ++v) {
                    const F g = gg[sl][v];
                    const F dd = a[v] * fwd_scale_f - g;
                    const F w = zw[sl][v][2];
                    p0 += dd * dd * w;
                    p1 += g * zw[sl][v][3];
                    p2 += g * g * w;
                    const F gn = g + omega * dd;
                    g4[v] = gn;
                    xo[v] = zw[sl][v][0] + gn * zw[sl][v][1];
                  }
                  Gt[tidx(q, lane)] = g4;
                  res_store_x(buf0, xr, gbX, TP > 1, tB.aout(q, lane), tB.tcol(lane),
                              xo);   // [m][n]
                  p0 = res_row_sum(p0);
                  p1 = res_row_sum(p1);
                  p2 = res_row_sum(p2);
                  if ((lane & 15) == 15) {
                    const int slot = (tB.atile(q) * 8 + tB.tt) * 4 + (lane >> 4);
                    if (TP == 1) {
                      nsl[0][slot] = p0;
                      nsl[1][slot] = p1;
                      nsl[2][slot] = p2;
                    } else {
                      st_x<float, true>(xN + slot, p0);
                      st_x<float, true>(xN + kResNormSlots + slot, p1);
                      st_x<float, true>(xN + 2 * kResNormSlots + slot, p2);
                    }
                  }
                }, k == 5 ? 150 : -1);
    }
    if (k < 20) RES_TS(7 + 7 * k);
    // the three sums over all tiles of the surface: lane a·8 + t sums the four row groups of
    // tile (a, t) in FP64, then the lanes in a fixed order (every wave, every member); a
    // team's per-tile sums load from the exchange area while the X field is copied
    const bool tin = (lane >> 3) < NAY && (lane & 7) < NTB;
    f4v r0 = {}, r1 = {}, r2 = {};
    if (TP > 1)
      exchange(buf0, xX, true, [&] {
        if (tin) {
          const int nb = (int)kResXchNorm + 16 * lane;
          r0 = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(xr, nb, 0,
                                                                            kCoherentAux));
          r1 = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(
                                           xr, nb + 4 * kResNormSlots, 0, kCoherentAux));
          r2 = __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(
                                           xr, nb + 8 * kResNormSlots, 0, kCoherentAux));
        }
      });
    if (k < 20) RES_TS(8 + 7 * k);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (tin) {
      if (TP == 1) {
        r0 = *reinterpret_cast<const f4v*>(&nsl[0][4 * lane]);
        r1 = *reinterpret_cast<const f4v*>(&nsl[1][4 * lane]);
        r2 = *reinterpret_cast<const f4v*>(&nsl[2][4 * lane]);
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        s0 += (double)r0[g];
        s1 += (double)r1[g];
        s2 += (double)r2[g];
      }
    }
    const double rr = res_wave_sum(s0);
    const double gu = res_wave_sum(s1);
    const double g2w = res_wave_sum(s2);
    const double a2 = AA + 2.0 * gu + g2w;
    const double rms_diff = sqrt(rr / n_elem);
    const double rms_old = sqrt(a2 / n_elem);
    rel = rms_diff / (rms_old > 1e-9 ? rms_old : 1e-9);
    conv = rel < tol;
    n_it = k + 1;
    if (k < 20) RES_TS(9 + 7 * k);
    if (conv) break;
  }

  // ---- step end: T_s of the final iterate and its maximum; G and the last iteration's
  // q_evap from the work slot into the surface arrays for the volume pass ----
  stage_inv1();
  double tmax = -1.0 / 0.0;
  res_stage(tB, W, KX, buf1, lane, w_none, no_pre,
            [&](int q, int, const f4v& a, int lane) __attribute__((always_inline)) {
              unsigned vo[4], vo16[4];
              res_voffs(tB, q, lane, nx, ny, vo, vo16);
              const f4v g4 = Gt[tidx(q, lane)], qe4 = QEt[tidx(q, lane)];
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                fld.store(0, vo[v], g4[v]);
                fld.store(7, vo[v], qe4[v]);
              }
#pragma unroll
              for (int v = 0; v < 4; ++v)
                if (vo[v] != kResOut) {
                  const double T = (double)(a[v] * inv_scale_f);
                  tmax = T > tmax ? T : tmax;
                }
            });
  tmax = wave_max(tmax);
  if (lane == 0) sh[0][wave] = tmax;
  __syncthreads();
  RES_TS(143);
#ifdef SGH_RESIDENT_TIMESTAMPS
  if (tid == 0 && blockIdx.x < kResTsBlocks) {
    g_res_ts[blockIdx.x * kResTsSlots + 144] = (unsigned long long)(nx | (TP << 16));
    g_res_ts[blockIdx.x * kResTsSlots + 145] = (unsigned long long)(ny | (mp << 16));
  }
#endif
  double tm_all = -1.0 / 0.0;
  if (tid == 0)
    for (int w = 0; w < kResWaves; ++w) tm_all = sh[0][w] > tm_all ? sh[0][w] : tm_all;
  if (TP > 1) {
    // every member's maximum to its slot; member 0 waits for all and finishes the step
    if (tid == 0) st_x<double, true>(xT + mp, tm_all);
    arrivals += TP;
    cuboid_barrier(P.bar + c, arrivals, P.loop_err, ticks, mp == 0);
    if (mp != 0) return;
    if (tid == 0)
      for (int m = 0; m < TP; ++m) {
        const double t = ld_x<double, true>(xT + m);
        tm_all = t > tm_all ? t : tm_all;
      }
  }
  if (tid == 0) {
    // the laser power: band partials in band order, the same for every team size (a single
    // member's own partials are in psl, written before the preparation's barrier)
    double P_sum = 0.0;
    for (int b = 0; b < NAY; ++b) P_sum += TP > 1 ? ld_x<double, true>(xP + b) : psl[b][0];
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
                           unsigned nblocks, unsigned long long ticks, hipStream_t st) {
  hipLaunchKernelGGL(k_b_resident, dim3(nblocks), dim3(kResThreads), kResLdsBytes, st, P,
                     (const __bf16*)m16, list, ticks);
  return hipGetLastError();
}

hipError_t resident_allow_lds() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_b_resident),
                             hipFuncAttributeMaxDynamicSharedMemorySize, kResLdsBytes);
}

}  // namespace sgh

#ifdef SGH_RESIDENT_TIMESTAMPS
// Copies the stamps of the last resident launch (kResTsBlocks × kResTsSlots values).
extern "C" int sgh_debug_resident_timestamps(unsigned long long* host_out) {
  return hipMemcpyFromSymbol(host_out, HIP_SYMBOL(sgh::g_res_ts),
                             sizeof(unsigned long long) * sgh::kResTsBlocks * sgh::kResTsSlots) ==
                 hipSuccess
             ? 0
             : -2;
}
#endif
