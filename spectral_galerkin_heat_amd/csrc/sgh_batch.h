// Definitions shared by the batched engine's translation units: the per-cuboid descriptor,
// work items, the packed-array pointer block, and the resident step kernel's launcher.
#pragma once

#include <hip/hip_runtime.h>

#include "sgh_common.h"
#include "sgh_mfma.h"

namespace sgh {

constexpr int kBT = 256;        // threads per block
constexpr int kRowsMin = 16;    // fewest surface rows per transform item (one MFMA row tile)
constexpr int kBMaxN = 256;     // widest surface axis
constexpr int kStatusStride = 4;  // n_it, converged, rel, AA of the current step
constexpr int kMetricStride = 5;  // P_laser, T_max, n_it, converged, rel
constexpr int kSurfArrays = 10;
constexpr int kStageAhead = 4;   // global loads in flight per thread while staging into LDS
constexpr int kSurfLds = 150 * 1024;  // default LDS bytes per surface workgroup
constexpr int kLdsMax = 159 * 1024;   // dynamic LDS cap: gfx950's 160 KiB less the static part
constexpr int kLoopLdsMax = 158 * 1024;  // the same for the persistent step kernel (four passes'
                                         // static reduction buffers)
constexpr int kXcds = 8;          // MI355X: workgroup b runs on XCD b mod 8
constexpr int kCusPerXcd = 32;    // MI355X: 256 CUs

struct CuboidDesc {
  int nx, ny, nz, S;
  long long a_off;
  int s_off;
  int bx_off, by_off;      // offsets of B̂ (n×n) and B̂ᵀ for x and y in the matrix table
  int bx16_off, by16_off;  // offsets of the split-bf16 tables for x and y (−1: none)
  int bxT_off, byT_off;
  int t1_off;              // offset of the 1-D tables (ex, kx2 | ey, ky2 | ez, kz2)
  int row_item0, nrowblk;  // first per-cuboid transform slot and block count
  int col_item0, ncolblk;  // first per-cuboid volume slot and block count
  int max_it;
  int n_steps;             // laser-table length
  long long laser_off;     // row offset into the laser table (6 doubles per row)
  double dx, dy, inv_sqrt_LxLy, Lz;
  double alpha, rho_cp, dt;
  double w0, w1, inv_arc;  // √(1/Lz), √(2/Lz), 1/(α·ρc)
  double omega, tol, n_elem;
  double f1, f2, T_boil, T_liq;
  double absorptivity, rx, ry, order, area_factor;
  int cell_integrated, live;  // live: the slot holds a cuboid
  int item_rows, ncb;         // surface rows per transform item (16 or 32); column blocks
};

// Wide surfaces split their transform items by output columns too (ncb blocks of cb_width
// columns, a multiple of 16): each item still forms all of Y = L[rows,:]·X, but stages and
// multiplies only its own columns of R, so a single large cuboid spreads over ncb times as
// many workgroups. Item blk covers row block blk / ncb and column block blk % ncb.
__host__ __device__ inline int cb_width(const CuboidDesc& d) {
  const int nb = d.ncb > 1 ? d.ncb : 1;
  return ((d.nx + nb - 1) / nb + 15) & ~15;
}
__host__ __device__ inline int n_items(const CuboidDesc& d) { return d.nrowblk * d.ncb; }

struct Item {
  int c, blk;
};

// A surface work item with a copy of its cuboid's descriptor, so that a surface
// workgroup finds everything it needs in one load (no item -> descriptor chain).
struct ItemDesc {
  CuboidDesc d;
  int c, blk;
};

template <typename F>
struct BatchPtrs {
  F* a;            // packed modal state
  F* G;            // surface forcing estimate (modal)
  F* ZE;           // Σ w·aE (modal)
  F* U;            // Σ aE·KK·w (modal)
  F* Ztt;          // Σ w²·KK
  F* Wtt;          // Σ w²·KK²
  F* q_las;        // laser flux (physical)
  F* q_net;        // net top flux (physical)
  F* q_evap;       // evaporation flux of the current iterate (physical)
  F* q_evap_old;   // previous step's evaporation flux
  F* Xinv;         // ZE + G∘Ztt: modal input of the inverse surface transform
  const F* mats;   // B̂ and B̂ᵀ matrices
  const double* t1;        // 1-D tables
  const double* laser;     // laser rows
  const CuboidDesc* desc;
  const Item* row_items;
  const ItemDesc* row_descs;  // row_items[i] with its cuboid's descriptor
  const Item* col_items;
  double* row_slots;       // per row item: rr, gu, g2w  (3 doubles)
  double* col_slots;       // per col item: AA
  double* aux_slots;       // per row item: P_laser, T_max (2 doubles)
  int* done_at;            // per cuboid: iteration at which the step finished, −1 active
  double* status;          // per cuboid: n_it, converged, rel
  int* step_idx;           // per cuboid: step counter
  double* metrics;         // [step][cuboid][5]
  unsigned* bar;           // per cuboid: barrier arrivals of the persistent step kernel
  int* loop_err;           // set when a barrier of the persistent step kernel timed out
  long long surf_stride;   // elements from one surface array to the next (G, ZE, U, Ztt, …)
  float* zwu;              // FP32 pools: (ZE, Ztt, Wtt, U) of each surface element, packed
                           // for the resident step kernel; null otherwise
  float* res_work;         // FP32 pools: per cuboid slot, kResWork floats of the resident
                           // step kernel's tiled G, q_evap and q_las; null otherwise
  unsigned char* res_xch;  // FP32 pools: per cuboid slot, kResXchBytes through which the
                           // workgroups of a resident team exchange fields and sums
  int metrics_rows;        // capacity in steps
  int n_cuboids;
};

// Resident step kernel (sgh_resident.hip): a team of 1 to kResTeamMax 512-thread
// workgroups per cuboid, mixed precision, surfaces up to kResMaxN × kResMaxN.
constexpr int kResMaxN = 128;            // widest surface axis
constexpr int kResLd = kResMaxN + 8;     // bf16 per LDS operand row: 16-byte rows, staggered banks
constexpr int kResThreads = 512;
constexpr int kResLdsBytes = 4 * kResMaxN * kResLd * 2;   // two hi / lo operand buffer pairs
constexpr int kResWork = 3 * kResMaxN * kResMaxN;          // work floats per cuboid slot
constexpr int kResTeamMax = 8;           // workgroups per cuboid (one per 16-row band at most)
constexpr int kWavesPerRes = kResThreads / 64;   // waves of a resident workgroup
constexpr int kResTiles = 64;            // 16×16 tiles of a kResMaxN² surface
constexpr int kResNormSlots = 4 * kResTiles;   // norm sums: one per 4-row group of a tile
// Exchange area of a cuboid slot: the X and q operand fields (hi, lo; LDS layout), the
// norm sums (3 × kResNormSlots floats), one T_max per member and the laser power of every
// 16-row band (doubles), 256-byte aligned.
constexpr long long kResXchField = 2LL * kResMaxN * kResLd;   // bf16 elements of a hi+lo field
constexpr long long kResXchNorm = 2 * kResXchField * 2;        // byte offset of the norm sums
constexpr long long kResXchBytes =
    (kResXchNorm + 4LL * 3 * kResNormSlots + 8LL * kResTeamMax + 8LL * (kResMaxN / 16) + 255) /
    256 * 256;

__host__ __device__ inline int pad32(int n) { return (n + 31) & ~31; }

// Work-list entry of member p of a team of tp workgroups for cuboid c (−1: idle).
__host__ __device__ inline int res_entry(int c, int tp, int p) { return c * 64 + (tp - 1) * 8 + p; }

// Barrier of the workgroups of one cuboid (the items of the persistent step kernel, the
// members of a resident team). What they exchange is written and read device-coherently
// (sc1 stores and loads, LDS DMA with sc1), so no cache flush is needed: every wave waits
// until its stores have completed, the workgroup meets, and one thread adds its arrival to
// the cuboid's counter (a device-scope atomic) and, unless `wait` is false, polls it with
// device-scope loads until `target` arrivals. (Device-scope release/acquire fences
// instead write back and invalidate the XCD's L2 at every barrier: measured 2.5× slower.)
// The wait is bounded: past `ticks` of the 100 MHz wall clock (or once another barrier has
// timed out) it records the failure in *err and goes on, so the kernel always ends; the
// host reports it.
constexpr unsigned long long kBarrierTicks = 2000000ULL;   // default: 20 ms
__device__ __forceinline__ void cuboid_barrier(unsigned* cnt, unsigned target, int* err,
                                               unsigned long long ticks, bool wait = true) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t0 = wall_clock64();
    while (wait && __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (wall_clock64() - t0 > ticks) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
}

// Enqueue the resident step kernel over `nblocks` entries of `list` (res_entry per
// workgroup, −1 idle); m16 holds the split-bf16 cosine tables the descriptors point into;
// `ticks` bounds a team barrier's wait.
hipError_t launch_resident(const BatchPtrs<float>& P, const void* m16, const int* list,
                           unsigned nblocks, unsigned long long ticks, hipStream_t st);
// Allow the resident kernel its dynamic LDS (once per process is enough).
hipError_t resident_allow_lds();

}  // namespace sgh
