// Pointwise, reduction and forcing kernels of the spectral ETD1 + Picard step.
//
// Memory layout: modal volumes are C-contiguous [p, n, m] = [z, y, x] with x fastest,
// surfaces [n, m]. A "column" is the nz values sharing one surface index s = n·nx + m;
// kernels that need a z-contraction assign one thread per column (coalesced along x)
// and split z into chunks across blockIdx.y so small surfaces still fill the GPU.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <mutex>
#include <string>

#include "sgh_common.h"
#include "sgh_props.h"

namespace sgh {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

constexpr int kThreads = 256;
static_assert(kThreads == kProgThreads, "property kernels index the LDS evaluation stack by thread");

// ---------------------------------------------------------------------------------
// Propagator tables (FP64 math, expm1 near λ → 0)
// ---------------------------------------------------------------------------------
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_etd_tables(F* __restrict__ E, F* __restrict__ KK, int nx, int ny, int nz, double px,
             double py, double pz, double alpha, double rho_cp, double dt) {
  const long long N = (long long)nx * ny * nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(i % nx);
    const int n = (int)((i / nx) % ny);
    const int p = (int)(i / ((long long)nx * ny));
    const double kz = pz * p, ky = py * n, kx = px * m;
    const double denom = alpha * (kz * kz + ky * ky + kx * kx);
    const double z = denom * dt;
    E[i] = (F)exp(-z);
    KK[i] = (F)(denom == 0.0 ? dt / rho_cp : (-expm1(-z) / denom) / rho_cp);
  }
}

// ---------------------------------------------------------------------------------
// Modal pointwise updates
// ---------------------------------------------------------------------------------
// Volume kernels index in I: unsigned (32-bit division) when the volume has fewer than
// 2^31 elements, long long otherwise (64-bit division costs several times as much).
template <typename F, typename I>
__global__ void __launch_bounds__(kThreads)
k_etd1_update(const F* __restrict__ aK, const F* __restrict__ KK, const F* __restrict__ wt,
              const F* __restrict__ St, const F* __restrict__ wb, const F* __restrict__ Sb,
              const F* __restrict__ Q, F* __restrict__ out, int S, I N) {
  for (I i = blockIdx.x * (I)blockDim.x + threadIdx.x; i < N; i += (I)gridDim.x * blockDim.x) {
    const int p = (int)(i / (I)S);
    const int s = (int)(i - (I)p * S);
    F f = wt[p] * St[s];
    if (wb) f += wb[p] * Sb[s];
    if (Q) f += Q[i];
    out[i] = aK[i] + KK[i] * f;
  }
}

template <typename F>
__global__ void __launch_bounds__(kThreads)
k_add_volume(F* __restrict__ a, const F* __restrict__ KK, const F* __restrict__ Q, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x)
    a[i] += KK[i] * Q[i];
}

template <typename F, typename I>
__global__ void __launch_bounds__(kThreads)
k_add_surface(F* __restrict__ a, const F* __restrict__ KK, const F* __restrict__ w,
              const F* __restrict__ B, int S, I N) {
  for (I i = blockIdx.x * (I)blockDim.x + threadIdx.x; i < N; i += (I)gridDim.x * blockDim.x) {
    const int p = (int)(i / (I)S);
    a[i] += KK[i] * w[p] * B[i - (I)p * S];
  }
}

template <typename F>
__global__ void __launch_bounds__(kThreads)
k_mul_inplace(F* __restrict__ a, const F* __restrict__ E, long long N) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N;
       i += (long long)gridDim.x * blockDim.x)
    a[i] *= E[i];
}

// parts[c][s] = Σ_{p in chunk c} w[p]·a[p·S + s]
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_zcontract(const F* __restrict__ a, const F* __restrict__ w, F* __restrict__ parts, int S,
            int nz, int pc) {
  const int s = blockIdx.x * kThreads + threadIdx.x;
  if (s >= S) return;
  const int p0 = blockIdx.y * pc;
  const int p1 = min(nz, p0 + pc);
  F acc = F(0);
  for (int p = p0; p < p1; ++p) acc += w[p] * a[(long long)p * S + s];
  parts[(long long)blockIdx.y * S + s] = acc;
}

// parts[0][i] = Σ_q parts[q][i], in place (the other planes are left as they were).
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_sum_parts(F* __restrict__ parts, long long n, int nparts) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kThreads) {
    F v = parts[i];
    for (int q = 1; q < nparts; ++q) v += parts[q * n + i];
    parts[i] = v;
  }
}

// One Picard iteration, column-parallel:
//   a_raw = aK + KK·(wt⊗St + wb⊗Sb + Q); r = a_raw − a_old; a_new = a_old + ω r
// with FP64 partial sums of r² and a_old² per block and the next iterate's z-contraction.
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_picard_fused(const F* __restrict__ aK, const F* __restrict__ KK, const F* __restrict__ a_old,
               const F* __restrict__ wt, const F* __restrict__ St, const F* __restrict__ wb,
               const F* __restrict__ Sb, const F* __restrict__ Q, F omega,
               const F* __restrict__ wz, F* __restrict__ zc_parts, F* __restrict__ a_new,
               double* __restrict__ partials, int S, int nz, int pc,
               unsigned int* __restrict__ counter, double n_elements,
               double* __restrict__ norms, double* __restrict__ host_out, double seq) {
  __shared__ double sh[kThreads / kWave];
  __shared__ int s_last;
  const int s = blockIdx.x * kThreads + threadIdx.x;
  const int p0 = blockIdx.y * pc;
  const int p1 = min(nz, p0 + pc);
  double rr = 0.0, aa = 0.0;
  if (s < S) {
    const F st = St[s];
    const F sb = wb ? Sb[s] : F(0);
    F zc = F(0);
    for (int p = p0; p < p1; ++p) {
      const long long i = (long long)p * S + s;
      F f = wt[p] * st;
      if (wb) f += wb[p] * sb;
      if (Q) f += Q[i];
      const F old = a_old[i];
      const F r = (aK[i] + KK[i] * f) - old;
      const F nw = old + omega * r;
      a_new[i] = nw;
      rr += (double)r * (double)r;
      aa += (double)old * (double)old;
      if (wz) zc += wz[p] * nw;
    }
    if (wz) zc_parts[(long long)blockIdx.y * S + s] = zc;
  }
  const double trr = block_sum(rr, sh);
  const double taa = block_sum(aa, sh);
  if (norms == nullptr) {
    if (threadIdx.x == 0) {
      const long long b = (long long)blockIdx.y * gridDim.x + blockIdx.x;
      partials[2 * b] = trr;
      partials[2 * b + 1] = taa;
    }
    return;
  }
  // The last block to finish sums every block's pair in the fixed order of
  // k_reduce_norms and resets the counter for the next launch. The pairs and the counter
  // are accessed with device-scope atomics (coherent past the L1 and across the XCDs'
  // L2s), and a block's pair has completed (s_waitcnt) before the block is counted, so no
  // fence is needed: a device-scope fence writes back and invalidates the L2, which cost
  // more than the rest of the pass at these sizes.
  if (threadIdx.x == 0) {
    const long long b = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    __hip_atomic_store(partials + 2 * b, trr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(partials + 2 * b + 1, taa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    const unsigned int nb = gridDim.x * gridDim.y;
    s_last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
             nb - 1;
  }
  __syncthreads();
  if (!s_last) return;
  const int n_pairs = gridDim.x * gridDim.y;
  double r2 = 0.0, a2 = 0.0;
  for (int i = threadIdx.x; i < n_pairs; i += kThreads) {
    r2 += __hip_atomic_load(partials + 2 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    a2 += __hip_atomic_load(partials + 2 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const double tr = block_sum(r2, sh);
  const double ta = block_sum(a2, sh);
  if (threadIdx.x == 0) {
    const double rms_diff = sqrt(tr / n_elements);
    const double rms_old = sqrt(ta / n_elements);
    const double rel = rms_diff / (rms_old > 1e-9 ? rms_old : 1e-9);
    norms[0] = rms_diff;
    norms[1] = rms_old;
    norms[2] = rel;
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (host_out) {
      // straight to pinned host memory: the three norms, then (after a system-scope
      // fence) the sequence number the host polls for
      host_out[0] = rms_diff;
      host_out[1] = rms_old;
      host_out[2] = rel;
      __threadfence_system();
      host_out[3] = seq;
    }
  }
}

__global__ void __launch_bounds__(kThreads)
k_reduce_norms(const double* __restrict__ partials, int n_pairs, double n_elements,
               double* __restrict__ out) {
  __shared__ double sh[kThreads / kWave];
  double rr = 0.0, aa = 0.0;
  for (int i = threadIdx.x; i < n_pairs; i += kThreads) {
    rr += partials[2 * i];
    aa += partials[2 * i + 1];
  }
  const double trr = block_sum(rr, sh);
  const double taa = block_sum(aa, sh);
  if (threadIdx.x == 0) {
    const double rms_diff = sqrt(trr / n_elements);
    const double rms_old = sqrt(taa / n_elements);
    out[0] = rms_diff;
    out[1] = rms_old;
    out[2] = rms_diff / (rms_old > 1e-9 ? rms_old : 1e-9);
  }
}

template <typename F>
__global__ void __launch_bounds__(1024)
k_reduce(const F* __restrict__ x, long long n, int op, double* __restrict__ out) {
  __shared__ double sh[1024 / kWave];
  double acc = op == 0 ? 0.0 : -1.0 / 0.0;
  for (long long i = threadIdx.x; i < n; i += 1024) {
    const double v = (double)x[i];
    acc = op == 0 ? acc + v : (v > acc ? v : acc);
  }
  const double t = op == 0 ? block_sum(acc, sh) : block_max(acc, sh);
  if (threadIdx.x == 0) out[0] = t;
}

// ---------------------------------------------------------------------------------
// Surface forcing
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double derf(double x) { return erf(x); }
__device__ __forceinline__ float derf(float x) { return erff(x); }

template <typename F>
__global__ void __launch_bounds__(kThreads)
k_laser_flux(F* __restrict__ out, int nx, int ny, F dx, F dy, F x0, F y0, F rx, F ry,
             F order, F I0, int cell_integrated) {
  const int S = nx * ny;
  for (int s = blockIdx.x * kThreads + threadIdx.x; s < S; s += gridDim.x * kThreads) {
    const int j = s / nx, i = s - j * nx;
    const F x = ((F)i + F(0.5)) * dx;
    const F y = ((F)j + F(0.5)) * dy;
    F q;
    if (cell_integrated) {
      const F s2 = dsqrt(F(2));
      const F c = dsqrt(F(M_PI) / F(8));
      const F ax = (x - x0 - dx / F(2)) * s2 / rx, bx = (x - x0 + dx / F(2)) * s2 / rx;
      const F ay = (y - y0 - dy / F(2)) * s2 / ry, by = (y - y0 + dy / F(2)) * s2 / ry;
      const F mx = rx * c * (derf(bx) - derf(ax)) / dx;
      const F my = ry * c * (derf(by) - derf(ay)) / dy;
      q = I0 * (my * mx);
    } else {
      const F ux = (x - x0) / rx, uy = (y - y0) / ry;
      const F r2 = ux * ux + uy * uy;
      const F shape = order == F(2) ? dexp(F(-2) * r2) : dexp(F(-2) * mpow(r2, order / F(2)));
      q = I0 * shape;
    }
    out[s] = q;
  }
}

// Bilinear translate with scipy's order-1 'constant' semantics: out[o] = in(o − s).
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_shift2d(const F* __restrict__ in, int ny, int nx, double sy, double sx,
          const F* __restrict__ base, F alpha, F* __restrict__ out) {
  const int S = nx * ny;
  for (int s = blockIdx.x * kThreads + threadIdx.x; s < S; s += gridDim.x * kThreads) {
    const int oy = s / nx, ox = s - oy * nx;
    const double py = oy - sy, px = ox - sx;
    const double fy = floor(py), fx = floor(px);
    const F ty = (F)(py - fy), tx = (F)(px - fx);
    const int iy0 = (int)fy, ix0 = (int)fx;
    const bool inr = (py >= 0.0) && (py <= ny - 1) && (px >= 0.0) && (px <= nx - 1);
    const int a = min(max(iy0, 0), ny - 1), b = min(max(iy0 + 1, 0), ny - 1);
    const int c = min(max(ix0, 0), nx - 1), d = min(max(ix0 + 1, 0), nx - 1);
    const F w0y = F(1) - ty, w0x = F(1) - tx;
    F val = w0y * (w0x * in[a * nx + c] + tx * in[a * nx + d]) +
            ty * (w0x * in[b * nx + c] + tx * in[b * nx + d]);
    if (!inr) val = F(0);
    out[s] = (base ? base[s] : F(0)) + alpha * val;
  }
}

// The same translate on every z slice of a box: 'nearest' clamps the source to the
// slice (edge value), otherwise outside the slice is zero.  One thread per output
// element; z is blockIdx.y so each block stays within a slice.
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_shift3d_xy(const F* __restrict__ in, int ny, int nx, double sy, double sx, int nearest,
             F* __restrict__ out) {
  const int S = nx * ny;
  const long long zoff = (long long)blockIdx.y * S;
  const F* src = in + zoff;
  for (int s = blockIdx.x * kThreads + threadIdx.x; s < S; s += gridDim.x * kThreads) {
    const int oy = s / nx, ox = s - oy * nx;
    const double py = oy - sy, px = ox - sx;
    const double fy = floor(py), fx = floor(px);
    const F ty = (F)(py - fy), tx = (F)(px - fx);
    const int iy0 = (int)fy, ix0 = (int)fx;
    const bool inr = (py >= 0.0) && (py <= ny - 1) && (px >= 0.0) && (px <= nx - 1);
    const int a = min(max(iy0, 0), ny - 1), b = min(max(iy0 + 1, 0), ny - 1);
    const int c = min(max(ix0, 0), nx - 1), d = min(max(ix0 + 1, 0), nx - 1);
    const F w0y = F(1) - ty, w0x = F(1) - tx;
    F val = w0y * (w0x * src[a * nx + c] + tx * src[a * nx + d]) +
            ty * (w0x * src[b * nx + c] + tx * src[b * nx + d]);
    if (!nearest && !inr) val = F(0);
    out[zoff + s] = val;
  }
}

// max |Q| over the box, its floor (z = 0), its x faces and its y faces: one block per
// z slice writes that slice's four maxima to part[4z..4z+3]; k_face_peaks_final folds them.
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_box_face_peaks(const F* __restrict__ Q, int ny, int nx, double* __restrict__ part) {
  __shared__ double sh[kThreads / kWave];
  const int z = blockIdx.x;
  const F* q = Q + (long long)z * nx * ny;
  double all = 0.0, xf = 0.0, yf = 0.0;
  for (int j = threadIdx.x / 64; j < ny; j += kThreads / 64) {
    const bool yface = j == 0 || j == ny - 1;
    for (int i = threadIdx.x % 64; i < nx; i += 64) {
      const double v = fabs((double)q[j * nx + i]);
      all = v > all ? v : all;
      if (i == 0 || i == nx - 1) xf = v > xf ? v : xf;
      if (yface) yf = v > yf ? v : yf;
    }
  }
  const double r0 = block_max(all, sh);
  const double r2 = block_max(xf, sh);
  const double r3 = block_max(yf, sh);
  if (threadIdx.x == 0) {
    part[4 * z + 0] = r0;
    part[4 * z + 1] = z == 0 ? r0 : 0.0;
    part[4 * z + 2] = r2;
    part[4 * z + 3] = r3;
  }
}

__global__ void __launch_bounds__(kThreads)
k_face_peaks_final(const double* __restrict__ part, int nz, double* __restrict__ out) {
  __shared__ double sh[kThreads / kWave];
  double m[4] = {0.0, 0.0, 0.0, 0.0};
  for (int z = threadIdx.x; z < nz; z += kThreads)
#pragma unroll
    for (int c = 0; c < 4; ++c) m[c] = part[4 * z + c] > m[c] ? part[4 * z + c] : m[c];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double r = block_max(m[c], sh);
    if (threadIdx.x == 0) out[c] = r;
  }
}

// out[i] = Σ_p Bz[p][i] · Σ_n By[n][i] · U[p][n][i]: the y and z sums of a point
// evaluation after the x contraction U = a ·ₓ Bx (one thread per point).
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_points_yz(const F* __restrict__ U, const F* __restrict__ By, const F* __restrict__ Bz,
            int nz, int ny, int P, F* __restrict__ out) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= P) return;
  F t = F(0);
  for (int p = 0; p < nz; ++p) {
    F u = F(0);
    const F* Up = U + (long long)p * ny * P + i;
    for (int n = 0; n < ny; ++n) u += By[(long long)n * P + i] * Up[(long long)n * P];
    t += Bz[(long long)p * P + i] * u;
  }
  out[i] = t;
}

template <typename F>
__global__ void __launch_bounds__(kThreads)
k_surface_forcing(const F* __restrict__ T, const F* __restrict__ q_las, long long S,
                  ForcingParams<F> fp, F* __restrict__ q_evap, F* __restrict__ q_net) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < S;
       s += (long long)gridDim.x * kThreads) {
    F qe, qn;
    surface_forcing_value<F>(fp, T[s], q_las ? q_las[s] : F(0), qe, qn);
    if (q_evap) q_evap[s] = qe;
    q_net[s] = qn;
  }
}

template <typename F>
__global__ void __launch_bounds__(kThreads)
k_convection(const F* __restrict__ T, long long S, F h, F T0, const PropProgram* __restrict__ prog,
             F kbar, F* __restrict__ out) {
  for (long long s = blockIdx.x * (long long)kThreads + threadIdx.x; s < S;
       s += (long long)gridDim.x * kThreads) {
    const F t = T[s];
    F q = (-h) * (t - T0);
    if (prog) {
      const F fl = prog_fl<F>(prog, t);
      q = q * (kbar / prog_blend<F>(prog, kKSolid, t, fl));
    }
    out[s] = q;
  }
}

// ---------------------------------------------------------------------------------
// Volume forcing: latent heat and the property-correction stencil
// ---------------------------------------------------------------------------------
template <typename F>
__global__ void __launch_bounds__(kThreads)
k_latent(const F* __restrict__ T, const F* __restrict__ Tp, long long N, LatentParams<F> lp,
         F* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < N;
       i += (long long)gridDim.x * kThreads)
    out[i] = latent_value<F>(lp, T[i], Tp[i]);
}

// numpy.gradient (edge_order=1) along one axis of stride `st` and length `n` at index c.
template <typename F, typename I>
__device__ __forceinline__ F grad1(const F* __restrict__ f, I i, int c, int n, I st, F inv_d) {
  if (n < 2) return F(0);
  if (c == 0) return (f[i + st] - f[i]) * inv_d;
  if (c == n - 1) return (f[i] - f[i - st]) * inv_d;
  return (f[i + st] - f[i - st]) * (F(0.5) * inv_d);
}

// (x, y, z) of flat index i of an nx × ny × nz volume
template <typename I>
__device__ __forceinline__ void xyz_of(I i, int nx, int ny, int& x, int& y, int& z) {
  const I r = i / (I)nx;
  const I q = r / (I)ny;
  x = (int)(i - r * (I)nx);
  y = (int)(r - q * (I)ny);
  z = (int)q;
}

template <typename F, typename I>
__global__ void __launch_bounds__(kThreads)
k_corr_grad(const F* __restrict__ T, const PropProgram* __restrict__ prog, F kbar, int nx,
            int ny, int nz, F idx, F idy, F idz, F* __restrict__ gx, F* __restrict__ gy,
            F* __restrict__ gz) {
  const I S = (I)nx * ny, N = S * nz;
  for (I i = blockIdx.x * (I)kThreads + threadIdx.x; i < N; i += (I)gridDim.x * kThreads) {
    int x, y, z;
    xyz_of<I>(i, nx, ny, x, y, z);
    const F t = T[i];
    const F kp = prog_blend<F>(prog, kKSolid, t, prog_fl<F>(prog, t)) - kbar;
    gx[i] = kp * grad1<F, I>(T, i, x, nx, 1, idx);
    gy[i] = kp * grad1<F, I>(T, i, y, ny, (I)nx, idy);
    gz[i] = kp * grad1<F, I>(T, i, z, nz, S, idz);
  }
}

template <typename F, typename I>
__global__ void __launch_bounds__(kThreads)
k_corr_div(const F* __restrict__ gx, const F* __restrict__ gy, const F* __restrict__ gz,
           const F* __restrict__ T, const F* __restrict__ Tp,
           const PropProgram* __restrict__ prog, F abar, int nx, int ny, int nz, F idx, F idy,
           F idz, F idt, F* __restrict__ out) {
  const I S = (I)nx * ny, N = S * nz;
  for (I i = blockIdx.x * (I)kThreads + threadIdx.x; i < N; i += (I)gridDim.x * kThreads) {
    int x, y, z;
    xyz_of<I>(i, nx, ny, x, y, z);
    const F t = T[i];
    const F fl = prog_fl<F>(prog, t);
    const F ap = prog_blend<F>(prog, kRhoSolid, t, fl) * prog_blend<F>(prog, kCSolid, t, fl) - abar;
    const F div = grad1<F, I>(gx, i, x, nx, 1, idx) + grad1<F, I>(gy, i, y, ny, (I)nx, idy) +
                  grad1<F, I>(gz, i, z, nz, S, idz);
    out[i] = div - ap * ((t - Tp[i]) * idt);
  }
}

inline unsigned grid_for(long long n, int threads = kThreads) {
  long long b = (n + threads - 1) / threads;
  if (b > 65535LL * 4) b = 65535LL * 4;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace sgh

using namespace sgh;

#define STREAM(s) reinterpret_cast<hipStream_t>(s)

extern "C" {

int sgh_version(void) { return 100; }

int sgh_host_alloc(size_t bytes, void** out) {
  SGH_REQUIRE(bytes > 0 && out, SGH_ERR_ARG, "sgh_host_alloc: bad args");
  hipError_t e = hipHostMalloc(out, bytes, hipHostMallocCoherent);
  if (e != hipSuccess) return fail_hip(e, "sgh_host_alloc");
  return SGH_OK;
}

int sgh_host_free(void* p) {
  if (p) {
    hipError_t e = hipHostFree(p);
    if (e != hipSuccess) return fail_hip(e, "sgh_host_free");
  }
  return SGH_OK;
}

const char* sgh_last_error(void) { return g_last_error.c_str(); }

int sgh_property_program_bytes(void) { return (int)sizeof(PropProgram); }

int sgh_etd_tables(int dtype, int nx, int ny, int nz, double Lx, double Ly, double Lz,
                   double alpha, double rho_cp, double dt, void* E, void* KK, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && E && KK, SGH_ERR_ARG, "sgh_etd_tables: bad args");
  SGH_REQUIRE(Lx > 0 && Ly > 0 && Lz > 0 && rho_cp > 0 && alpha >= 0, SGH_ERR_ARG,
              "sgh_etd_tables: non-positive extent or material constant");
  const long long N = (long long)nx * ny * nz;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_etd_tables<F>, dim3(grid_for(N)), dim3(kThreads), 0, STREAM(stream),
                       (F*)E, (F*)KK, nx, ny, nz, M_PI / Lx, M_PI / Ly, M_PI / Lz, alpha,
                       rho_cp, dt);
  });
  SGH_CHECK_LAUNCH("sgh_etd_tables");
  return SGH_OK;
}

int sgh_etd1_update(int dtype, int nx, int ny, int nz, const void* aK, const void* KK,
                    const void* wt, const void* St, const void* wb, const void* Sb,
                    const void* Q, void* out, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && aK && KK && wt && St && out, SGH_ERR_ARG,
              "sgh_etd1_update: bad args");
  SGH_REQUIRE((wb == nullptr) == (Sb == nullptr), SGH_ERR_ARG,
              "sgh_etd1_update: wb and Sb must both be given or both be NULL");
  const int S = nx * ny;
  const long long N = (long long)S * nz;
  SGH_DISPATCH(dtype, {
    if (N < (1LL << 31))
      hipLaunchKernelGGL((k_etd1_update<F, unsigned>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), (const F*)aK, (const F*)KK, (const F*)wt,
                         (const F*)St, (const F*)wb, (const F*)Sb, (const F*)Q, (F*)out, S,
                         (unsigned)N);
    else
      hipLaunchKernelGGL((k_etd1_update<F, long long>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), (const F*)aK, (const F*)KK, (const F*)wt,
                         (const F*)St, (const F*)wb, (const F*)Sb, (const F*)Q, (F*)out, S, N);
  });
  SGH_CHECK_LAUNCH("sgh_etd1_update");
  return SGH_OK;
}

int sgh_add_volume_source(int dtype, long long n, void* a, const void* KK, const void* Q,
                          void* stream) {
  SGH_REQUIRE(n > 0 && a && KK && Q, SGH_ERR_ARG, "sgh_add_volume_source: bad args");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_add_volume<F>, dim3(grid_for(n)), dim3(kThreads), 0, STREAM(stream),
                       (F*)a, (const F*)KK, (const F*)Q, n);
  });
  SGH_CHECK_LAUNCH("sgh_add_volume_source");
  return SGH_OK;
}

int sgh_add_surface_source(int dtype, int nx, int ny, int nz, void* a, const void* KK,
                           const void* w, const void* B, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && a && KK && w && B, SGH_ERR_ARG,
              "sgh_add_surface_source: bad args");
  const int S = nx * ny;
  const long long N = (long long)S * nz;
  SGH_DISPATCH(dtype, {
    if (N < (1LL << 31))
      hipLaunchKernelGGL((k_add_surface<F, unsigned>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), (F*)a, (const F*)KK, (const F*)w, (const F*)B, S,
                         (unsigned)N);
    else
      hipLaunchKernelGGL((k_add_surface<F, long long>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), (F*)a, (const F*)KK, (const F*)w, (const F*)B, S, N);
  });
  SGH_CHECK_LAUNCH("sgh_add_surface_source");
  return SGH_OK;
}

int sgh_mul_inplace(int dtype, long long n, void* a, const void* E, void* stream) {
  SGH_REQUIRE(n > 0 && a && E, SGH_ERR_ARG, "sgh_mul_inplace: bad args");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_mul_inplace<F>, dim3(grid_for(n)), dim3(kThreads), 0, STREAM(stream),
                       (F*)a, (const F*)E, n);
  });
  SGH_CHECK_LAUNCH("sgh_mul_inplace");
  return SGH_OK;
}

int sgh_zcontract(int dtype, int nx, int ny, int nz, const void* a, const void* w, int nchunks,
                  void* parts, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && a && w && parts && nchunks >= 1 && nchunks <= nz,
              SGH_ERR_ARG, "sgh_zcontract: bad args");
  const int S = nx * ny;
  const int pc = (nz + nchunks - 1) / nchunks;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_zcontract<F>, dim3(cdiv(S, kThreads), nchunks), dim3(kThreads), 0,
                       STREAM(stream), (const F*)a, (const F*)w, (F*)parts, S, nz, pc);
  });
  SGH_CHECK_LAUNCH("sgh_zcontract");
  return SGH_OK;
}

int sgh_sum_parts(int dtype, long long n, int nparts, void* parts, void* stream) {
  SGH_REQUIRE(n > 0 && nparts >= 1 && parts, SGH_ERR_ARG, "sgh_sum_parts: bad args");
  if (nparts == 1) return SGH_OK;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_sum_parts<F>, dim3(grid_for(n)), dim3(kThreads), 0, STREAM(stream),
                       (F*)parts, n, nparts);
  });
  SGH_CHECK_LAUNCH("sgh_sum_parts");
  return SGH_OK;
}

int sgh_picard_partials_count(int nx, int ny, int nz, int nchunks) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nchunks < 1) return SGH_ERR_ARG;
  return (int)cdiv((long long)nx * ny, kThreads) * nchunks;
}

int sgh_picard_iter_fused_norms(int dtype, int nx, int ny, int nz, const void* aK,
                                const void* KK, const void* a_old, const void* wt,
                                const void* St, const void* wb, const void* Sb, const void* Q,
                                double omega, const void* wz, int nchunks, void* zc_parts,
                                void* a_new, double* partials, unsigned int* counter,
                                double* norms, double* host_out, double seq, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && aK && KK && a_old && wt && St && a_new && partials,
              SGH_ERR_ARG, "sgh_picard_iter_fused: bad args");
  SGH_REQUIRE((counter == nullptr) == (norms == nullptr), SGH_ERR_ARG,
              "sgh_picard_iter_fused: counter and norms must both be given or both be NULL");
  SGH_REQUIRE(host_out == nullptr || norms != nullptr, SGH_ERR_ARG,
              "sgh_picard_iter_fused: host_out needs the norm reduction");
  SGH_REQUIRE(nchunks >= 1 && nchunks <= nz, SGH_ERR_ARG, "sgh_picard_iter_fused: nchunks");
  SGH_REQUIRE((wb == nullptr) == (Sb == nullptr), SGH_ERR_ARG,
              "sgh_picard_iter_fused: wb and Sb must both be given or both be NULL");
  SGH_REQUIRE(wz == nullptr || zc_parts != nullptr, SGH_ERR_ARG,
              "sgh_picard_iter_fused: wz given without zc_parts");
  SGH_REQUIRE(a_new != a_old, SGH_ERR_ARG, "sgh_picard_iter_fused: a_new must not alias a_old");
  const int S = nx * ny;
  const int pc = (nz + nchunks - 1) / nchunks;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_picard_fused<F>, dim3(cdiv(S, kThreads), nchunks), dim3(kThreads), 0,
                       STREAM(stream), (const F*)aK, (const F*)KK, (const F*)a_old,
                       (const F*)wt, (const F*)St, (const F*)wb, (const F*)Sb, (const F*)Q,
                       (F)omega, (const F*)wz, (F*)zc_parts, (F*)a_new, partials, S, nz, pc,
                       counter, (double)nx * ny * nz, norms, host_out, seq);
  });
  SGH_CHECK_LAUNCH("sgh_picard_iter_fused");
  return SGH_OK;
}

int sgh_picard_iter_fused(int dtype, int nx, int ny, int nz, const void* aK, const void* KK,
                          const void* a_old, const void* wt, const void* St, const void* wb,
                          const void* Sb, const void* Q, double omega, const void* wz,
                          int nchunks, void* zc_parts, void* a_new, double* partials,
                          void* stream) {
  return sgh_picard_iter_fused_norms(dtype, nx, ny, nz, aK, KK, a_old, wt, St, wb, Sb, Q, omega,
                                     wz, nchunks, zc_parts, a_new, partials, nullptr, nullptr,
                                     nullptr, 0.0, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Allocation-free form of the fused Picard pass: its scratch (the per-block norm pairs
// and the last-workgroup counter) is a fixed device array of this library instead of
// caller buffers. Every GPU holds its own copy until the process ends; the host reaches
// it only through the address hipGetSymbolAddress gives on that device.
// ---------------------------------------------------------------------------------
namespace sgh {

constexpr int kNoAllocMaxPairs = SGH_NOALLOC_MAX_PAIRS;
// [0, 2·kNoAllocMaxPairs): norm pairs; then one 32-bit counter word (zero at load, and
// the pass's last workgroup puts it back to zero)
static __device__ __align__(256) double scratch[2 * kNoAllocMaxPairs + 32];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_zero_word(unsigned int* w) { *w = 0u; }

// Device address of the current device's scratch array; on the first call for a device
// it also enqueues a one-thread kernel on `stream` that zeroes the counter word.
static int scratch_for_current_device(hipStream_t stream, double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail_hip(e, "sgh_picard_iter_fused_norms_noalloc: hipGetDevice");
  SGH_REQUIRE(dev >= 0 && dev < kMaxDevices, SGH_ERR_STATE,
              "sgh_picard_iter_fused_norms_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    double* ptr = nullptr;
    e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return fail_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                      "sgh_picard_iter_fused_norms_noalloc: hipGetSymbolAddress");
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(1), 0, stream,
                       reinterpret_cast<unsigned int*>(ptr + 2 * kNoAllocMaxPairs));
    SGH_CHECK_LAUNCH("sgh_picard_iter_fused_norms_noalloc: counter reset");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return SGH_OK;
}

}  // namespace sgh

extern "C" {

int sgh_picard_iter_fused_norms_noalloc(int dtype, int nx, int ny, int nz, const void* aK,
                                        const void* KK, const void* a_old, const void* wt,
                                        const void* St, const void* wb, const void* Sb,
                                        const void* Q, double omega, const void* wz,
                                        int nchunks, void* zc_parts, void* a_new,
                                        double* norms, double* host_out, double seq,
                                        void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && nchunks >= 1 && nchunks <= nz && norms,
              SGH_ERR_ARG, "sgh_picard_iter_fused_norms_noalloc: bad args");
  const long long pairs = (long long)cdiv((long long)nx * ny, kThreads) * nchunks;
  SGH_REQUIRE(pairs <= kNoAllocMaxPairs, SGH_ERR_ARG,
              "sgh_picard_iter_fused_norms_noalloc: %lld workgroups (nx·ny/256 · nchunks) "
              "exceed the fixed scratch of %d norm pairs", pairs, kNoAllocMaxPairs);
  double* scr = nullptr;
  const int rc = scratch_for_current_device(STREAM(stream), &scr);
  if (rc) return rc;
  return sgh_picard_iter_fused_norms(
      dtype, nx, ny, nz, aK, KK, a_old, wt, St, wb, Sb, Q, omega, wz, nchunks, zc_parts, a_new,
      scr, reinterpret_cast<unsigned int*>(scr + 2 * kNoAllocMaxPairs), norms, host_out, seq,
      stream);
}

int sgh_reduce_norms(const double* partials, int n_pairs, double n_elements, double* out,
                     void* stream) {
  SGH_REQUIRE(partials && out && n_pairs > 0 && n_elements > 0, SGH_ERR_ARG,
              "sgh_reduce_norms: bad args");
  hipLaunchKernelGGL(k_reduce_norms, dim3(1), dim3(kThreads), 0, STREAM(stream), partials,
                     n_pairs, n_elements, out);
  SGH_CHECK_LAUNCH("sgh_reduce_norms");
  return SGH_OK;
}

int sgh_reduce(int dtype, long long n, const void* x, int op, double* out, void* stream) {
  SGH_REQUIRE(n > 0 && x && out && (op == 0 || op == 1), SGH_ERR_ARG, "sgh_reduce: bad args");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_reduce<F>, dim3(1), dim3(1024), 0, STREAM(stream), (const F*)x, n, op,
                       out);
  });
  SGH_CHECK_LAUNCH("sgh_reduce");
  return SGH_OK;
}

int sgh_laser_flux(int dtype, int nx, int ny, double dx, double dy, double x0, double y0,
                   double rx, double ry, double order, double I0, int cell_integrated,
                   void* out, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && out && rx > 0 && ry > 0 && dx > 0 && dy > 0, SGH_ERR_ARG,
              "sgh_laser_flux: bad args");
  SGH_REQUIRE(!cell_integrated || order == 2.0, SGH_ERR_ARG,
              "sgh_laser_flux: cell_integrated flux is only defined for order 2");
  const int S = nx * ny;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_laser_flux<F>, dim3(grid_for(S)), dim3(kThreads), 0, STREAM(stream),
                       (F*)out, nx, ny, (F)dx, (F)dy, (F)x0, (F)y0, (F)rx, (F)ry, (F)order,
                       (F)I0, cell_integrated);
  });
  SGH_CHECK_LAUNCH("sgh_laser_flux");
  return SGH_OK;
}

int sgh_shift2d(int dtype, int ny, int nx, const void* in, double sy, double sx,
                const void* base, double alpha, void* out, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && in && out && in != out, SGH_ERR_ARG, "sgh_shift2d: bad args");
  SGH_REQUIRE(std::isfinite(sy) && std::isfinite(sx), SGH_ERR_ARG, "sgh_shift2d: non-finite shift");
  const int S = nx * ny;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_shift2d<F>, dim3(grid_for(S)), dim3(kThreads), 0, STREAM(stream),
                       (const F*)in, ny, nx, sy, sx, (const F*)base, (F)alpha, (F*)out);
  });
  SGH_CHECK_LAUNCH("sgh_shift2d");
  return SGH_OK;
}

int sgh_points_yz(int dtype, int nz, int ny, int P, const void* U, const void* By,
                  const void* Bz, void* out, void* stream) {
  SGH_REQUIRE(nz > 0 && ny > 0 && P > 0 && U && By && Bz && out, SGH_ERR_ARG,
              "sgh_points_yz: bad args");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_points_yz<F>, dim3(cdiv(P, kThreads)), dim3(kThreads), 0,
                       STREAM(stream), (const F*)U, (const F*)By, (const F*)Bz, nz, ny, P,
                       (F*)out);
  });
  SGH_CHECK_LAUNCH("sgh_points_yz");
  return SGH_OK;
}

int sgh_shift3d_xy(int dtype, int nz, int ny, int nx, const void* in, double sy, double sx,
                   int nearest, void* out, void* stream) {
  SGH_REQUIRE(nz > 0 && ny > 0 && nx > 0 && in && out && in != out, SGH_ERR_ARG,
              "sgh_shift3d_xy: bad args");
  SGH_REQUIRE(nz <= 65535 && (long long)nx * ny <= (1LL << 31) - 1, SGH_ERR_ARG,
              "sgh_shift3d_xy: box too large");
  SGH_REQUIRE(std::isfinite(sy) && std::isfinite(sx), SGH_ERR_ARG,
              "sgh_shift3d_xy: non-finite shift");
  const int S = nx * ny;
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_shift3d_xy<F>, dim3(std::min(grid_for(S), 1024u), (unsigned)nz),
                       dim3(kThreads), 0, STREAM(stream), (const F*)in, ny, nx, sy, sx,
                       nearest, (F*)out);
  });
  SGH_CHECK_LAUNCH("sgh_shift3d_xy");
  return SGH_OK;
}

int sgh_box_face_peaks(int dtype, int nz, int ny, int nx, const void* Q, double* out,
                       void* stream) {
  SGH_REQUIRE(nz > 0 && ny > 0 && nx > 0 && Q && out, SGH_ERR_ARG,
              "sgh_box_face_peaks: bad args");
  SGH_REQUIRE(nz <= 65535 && (long long)nx * ny <= (1LL << 31) - 1, SGH_ERR_ARG,
              "sgh_box_face_peaks: box too large");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_box_face_peaks<F>, dim3(nz), dim3(kThreads), 0, STREAM(stream),
                       (const F*)Q, ny, nx, out + 4);
  });
  hipLaunchKernelGGL(k_face_peaks_final, dim3(1), dim3(kThreads), 0, STREAM(stream), out + 4,
                     nz, out);
  SGH_CHECK_LAUNCH("sgh_box_face_peaks");
  return SGH_OK;
}

int sgh_surface_forcing(int dtype, long long S, const void* T, const void* q_las, double P0,
                        double T_boil, double dH, double R_v, double T_liq, double h_top,
                        double T0, const void* prog, double kbar, void* q_evap, void* q_net,
                        void* stream) {
  SGH_REQUIRE(S > 0 && T && q_net, SGH_ERR_ARG, "sgh_surface_forcing: bad args");
  SGH_REQUIRE(R_v > 0 && T_boil > 0, SGH_ERR_ARG, "sgh_surface_forcing: R_v and T_boil must be > 0");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_surface_forcing<F>, dim3(grid_for(S)), dim3(kThreads), 0,
                       STREAM(stream), (const F*)T, (const F*)q_las, S,
                       make_forcing<F>(P0, T_boil, dH, R_v, T_liq, h_top, T0, prog, kbar),
                       (F*)q_evap, (F*)q_net);
  });
  SGH_CHECK_LAUNCH("sgh_surface_forcing");
  return SGH_OK;
}

int sgh_convection_flux(int dtype, long long S, const void* T, double h, double T0,
                        const void* prog, double kbar, void* out, void* stream) {
  SGH_REQUIRE(S > 0 && T && out, SGH_ERR_ARG, "sgh_convection_flux: bad args");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_convection<F>, dim3(grid_for(S)), dim3(kThreads), 0, STREAM(stream),
                       (const F*)T, S, (F)h, (F)T0, (const PropProgram*)prog, (F)kbar, (F*)out);
  });
  SGH_CHECK_LAUNCH("sgh_convection_flux");
  return SGH_OK;
}

int sgh_latent_source(int dtype, long long n, const void* T, const void* T_prev, double T_S,
                      double T_L, double rho, double L, double dt, const void* prog,
                      void* out, void* stream) {
  SGH_REQUIRE(n > 0 && T && T_prev && out && dt > 0, SGH_ERR_ARG, "sgh_latent_source: bad args");
  SGH_DISPATCH(dtype, {
    hipLaunchKernelGGL(k_latent<F>, dim3(grid_for(n)), dim3(kThreads), 0, STREAM(stream),
                       (const F*)T, (const F*)T_prev, n,
                       make_latent<F>(T_S, T_L, rho, L, dt, prog), (F*)out);
  });
  SGH_CHECK_LAUNCH("sgh_latent_source");
  return SGH_OK;
}

int sgh_property_correction_source(int dtype, int nx, int ny, int nz, const void* T,
                                   const void* T_prev, const void* prog, double kbar,
                                   double abar, double dt, double dx, double dy, double dz,
                                   void* work, void* out, void* stream) {
  SGH_REQUIRE(nx > 0 && ny > 0 && nz > 0 && T && T_prev && prog && work && out, SGH_ERR_ARG,
              "sgh_property_correction_source: bad args");
  SGH_REQUIRE(dt > 0 && dx > 0 && dy > 0 && dz > 0, SGH_ERR_ARG,
              "sgh_property_correction_source: non-positive spacing or dt");
  const long long N = (long long)nx * ny * nz;
  SGH_DISPATCH(dtype, {
    F* gx = (F*)work;
    F* gy = gx + N;
    F* gz = gy + N;
    const F idx = (F)(1.0 / dx), idy = (F)(1.0 / dy), idz = (F)(1.0 / dz), idt = (F)(1.0 / dt);
    if (N < (1LL << 31)) {
      hipLaunchKernelGGL((k_corr_grad<F, unsigned>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), (const F*)T, (const PropProgram*)prog, (F)kbar, nx,
                         ny, nz, idx, idy, idz, gx, gy, gz);
      hipLaunchKernelGGL((k_corr_div<F, unsigned>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), gx, gy, gz, (const F*)T, (const F*)T_prev,
                         (const PropProgram*)prog, (F)abar, nx, ny, nz, idx, idy, idz, idt,
                         (F*)out);
    } else {
      hipLaunchKernelGGL((k_corr_grad<F, long long>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), (const F*)T, (const PropProgram*)prog, (F)kbar, nx,
                         ny, nz, idx, idy, idz, gx, gy, gz);
      hipLaunchKernelGGL((k_corr_div<F, long long>), dim3(grid_for(N)), dim3(kThreads), 0,
                         STREAM(stream), gx, gy, gz, (const F*)T, (const F*)T_prev,
                         (const PropProgram*)prog, (F)abar, nx, ny, nz, idx, idy, idz, idt,
                         (F*)out);
    }
  });
  SGH_CHECK_LAUNCH("sgh_property_correction_source");
  return SGH_OK;
}

}  // extern "C"
