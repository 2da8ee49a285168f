// Device evaluator for temperature-dependent material properties.
//
// A property branch is a whitelisted expression of T (the reference's PropertyExpr,
// sgh/core/property_expr.py:69-147). The host compiles each of the six branches
// (k, rho, c; solid and liquid) into a short stack program; the kernels here
// interpret it per cell and blend the branches with the clamped liquid fraction
// (sgh/core/properties.py:61-108). The layout must match
// spectral_galerkin_heat_amd/core/properties.py:pack_property_program.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace sgh {

constexpr int kProgMaxOps = 64;
constexpr int kProgMaxConsts = 24;
constexpr int kProgMaxStack = 16;

struct PropProgram {
  int32_t len[6];
  int32_t ops[6][kProgMaxOps];
  double consts[6][kProgMaxConsts];
  double t_sol;
  double inv_band;
  int32_t degenerate;
  int32_t same;   // bit b: solid and liquid programs of property b (k, rho, c) identical
};
static_assert(sizeof(PropProgram) == 2736, "property program layout");

enum PropBranch { kKSolid = 0, kKLiquid, kRhoSolid, kRhoLiquid, kCSolid, kCLiquid };

enum Op {
  OP_CONST = 0, OP_T = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_DIV = 5, OP_POW = 6,
  OP_NEG = 7, OP_FN = 8, OP_ATAN2 = 9, OP_SQR = 10, OP_CUBE = 11, OP_RECIP = 12
};

// Elementary functions of the whitelist. The cheap ones are inlined into the
// interpreter; the transcendental ones are out-of-line calls, so the interpreter loop
// keeps a small register footprint (inlining all of them made every property kernel
// spill to scratch memory).
template <typename F>
__device__ __attribute__((noinline)) F mfn_slow(int f, F x) {
  switch (f) {
    case 1: return exp(x);
    case 2: return log(x);
    case 3: return log2(x);
    case 4: return log10(x);
    case 5: return sin(x);
    case 6: return cos(x);
    case 7: return tan(x);
    case 8: return asin(x);
    case 9: return acos(x);
    case 10: return atan(x);
    case 11: return sinh(x);
    case 12: return cosh(x);
    case 13: return tanh(x);
    case 14: return asinh(x);
    case 15: return acosh(x);
    case 16: return atanh(x);
    case 21: return log1p(x);
    default: return expm1(x);
  }
}

template <typename F>
__device__ __attribute__((noinline)) F mpow(F a, F b) { return pow(a, b); }
template <typename F>
__device__ __attribute__((noinline)) F matan2(F a, F b) { return atan2(a, b); }

template <typename F>
__device__ __forceinline__ F mfn(int f, F x) {
  switch (f) {
    case 0: return sqrt(x);
    case 17: return fabs(x);
    case 18: return ceil(x);
    case 19: return floor(x);
    case 20: return trunc(x);
    default: return mfn_slow<F>(f, x);
  }
}

// Per-thread evaluation stack in LDS (kProgMaxStack slots × kProgThreads lanes), indexed
// [slot][thread]: a dynamically indexed private array would live in scratch memory.
constexpr int kProgThreads = 256;

// Evaluate branch `w` of the program at temperature T (in precision F). Kernels that call
// it run kProgThreads threads per block. The stack depth after each op depends on the
// program only (it is uniform), so the top four entries live in registers (t0 the top) and
// only deeper entries in LDS: an expression of depth ≤ 4 runs without memory traffic, each
// op a few register moves, where a stack in LDS put a dependent LDS round trip in every op.
// Every op computes exactly what the LDS stack computed (same operands, same order).
template <typename F>
__device__ F prog_eval(const PropProgram* __restrict__ P, int w, F T) {
  __shared__ F stk[kProgMaxStack * kProgThreads];
  F* st = stk + threadIdx.x;
  F t0 = F(0), t1 = F(0), t2 = F(0), t3 = F(0);
  int sp = 0;
  const int n = P->len[w];
  const int32_t* ops = P->ops[w];
  auto push = [&](F v) __attribute__((always_inline)) {
    if (sp >= 4) st[(sp - 4) * kProgThreads] = t3;
    t3 = t2;
    t2 = t1;
    t1 = t0;
    t0 = v;
    ++sp;
  };
  // the top two entries replaced by r
  auto pop2 = [&](F r) __attribute__((always_inline)) {
    t0 = r;
    t1 = t2;
    t2 = t3;
    --sp;
    if (sp >= 4) t3 = st[(sp - 4) * kProgThreads];
  };
  for (int i = 0; i < n; ++i) {
    const int op = ops[i];
    switch (op) {
      case OP_CONST: push((F)P->consts[w][ops[++i]]); break;
      case OP_T: push(T); break;
      case OP_ADD: pop2(t1 + t0); break;
      case OP_SUB: pop2(t1 - t0); break;
      case OP_MUL: pop2(t1 * t0); break;
      case OP_DIV: pop2(t1 / t0); break;
      case OP_POW: pop2(mpow<F>(t1, t0)); break;
      case OP_NEG: t0 = -t0; break;
      case OP_FN: { const int f = ops[++i]; t0 = mfn<F>(f, t0); } break;
      case OP_ATAN2: pop2(matan2<F>(t1, t0)); break;
      case OP_SQR: t0 = t0 * t0; break;
      case OP_CUBE: t0 = t0 * t0 * t0; break;
      case OP_RECIP: t0 = F(1) / t0; break;
      default: break;
    }
  }
  return sp > 0 ? t0 : F(0);
}

// Liquid fraction clamped to [0, 1]; a step at t_sol when the band is degenerate.
template <typename F>
__device__ __forceinline__ F prog_fl(const PropProgram* __restrict__ P, F T) {
  if (P->degenerate) return T >= (F)P->t_sol ? F(1) : F(0);
  F f = (T - (F)P->t_sol) * (F)P->inv_band;
  return f < F(0) ? F(0) : (f > F(1) ? F(1) : f);
}

// Blended property (1 - f_l)·solid + f_l·liquid for branch pair `solid_idx`.
// With identical solid and liquid programs the one value is blended with itself: the same
// arithmetic as evaluating both, with half the interpretation.
template <typename F>
__device__ __forceinline__ F prog_blend(const PropProgram* __restrict__ P, int solid_idx, F T, F fl) {
  if ((P->same >> (solid_idx >> 1)) & 1) {
    const F v = prog_eval<F>(P, solid_idx, T);
    return (F(1) - fl) * v + fl * v;
  }
  return (F(1) - fl) * prog_eval<F>(P, solid_idx, T) + fl * prog_eval<F>(P, solid_idx + 1, T);
}

// Latent-heat source −ρ·L·(f_l(T) − f_l(T_prev))/dt. With a property program: blended
// ρ(T) and the program's liquid fraction; otherwise constant ρ and f_l = (T − T_S)·inv_band
// clamped to [0, 1] (a step at T_S when the band is degenerate).
template <typename F>
struct LatentParams {
  F T_S, inv_band, rho, L, dt;
  int degenerate;
  const PropProgram* prog;
};

template <typename F>
__device__ __forceinline__ F latent_value(const LatentParams<F>& p, F t, F tp) {
  if (p.prog) {
    const F flc = prog_fl<F>(p.prog, t), flp = prog_fl<F>(p.prog, tp);
    const F r = prog_blend<F>(p.prog, kRhoSolid, t, flc);
    return -r * p.L * (flc - flp) * (F(1) / p.dt);
  }
  F fc, fp;
  if (p.degenerate) {
    fc = t >= p.T_S ? F(1) : F(0);
    fp = tp >= p.T_S ? F(1) : F(0);
  } else {
    fc = (t - p.T_S) * p.inv_band;
    fc = fc < F(0) ? F(0) : (fc > F(1) ? F(1) : fc);
    fp = (tp - p.T_S) * p.inv_band;
    fp = fp < F(0) ? F(0) : (fp > F(1) ? F(1) : fp);
  }
  return (-p.rho * p.L / p.dt) * (fc - fp);
}

// Top-surface forcing of one cell at temperature t: the evaporation flux (0.82·ΔH·P0 /
// √(2πR_v) · T^-½ · exp(ΔH/(R_v T_b)·(1 − T_b/T)), zero below T_liq) and the net flux
// q_las − q_evap − h_top(T − T0), rescaled by k̄/k(T) with a property program.
template <typename F>
struct ForcingParams {
  F f1, f2, T_boil, T_liq, h_top, T0, kbar;
  const PropProgram* prog;
};

__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ float fexp(float x) { return expf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return sqrtf(x); }

template <typename F>
__device__ __forceinline__ F evap_flux(F T, F f1, F f2, F T_boil, F T_liq) {
  if (T < T_liq) return F(0);
  return f1 * ((F(1) / fsqrt(T)) * fexp(f2 * (F(1) - T_boil / T)));
}

template <typename F>
__device__ __forceinline__ void surface_forcing_value(const ForcingParams<F>& p, F t, F q_las,
                                                      F& q_evap, F& q_net) {
  const F qe = evap_flux(t, p.f1, p.f2, p.T_boil, p.T_liq);
  q_evap = qe;
  F qn = q_las - qe;
  if (p.h_top > F(0)) qn -= p.h_top * (t - p.T0);
  if (p.prog) {
    const F fl = prog_fl<F>(p.prog, t);
    qn = qn * (p.kbar / prog_blend<F>(p.prog, kKSolid, t, fl));
  }
  q_net = qn;
}

template <typename F>
inline ForcingParams<F> make_forcing(double P0, double T_boil, double dH, double R_v,
                                     double T_liq, double h_top, double T0, const void* prog,
                                     double kbar) {
  ForcingParams<F> p;
  p.f1 = (F)(0.82 * dH * P0 / sqrt(2.0 * M_PI * R_v));
  p.f2 = (F)(dH / (R_v * T_boil));
  p.T_boil = (F)T_boil;
  p.T_liq = (F)T_liq;
  p.h_top = (F)h_top;
  p.T0 = (F)T0;
  p.kbar = (F)kbar;
  p.prog = (const PropProgram*)prog;
  return p;
}

template <typename F>
inline LatentParams<F> make_latent(double T_S, double T_L, double rho, double L, double dt,
                                   const void* prog) {
  const int degenerate = T_L <= T_S;
  LatentParams<F> p;
  p.T_S = (F)T_S;
  p.inv_band = (F)(degenerate ? 0.0 : 1.0 / (T_L - T_S));
  p.rho = (F)rho;
  p.L = (F)L;
  p.dt = (F)dt;
  p.degenerate = degenerate;
  p.prog = (const PropProgram*)prog;
  return p;
}

}  // namespace sgh
