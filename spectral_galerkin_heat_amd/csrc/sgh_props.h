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
  int32_t pad;
};
static_assert(sizeof(PropProgram) == 2736, "property program layout");

enum PropBranch { kKSolid = 0, kKLiquid, kRhoSolid, kRhoLiquid, kCSolid, kCLiquid };

enum Op {
  OP_CONST = 0, OP_T = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_DIV = 5, OP_POW = 6,
  OP_NEG = 7, OP_FN = 8, OP_ATAN2 = 9, OP_SQR = 10, OP_CUBE = 11, OP_RECIP = 12
};

__device__ __forceinline__ double mfn(int f, double x) {
  switch (f) {
    case 0: return sqrt(x);
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
    case 17: return fabs(x);
    case 18: return ceil(x);
    case 19: return floor(x);
    case 20: return trunc(x);
    case 21: return log1p(x);
    default: return expm1(x);
  }
}

__device__ __forceinline__ float mfn(int f, float x) {
  switch (f) {
    case 0: return sqrtf(x);
    case 1: return expf(x);
    case 2: return logf(x);
    case 3: return log2f(x);
    case 4: return log10f(x);
    case 5: return sinf(x);
    case 6: return cosf(x);
    case 7: return tanf(x);
    case 8: return asinf(x);
    case 9: return acosf(x);
    case 10: return atanf(x);
    case 11: return sinhf(x);
    case 12: return coshf(x);
    case 13: return tanhf(x);
    case 14: return asinhf(x);
    case 15: return acoshf(x);
    case 16: return atanhf(x);
    case 17: return fabsf(x);
    case 18: return ceilf(x);
    case 19: return floorf(x);
    case 20: return truncf(x);
    case 21: return log1pf(x);
    default: return expm1f(x);
  }
}

__device__ __forceinline__ double mpow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float mpow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double matan2(double a, double b) { return atan2(a, b); }
__device__ __forceinline__ float matan2(float a, float b) { return atan2f(a, b); }

// Evaluate branch `w` of the program at temperature T (in precision F).
template <typename F>
__device__ F prog_eval(const PropProgram* __restrict__ P, int w, F T) {
  F st[kProgMaxStack];
  int sp = 0;
  const int n = P->len[w];
  const int32_t* ops = P->ops[w];
  for (int i = 0; i < n; ++i) {
    const int op = ops[i];
    switch (op) {
      case OP_CONST: st[sp++] = (F)P->consts[w][ops[++i]]; break;
      case OP_T: st[sp++] = T; break;
      case OP_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
      case OP_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
      case OP_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
      case OP_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
      case OP_POW: --sp; st[sp - 1] = mpow(st[sp - 1], st[sp]); break;
      case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case OP_FN: { const int f = ops[++i]; st[sp - 1] = mfn(f, st[sp - 1]); } break;
      case OP_ATAN2: --sp; st[sp - 1] = matan2(st[sp - 1], st[sp]); break;
      case OP_SQR: st[sp - 1] = st[sp - 1] * st[sp - 1]; break;
      case OP_CUBE: st[sp - 1] = st[sp - 1] * st[sp - 1] * st[sp - 1]; break;
      case OP_RECIP: st[sp - 1] = F(1) / st[sp - 1]; break;
      default: break;
    }
  }
  return sp > 0 ? st[0] : F(0);
}

// Liquid fraction clamped to [0, 1]; a step at t_sol when the band is degenerate.
template <typename F>
__device__ __forceinline__ F prog_fl(const PropProgram* __restrict__ P, F T) {
  if (P->degenerate) return T >= (F)P->t_sol ? F(1) : F(0);
  F f = (T - (F)P->t_sol) * (F)P->inv_band;
  return f < F(0) ? F(0) : (f > F(1) ? F(1) : f);
}

// Blended property (1 - f_l)·solid + f_l·liquid for branch pair `solid_idx`.
template <typename F>
__device__ __forceinline__ F prog_blend(const PropProgram* __restrict__ P, int solid_idx, F T, F fl) {
  return (F(1) - fl) * prog_eval<F>(P, solid_idx, T) + fl * prog_eval<F>(P, solid_idx + 1, T);
}

}  // namespace sgh
