// near_math.hpp -- device evaluation of the built-in kernels for the near
// field: the FP64 forms shared by the single-RHS and the blocked near-field
// kernels, and the FP32 forms of the reduced-precision blocked passes.
#pragma once

#include "common.hpp"

namespace afmm {

// The log kernel's near field works in log2 units: log2_near gives
// log2(R^2), and every sum of such values is multiplied once by
// kHalfLn2 = ln(2)/2 (log R = ln(2)/2 log2 R^2).
constexpr double kHalfLn2 = 0.34657359027997264;

// log2(r2) for r2 >= 0 by a table method after Tang (ACM TOMS 16(4), 1990):
// r2 = 2^e m, m in [1, 2); j = the top 8 bits of m's fraction; r = m RN(1/c_j)
// - 1, |r| <= 2^-9 (j >= 1) or in [0, 2^-8) (j = 0, where c_0 = 1);
// log2 r2 = (e + 1) + (-log2 RN(1/c_j) - 1) + log2(1 + r), the last by its
// Taylor polynomial to degree 5 (truncation below 1.3e-17, and 8.5e-16 in
// bin 0).  e + 1 is the exponent frexp returns.  r2 below 2^-1022 (zero:
// the diagonal and coincident points) is read as 1 and gives 0 exactly (j =
// 0, m = 1, r = 0).  About one ulp of the result; no branch.
__device__ __forceinline__ double log2_near(double r2, const double2* __restrict__ tab) {
  int lo = __double2loint(r2);
  int hi = __double2hiint(r2);
  const bool tiny = hi < 0x00100000;
  hi = tiny ? 0x3FF00000 : hi;
  lo = tiny ? 0 : lo;
  const int fe = __builtin_amdgcn_frexp_exp(__hiloint2double(hi, lo));
  const double m = __hiloint2double((hi & 0x000FFFFF) | 0x3FF00000, lo);
  const double2 t = tab[(hi >> 12) & (kNearLogTab - 1)];
  const double r = fma(m, t.x, -1.0);
  // (-1)^(k+1) / (k ln 2), k = 5 .. 1
  double q = fma(r, 0.28853900817779266, -0.36067376022224085);
  q = fma(r, q, 0.48089834696298778);
  q = fma(r, q, -0.72134752044448169);
  q = fma(r, q, 1.4426950408889634);
  return (double)fe + fma(r, q, t.y);
}

// exp(x) for x <= 0 by a table method after Tang (ACM TOMS 15(2), 1989):
// x = (64 k + j) ln(2)/64 + r with n = 64 k + j = rint(64 x / ln 2), |r| <=
// ln(2)/128; exp x = 2^k 2^(j/64) (1 + p(r)), p the Taylor polynomial of
// e^r - 1 to degree 5 (truncation below 3.5e-17 relative).  ln(2)/64 is
// split in two so that n times its leading part is exact; x below -746 is
// raised to -746, where the result underflows to 0 either way.  About one
// ulp; exp(0) = 1 exactly.  tab: RN(2^(j/64)).
__device__ __forceinline__ double exp_near(double x, const double* __restrict__ tab) {
  x = fmax(x, -746.0);
  const double n = rint(x * 92.332482616893657);  // 64 / ln 2
  double r = fma(-n, 0.010830424726009369, x);   // ln(2)/64 to 24 bits
  r = fma(-n, -2.9760223436661086e-11, r);       // and the rest
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  const double p = fma(r * r, q, r);
  const int ni = (int)n;
  const double t = tab[ni & (kNearExpTab - 1)];
  return ldexp(fma(t, p, t), ni >> 6);
}

// double2 entries of a kernel's near-field table in LDS: the log table, the
// Gaussian's exp table (two values per entry), or none.
template <int KID>
constexpr int kNearTabLen = KID == K_LOG ? kNearLogTab : KID == K_GAUSS ? kNearExpTab / 2 : 1;

// A kernel's table from the engine's buffer (the log table, then the exp
// table) into LDS; the caller synchronises before the first use.
template <int KID>
__device__ __forceinline__ void load_near_tab(double2* tab, const double2* __restrict__ src,
                                              int tid, int nthreads) {
  if constexpr (KID == K_LOG || KID == K_GAUSS) {
    const double2* s = src + (KID == K_GAUSS ? kNearLogTab : 0);
    for (int u = tid; u < kNearTabLen<KID>; u += nthreads) tab[u] = s[u];
  }
}

// 1/sqrt(x) for normal x > 0: the hardware estimate y0 = (1 + d)/sqrt(x)
// (v_rsq_f64) refined by one second-order Newton step.  With
// e = 1 - x y0^2 = -2d - d^2, sqrt(x) y0 (1 + e/2 + 3e^2/8) = 1 + O(d^3);
// what remains is the rounding of x y0 in e and of the last FMA, about one
// ulp.  x = 0 gives +inf from the estimate and NaN after the step: callers
// select the diagonal value separately.
__device__ __forceinline__ double rsqrt_refined(double x) {
  const double y0 = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y0), y0, 1.0);
  return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// The kernel from the squared distance for the near field.  Same values as
// kernel_r2 to a few ulp (times near_scale<KID>), without its branches and
// correctly rounded divisions and square roots: the log as log2_near, the
// 1/R kernels through rsqrt_refined (one estimate, no division), the
// Gaussian as exp_near with the reciprocal of its width taken once
// (ninv_p0 = -1/p0).  tab: the kernel's table (kNearTabLen<KID>).
template <int KID>
__device__ __forceinline__ double near_kernel(double r2, const KParams& kp, double ninv_p0,
                                              double inv_a, const double2* tab) {
  if constexpr (KID == K_LOG) {
    return log2_near(r2, tab);
  } else if constexpr (KID == K_INV_R) {
    const double rs = rsqrt_refined(r2);
    return r2 == 0.0 ? 0.0 : rs;
  } else if constexpr (KID == K_GAUSS) {
    return exp_near(r2 * ninv_p0, reinterpret_cast<const double*>(tab));
  } else {
    // R < a ? R/a : a/R = min(R/a, a/R) = (1/R) min(R^2/a, a); r2 is kept
    // off zero inside the square root only, so R = 0 gives 1e150 x 0 = 0
    // exactly
    const double rs = rsqrt_refined(fmax(r2, 1e-300));
    return rs * fmin(r2 * inv_a, kp.a);
  }
}

// K(target, source) for the near field (times near_scale<KID>): the user
// kernel takes the points, the built-in ones the squared distance.
template <int KID>
__device__ __forceinline__ double pair_value(double xi, double yi, double sx, double sy,
                                             const KParams& kp, double ninv_p0, double inv_a,
                                             const double2* tab) {
  if constexpr (KID >= K_USER) {
    return kernel_xy<KID>(xi, yi, sx, sy, kp);
  } else {
    const double dx = xi - sx, dy = yi - sy;
    return near_kernel<KID>(fma(dx, dx, dy * dy), kp, ninv_p0, inv_a, tab);
  }
}

// The factor of the near field's FP64 sums: near_kernel<K_LOG> is log2 R^2.
template <int KID>
constexpr double near_scale = KID == K_LOG ? kHalfLn2 : 1.0;

// FP32 kernel value for the reduced-precision near field: the difference of
// the FP64 coordinates rounded to FP32, the squared distance in FP32, then
// the hardware log2 (v_log_f32), reciprocal square root (v_rsq_f32) or exp2
// (v_exp_f32), each about 1 ulp.  c0, c1: 0.5 ln 2 (log), -log2(e)/p0
// (Gaussian), 1/a and a (a/R).  R2 == 0 gives the diagonal value.
template <int KID>
__device__ __forceinline__ float near_kernel_f32(double xi, double yi, double sx, double sy,
                                                 const KParams& kp, float c0, float c1) {
  if constexpr (KID >= K_USER) {
    return (float)kernel_xy<KID>(xi, yi, sx, sy, kp);
  } else {
    const float dx = (float)(xi - sx), dy = (float)(yi - sy);
    const float r2 = fmaf(dx, dx, dy * dy);
    if constexpr (KID == K_LOG) {
      const float v = __builtin_amdgcn_logf(r2) * c0;
      return r2 == 0.0f ? 0.0f : v;
    } else if constexpr (KID == K_INV_R) {
      const float v = __builtin_amdgcn_rsqf(r2);
      return r2 == 0.0f ? 0.0f : v;
    } else if constexpr (KID == K_GAUSS) {
      return __builtin_amdgcn_exp2f(r2 * c0);
    } else {
      const float rs = __builtin_amdgcn_rsqf(fmaxf(r2, 1e-30f));
      return rs * fminf(r2 * c0, c1);
    }
  }
}

}  // namespace afmm
