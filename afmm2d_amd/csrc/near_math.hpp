// near_math.hpp -- device evaluation of the built-in kernels for the near
// field: the FP64 forms shared by the single-RHS and the blocked near-field
// kernels, and the FP32 forms of the reduced-precision blocked passes.
#pragma once

#include "common.hpp"

namespace afmm {

// log(x) for normal x > 0 by Tang's table method (ACM TOMS 16(4), 1990):
// x = 2^e m, m in [1,2); j = the top 7 bits of m's fraction;
// c_j = 1 + (j + 1/2)/128, t = (RN(1/c_j), -log RN(1/c_j));
// log x = e ln2 + t.y + log1p(m t.x - 1), |m t.x - 1| < 2^-8, log1p by its
// Taylor polynomial to degree 6 (truncation < 4e-18).
__device__ __forceinline__ double log_tang(double x, const double2* __restrict__ tab,
                                           double ln2_hi, double ln2_lo) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
  const int ex = (int)(bits >> 52) - 1023;
  const int j = (int)((bits >> 45) & 127);
  const double m = __longlong_as_double((long long)((bits & 0x000FFFFFFFFFFFFFull) |
                                                    0x3FF0000000000000ull));
  const double2 t = tab[j];
  const double r = fma(m, t.x, -1.0);
  double p = fma(r, -1.0 / 6.0, 1.0 / 5.0);
  p = fma(r, p, -0.25);
  p = fma(r, p, 1.0 / 3.0);
  p = fma(r, p, -0.5);
  p = r * p;
  p = fma(r, p, r);  // log1p(r)
  const double e = (double)ex;
  return fma(e, ln2_hi, t.y) + fma(e, ln2_lo, p);
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
// kernel_r2 to a few ulp, without its branches and correctly rounded
// divisions and square roots: the log by Tang's table, the 1/R kernels
// through rsqrt_refined (one estimate, no division), the Gaussian with the
// reciprocal of its width taken once (ninv_p0 = -1/p0).
template <int KID>
__device__ __forceinline__ double near_kernel(double r2, const KParams& kp, double ninv_p0,
                                              double inv_a, const double2* tab, double ln2_hi,
                                              double ln2_lo) {
  if constexpr (KID == K_LOG) {
    // 2 x the log kernel; the factor 1/2 is applied to the sums.  The log is
    // taken of max(r2, tiny) and the diagonal value selected after, so no
    // branch splits the wave
    const double lg = log_tang(fmax(r2, 1e-300), tab, ln2_hi, ln2_lo);
    return r2 == 0.0 ? 0.0 : lg;
  } else if constexpr (KID == K_INV_R) {
    const double rs = rsqrt_refined(r2);
    return r2 == 0.0 ? 0.0 : rs;
  } else if constexpr (KID == K_GAUSS) {
    return exp(r2 * ninv_p0);
  } else {
    // R < a ? R/a : a/R = min(R/a, a/R) with R = r2 / sqrt(r2); r2 is kept
    // off zero inside the square root only, so R = 0 gives R/a = 0 exactly
    const double rs = rsqrt_refined(fmax(r2, 1e-300));
    const double r = r2 * rs;
    return fmin(r * inv_a, kp.a * rs);
  }
}

// K(target, source) for the near field: the user kernel takes the points,
// the built-in ones the squared distance.
template <int KID>
__device__ __forceinline__ double pair_value(double xi, double yi, double sx, double sy,
                                             const KParams& kp, double ninv_p0, double inv_a,
                                             const double2* tab, double ln2_hi, double ln2_lo) {
  if constexpr (KID >= K_USER) {
    return kernel_xy<KID>(xi, yi, sx, sy, kp);
  } else {
    const double dx = xi - sx, dy = yi - sy;
    return near_kernel<KID>(fma(dx, dx, dy * dy), kp, ninv_p0, inv_a, tab, ln2_hi, ln2_lo);
  }
}

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
      return fminf(r2 * rs * c0, c1 * rs);
    }
  }
}

}  // namespace afmm
