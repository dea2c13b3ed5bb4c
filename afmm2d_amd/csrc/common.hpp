// common.hpp -- shared definitions of the MI355X AFMM2D engine: HIP error
// plumbing, the built-in kernel registry (device and host), launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace afmm {

// Error carried from the implementation to the C ABI.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define AFMM_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::afmm::Error(e_ == hipErrorOutOfMemory ? 4 : 5,                           \
                          std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define AFMM_LAUNCH_CHECK() AFMM_HIP(hipGetLastError())

// M2L block storage modes (Engine::m2l_store, afmm2d_set_m2l_storage).
enum M2LStore { M2L_STORE_F64 = 0, M2L_STORE_B48 = 1 };

// 48-bit block floating point: a block's entries v are stored as integers
// m = rint(v 2^-s) with one exponent s per block, the smallest for which
// every |m| <= 2^46 (s = e - 46 with max |v| < 2^e), so
// |v - m 2^s| <= 2^(s-1) <= 2^-46 max |v|.  m is kept biased, as the
// unsigned 48-bit u = m + 2^47: its low 32 bits and its top 16.  The double
// whose exponent field is that of 2^52 and whose mantissa field is u is
// exactly 2^52 + u, so one subtraction of kB48Bias = 2^52 + 2^47 returns m
// (exactly: an integer of magnitude <= 2^46).
constexpr double kB48Bias = 4503599627370496.0 + 140737488355328.0;  // 2^52 + 2^47
constexpr uint32_t kB48ExpWord = 0x43300000u;  // top word of 2^52
__host__ __device__ __forceinline__ int b48_exponent(double maxabs) {
  if (!(maxabs > 0.0)) return 0;
  int e;
  (void)frexp(maxabs, &e);  // maxabs = f 2^e, 1/2 <= f < 1
  return e - 46;
}
__host__ __device__ __forceinline__ void b48_split(double v, int sexp, uint32_t& lo32,
                                                   uint16_t& hi16) {
  const unsigned long long u = (unsigned long long)(llrint(ldexp(v, -sexp)) + (1ll << 47));
  lo32 = (uint32_t)u;
  hi16 = (uint16_t)(u >> 32);
}
__host__ __device__ __forceinline__ double b48_join(uint32_t lo32, uint16_t hi16, int sexp) {
  const long long u = (long long)(((unsigned long long)hi16 << 32) | lo32);
  return ldexp((double)(u - (1ll << 47)), sexp);
}

// Kernel ids match afmm2d_kernel_id.
// K_USER_SYM and K_USER_ANTI are the symmetric and antisymmetric parts of
// the user kernel, (K(x, y) + K(y, x)) / 2 and (K(x, y) - K(y, x)) / 2: a
// non-symmetric user kernel runs as the sum of two operators, one for each.
enum KernelId {
  K_LOG = 0,
  K_INV_R = 1,
  K_GAUSS = 2,
  K_REG_INV_R = 3,
  K_USER = 4,
  K_USER_SYM = 5,
  K_USER_ANTI = 6
};
constexpr int kNumKernels = 7;

// K(y, x) = kTransposeSign * K(x, y): -1 for the antisymmetric part, whose
// M2L blocks and near-field values are used with this sign in the
// transposed direction.
template <int KID>
constexpr double kTransposeSign = KID == K_USER_ANTI ? -1.0 : 1.0;

// Entries of the near-field log table (log2_near): (RN(1/c_j), -log2
// RN(1/c_j) - 1) for c_0 = 1 and c_j = 1 + (j + 1/2)/256, j = 1 .. 255,
// evaluated in extended precision on the host.
constexpr int kNearLogTab = 256;
// Entries of the near-field exp table (exp_near), after the log table in the
// same buffer: RN(2^(j/64)), j = 0 .. 63.
constexpr int kNearExpTab = 64;

// Parameters every kernel needs: p0 = Gaussian width (s^2), a = the
// regularisation radius of K_REG_INV_R (0.5^L), u = the kparams as given
// (zeros if none), for the user kernel.
struct KParams {
  double p0;
  double a;
  double u[4];
};
}  // namespace afmm

// The user kernel (K_USER), from the header named by AFMM2D_USER_KERNEL_HEADER
// at build time.  It is compiled without contracting multiplies and adds into
// fused operations, like the host compiler and the CPU oracle evaluate it.
#define AFMM2D_USER_FN __host__ __device__ inline
#pragma clang fp contract(off)
#include AFMM2D_USER_KERNEL_HEADER
#pragma clang fp contract(fast)

namespace afmm {

// Kernel value from the squared distance.  R2 == 0 (the diagonal) gives the
// kernel's diagonal value: 0 for the singular kernels, exp(0) = 1 for the
// Gaussian, 0/a for K_REG_INV_R.  Shared by host and device so the host
// plug-in and the device kernels evaluate the same formula.
template <int KID>
__host__ __device__ __forceinline__ double kernel_r2(double r2, const KParams& p) {
  if constexpr (KID == K_LOG) {
    return r2 == 0.0 ? 0.0 : 0.5 * log(r2);
  } else if constexpr (KID == K_INV_R) {
    return r2 == 0.0 ? 0.0 : 1.0 / sqrt(r2);
  } else if constexpr (KID == K_GAUSS) {
    return exp(-r2 / p.p0);
  } else {
    double r = sqrt(r2);
    return r < p.a ? r / p.a : p.a / r;
  }
}

// Squared distance with separate multiplies and an add (no fused
// multiply-add), matching the host evaluation bit for bit.
__host__ __device__ __forceinline__ double dist2_exact(double xi, double yi, double xj, double yj) {
#pragma clang fp contract(off)
  double dx = xi - xj;
  double dy = yi - yj;
  double a = dx * dx;
  double b = dy * dy;
  return a + b;
}

template <int KID>
__host__ __device__ __forceinline__ double kernel_xy(double xi, double yi, double xj, double yj,
                                                     const KParams& p) {
  if constexpr (KID == K_USER) {
    return ::afmm2d_user_kernel(xi, yi, xj, yj, p.u, p.a);
  } else if constexpr (KID == K_USER_SYM) {
    return 0.5 * (::afmm2d_user_kernel(xi, yi, xj, yj, p.u, p.a) +
                  ::afmm2d_user_kernel(xj, yj, xi, yi, p.u, p.a));
  } else if constexpr (KID == K_USER_ANTI) {
    return 0.5 * (::afmm2d_user_kernel(xi, yi, xj, yj, p.u, p.a) -
                  ::afmm2d_user_kernel(xj, yj, xi, yi, p.u, p.a));
  } else {
    return kernel_r2<KID>(dist2_exact(xi, yi, xj, yj), p);
  }
}

inline double kernel_host(int kid, const KParams& p, double xi, double yi, double xj, double yj) {
  switch (kid) {
    case K_LOG: return kernel_xy<K_LOG>(xi, yi, xj, yj, p);
    case K_INV_R: return kernel_xy<K_INV_R>(xi, yi, xj, yj, p);
    case K_GAUSS: return kernel_xy<K_GAUSS>(xi, yi, xj, yj, p);
    case K_REG_INV_R: return kernel_xy<K_REG_INV_R>(xi, yi, xj, yj, p);
    case K_USER: return kernel_xy<K_USER>(xi, yi, xj, yj, p);
    case K_USER_SYM: return kernel_xy<K_USER_SYM>(xi, yi, xj, yj, p);
    default: return kernel_xy<K_USER_ANTI>(xi, yi, xj, yj, p);
  }
}

// Dispatch a templated launcher on the runtime kernel id.
#define AFMM_DISPATCH_KERNEL(kid, KID, ...)                       \
  switch (kid) {                                                  \
    case ::afmm::K_LOG: { constexpr int KID = 0; __VA_ARGS__; break; }     \
    case ::afmm::K_INV_R: { constexpr int KID = 1; __VA_ARGS__; break; }   \
    case ::afmm::K_GAUSS: { constexpr int KID = 2; __VA_ARGS__; break; }   \
    case ::afmm::K_REG_INV_R: { constexpr int KID = 3; __VA_ARGS__; break; } \
    case ::afmm::K_USER: { constexpr int KID = 4; __VA_ARGS__; break; }    \
    case ::afmm::K_USER_SYM: { constexpr int KID = 5; __VA_ARGS__; break; } \
    default: { constexpr int KID = 6; __VA_ARGS__; break; }                \
  }

inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace afmm
