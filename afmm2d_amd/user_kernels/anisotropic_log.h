/*
 * anisotropic_log.h -- the user kernel compiled into libafmm2d.so by default
 * as AFMM2D_KERNEL_USER (kernel id 4), and into the CPU oracle.
 *
 * The reference takes its kernel from the user's own source file
 * (userkernel::getMatrixEntry, kernel.hpp:58; e.g. examples/testAFMM2D.cpp:
 * 9-20) at build time.  Here the user's kernel is a header like this one,
 * given at build time (python -m afmm2d_amd.build --user-kernel FILE, or
 * AFMM2D_USER_KERNEL=FILE); the same header builds the device kernels, the
 * host entry points and the oracle, so all three evaluate one formula.
 *
 * Contract of the header: define
 *
 *   AFMM2D_USER_FN double afmm2d_user_kernel(double xi, double yi,
 *                                            double xj, double yj,
 *                                            const double* kp, double a);
 *
 * K(point i, point j), finite at every pair of points including coincident
 * ones.  The kernel need not be symmetric: the reference's algorithm assumes
 * K(i, j) == K(j, i) (AFMM2DTree.cpp:1081-1082), so a non-symmetric kernel
 * (found at create by comparing K(i, j) with K(j, i) on sampled pairs) runs
 * as two operators, its symmetric and its antisymmetric part, whose results
 * are added.  kp: the 4 kparams given at create (zeros if none); a:
 * 0.5^nLevels.  AFMM2D_USER_FN is defined by the including file (a host and
 * device inline function in HIP, a static inline function in C); use only
 * <math.h> functions.
 *
 * This example: the Laplace kernel of an anisotropic medium plus a dipole
 * field along x,
 *   0.5 log(dx^2 + (g dy)^2) + b dx / (dx^2 + dy^2),   0 on the diagonal,
 * g = kp[0] (2 when kp[0] is 0), b = kp[1] (default 0: symmetric; any other
 * value makes the kernel non-symmetric).
 */
AFMM2D_USER_FN double afmm2d_user_kernel(double xi, double yi, double xj, double yj,
                                         const double* kp, double a) {
  (void)a;
  const double g = kp[0] > 0.0 ? kp[0] : 2.0;
  const double b = kp[1];
  const double dx = xi - xj;
  const double dy = yi - yj;
  const double gy = g * dy;
  const double r2g = dx * dx + gy * gy;
  if (r2g == 0.0) return 0.0;
  const double v = 0.5 * log(r2g);
  return b == 0.0 ? v : v + b * dx / (dx * dx + dy * dy);
}
