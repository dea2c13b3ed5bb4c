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
 * K(point i, point j) for a symmetric kernel (K(i, j) == K(j, i): the M2L
 * blocks are stored once and applied both ways, as the reference assumes at
 * AFMM2DTree.cpp:1081-1082), finite at every pair of points including
 * coincident ones.  kp: the 4 kparams given at create (zeros if none); a:
 * 0.5^nLevels.  AFMM2D_USER_FN is defined by the including file (a host and
 * device inline function in HIP, a static inline function in C); use only
 * <math.h> functions.
 *
 * This example: the Laplace kernel of an anisotropic medium,
 * 0.5 log(dx^2 + (g dy)^2), g = kp[0] (2 when kp[0] is 0), 0 on the diagonal.
 */
AFMM2D_USER_FN double afmm2d_user_kernel(double xi, double yi, double xj, double yj,
                                         const double* kp, double a) {
  (void)a;
  const double g = kp[0] > 0.0 ? kp[0] : 2.0;
  const double dx = xi - xj;
  const double dy = g * (yi - yj);
  const double r2 = dx * dx + dy * dy;
  return r2 == 0.0 ? 0.0 : 0.5 * log(r2);
}
