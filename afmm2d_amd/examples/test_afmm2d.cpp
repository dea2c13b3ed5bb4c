// test_afmm2d.cpp -- the driver flow of the reference's examples/testAFMM2D.cpp
// (arguments N, MinParticlesInLeaf, TOL_POW; build, assemble, matvec, error
// of one random leaf) on the MI355X engine through the C++ facade.  Points
// and b come from a seeded generator instead of rand(); the kernel is the
// log kernel (0.5 log R^2) unless a fourth argument names another:
// log | inv_r | gaussian | reg_inv_r.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "afmm2d.hpp"

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char* argv[]) {
  int N = 6400, M = 200, TOL_POW = 12;  // the reference's defaults
  int kid = AFMM2D_KERNEL_LOG;
  if (argc >= 4) {
    N = std::atoi(argv[1]);
    M = std::atoi(argv[2]);
    TOL_POW = std::atoi(argv[3]);
  } else {
    std::printf("All arguments weren't passed to executable!\nUsing Default Arguments:\n");
  }
  if (argc >= 5) {
    const char* names[4] = {"log", "inv_r", "gaussian", "reg_inv_r"};
    for (int i = 0; i < 4; i++)
      if (std::strcmp(argv[4], names[i]) == 0) kid = i;
  }
  std::mt19937_64 rng(2024);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::vector<double> b(N), loc(2 * (size_t)N);
  for (int i = 0; i < N; i++) b[i] = U(rng);
  for (int j = 0; j < N; j++) loc[j] = U(rng);            // column-major N x 2
  for (int j = 0; j < N; j++) loc[(size_t)N + j] = U(rng);
  try {
    double t0 = now();
    afmm2d::AFMM afmm(N, M, TOL_POW, loc, kid);
    double t1 = now();
    std::printf("\nAFMM tree creation time: %g\n", t1 - t0);
    std::printf("========================= Problem Parameters =========================\n");
    std::printf("Matrix Size                        :%d\n", N);
    std::printf("Leaf Size                          :%d\n", M);
    std::printf("Tolerance                          :%g\n\n", std::pow(10.0, -TOL_POW));
    t0 = now();
    afmm.assemble();
    t1 = now();
    std::printf("========================= Assembly Time =========================\n");
    std::printf("Time for assemble in AFMM form    :%g\n", t1 - t0);
    t0 = now();
    std::vector<double> y = afmm.computeMatVecProduct(b);
    t1 = now();
    std::printf("========================= Matrix-Vector Multiplication =========================\n");
    std::printf("Time for MatVec in AFMM form       :%g\n", t1 - t0);
    double err = afmm.evaluateError(1);
    std::printf("nLevels: %u\terr: %.6e\n", afmm.nLevels, err);
    double s = 0.0;
    for (double v : y) s += v * v;
    std::printf("norm of output: %.12e\n", std::sqrt(s));
    return err < 1e-3 ? 0 : 1;
  } catch (const afmm2d::Error& e) {
    std::fprintf(stderr, "afmm2d error %d: %s\n", e.status, e.what());
    return 2;
  }
}
