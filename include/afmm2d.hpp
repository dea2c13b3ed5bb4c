// afmm2d.hpp -- C++ facade over the C ABI (afmm2d.h), keeping the reference's
// class names and call flow (include/AFMM2D.hpp:14-33, include/kernel.hpp:12-61):
//
//   afmm2d::AFMM* afmm = new afmm2d::AFMM(N, MinParticlesInLeaf, TOL_POW, loc);
//   afmm->assemble();
//   std::vector<double> y = afmm->computeMatVecProduct(b);
//   afmm->evaluateError();
//
// Header-only; link with -lafmm2d.  Errors throw afmm2d::Error.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "afmm2d.h"

namespace afmm2d {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(int status) {
  if (status != AFMM2D_OK)
    throw Error(status, std::string(afmm2d_status_string(status)) + ": " + afmm2d_last_error());
}

// pts2D (definitions.hpp:5-7)
struct pts2D {
  double x, y;
};

// kernel (kernel.hpp:12-34): entries over the KD-sorted particles.  The device
// evaluates the same built-in formula (afmm2d_kernel_id); getMatrixEntry is
// the host evaluation used for checks such as the dense matrix of the
// reference's tutorial (examples/tutorial.cpp:97-131).
class kernel {
 public:
  double a = 0.0;
  std::vector<pts2D> particles_X;
  std::vector<pts2D> particles_Y;
  int kernel_id = AFMM2D_KERNEL_LOG;
  double kparams[4] = {0.0, 0.0, 0.0, 0.0};

  virtual ~kernel() = default;
  virtual double getMatrixEntry(const unsigned i, const unsigned j) const {
    double v = 0.0;
    const pts2D& p = particles_X[i];
    const pts2D& q = particles_Y[j];
    check(afmm2d_kernel_value(kernel_id, kparams, a, p.x, p.y, q.x, q.y, &v));
    return v;
  }
  std::vector<double> getRow(const int j, const std::vector<int>& col_indices) const {
    std::vector<double> r(col_indices.size());
    for (size_t k = 0; k < col_indices.size(); k++) r[k] = getMatrixEntry(j, col_indices[k]);
    return r;
  }
  std::vector<double> getCol(const int k, const std::vector<int>& row_indices) const {
    std::vector<double> c(row_indices.size());
    for (size_t j = 0; j < row_indices.size(); j++) c[j] = getMatrixEntry(row_indices[j], k);
    return c;
  }
  // rows x cols, row-major
  std::vector<double> getMatrix(const std::vector<int>& rows, const std::vector<int>& cols) const {
    std::vector<double> m(rows.size() * cols.size());
    for (size_t j = 0; j < rows.size(); j++)
      for (size_t k = 0; k < cols.size(); k++) m[j * cols.size() + k] = getMatrixEntry(rows[j], cols[k]);
    return m;
  }
};

// userkernel (kernel.hpp:37-61): a built-in kernel chosen by id.
class userkernel : public kernel {
 public:
  explicit userkernel(int id = AFMM2D_KERNEL_LOG, const double* params = nullptr) {
    kernel_id = id;
    if (params)
      for (int i = 0; i < 4; i++) kparams[i] = params[i];
  }
};

// AFMM (AFMM2D.hpp:14-33).
class AFMM {
 public:
  int N = 0;
  int MinParticlesInLeaf = 0;
  int TOL_POW = 0;
  unsigned nLevels = 0;
  std::vector<double> sorted_Locations;   // interleaved x,y in KD order
  std::vector<int64_t> sorted_Properties; // caller index of each KD-sorted point
  userkernel* mykernel = nullptr;

  // loc: N x 2, column-major (the layout of the reference's Eigen::MatrixXd)
  // unless interleaved is true.
  AFMM(int N_, int MinParticlesInLeaf_, int TOL_POW_, const std::vector<double>& loc,
       int kernel_id = AFMM2D_KERNEL_LOG, const double* kparams = nullptr,
       bool interleaved = false, int device = -1)
      : N(N_), MinParticlesInLeaf(MinParticlesInLeaf_), TOL_POW(TOL_POW_) {
    if ((int64_t)loc.size() != 2 * (int64_t)N) throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "loc size");
    check(afmm2d_create(N, MinParticlesInLeaf, TOL_POW, loc.data(),
                        interleaved ? AFMM2D_LOC_INTERLEAVED : AFMM2D_LOC_COLUMN_MAJOR, kernel_id,
                        kparams, device, &h_));
    int lv = 0;
    check(afmm2d_levels(h_, &lv));
    nLevels = (unsigned)lv;
    sorted_Locations.resize(2 * (size_t)N);
    sorted_Properties.resize(N);
    check(afmm2d_get_sorted_locations(h_, sorted_Locations.data()));
    check(afmm2d_get_permutation(h_, sorted_Properties.data()));
    mykernel = new userkernel(kernel_id, kparams);
    mykernel->a = std::ldexp(1.0, -(int)nLevels);
    if (kparams && kernel_id == AFMM2D_KERNEL_REG_INV_R && kparams[1] > 0.0) mykernel->a = kparams[1];
    mykernel->particles_X.resize(N);
    for (int i = 0; i < N; i++) mykernel->particles_X[i] = {sorted_Locations[2 * i], sorted_Locations[2 * i + 1]};
    mykernel->particles_Y = mykernel->particles_X;
  }
  AFMM(const AFMM&) = delete;
  AFMM& operator=(const AFMM&) = delete;
  ~AFMM() {
    delete mykernel;
    afmm2d_destroy(h_);
  }

  void assemble() { check(afmm2d_assemble(h_)); }

  // Dense M2L block storage, before assemble(): AFMM2D_M2L_FP64 (default) or
  // AFMM2D_M2L_BFP48 (48-bit block floating point, FP64 products).
  void setM2LStorage(int storage) { check(afmm2d_set_m2l_storage(h_, storage)); }

  // y = K b, both in the caller's point order.
  std::vector<double> computeMatVecProduct(const std::vector<double>& b) {
    if ((int)b.size() != N) throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "b size");
    std::vector<double> y(N);
    check(afmm2d_matvec(h_, b.data(), y.data(), 1));
    return y;
  }

  // The near-field term alone (FMM2DTree::evaluate_NearField); needs no assemble().
  std::vector<double> nearField(const std::vector<double>& b) {
    if ((int)b.size() != N) throw Error(AFMM2D_ERR_INVALID_ARGUMENT, "b size");
    std::vector<double> y(N);
    check(afmm2d_near_field(h_, b.data(), y.data()));
    return y;
  }

  // One leaf's relative error against the direct sum; returns it.
  double evaluateError(unsigned seed = 0) {
    int box = 0;
    double err = 0.0;
    check(afmm2d_evaluate_error(h_, seed, &box, &err));
    return err;
  }

  afmm2d_handle handle() const { return h_; }

 private:
  afmm2d_handle h_ = nullptr;
};

}  // namespace afmm2d
