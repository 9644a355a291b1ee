// amss_adjoint.h -- discrete adjoint of one RK4 step of one Block
// (bssn_class::Step with frozen ghost zones, GAUGE 0, vacuum), on device
// arrays: the stage states are re-evaluated from U^n, then the step is run
// backwards through the RK4 combination, chi's lower bound, enforce_ga and
// the right-hand side (reverse algebra + transposed stencils).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "bssn_kernels.h"

namespace amss {

class AdjointStep {
 public:
  AdjointStep(const int ex[3], const Coef& C, double chitiny);
  ~AdjointStep();
  bool ok() const { return err_.empty(); }
  const std::string& err() const { return err_; }
  // U^n -> U^{n+1} (Unext may be null: only the stage data are kept)
  int forward(double dT, const double* const* U, double* const* Unext, hipStream_t s);
  // lam: dJ/dU^{n+1} in, dJ/dU^n out (runs forward() first)
  int backward(double dT, const double* const* U, double* const* lam, hipStream_t s);
  // dJ/dU of sum_p lam(p) . rhs(U)(p) over the interior points
  int rhs_vjp(const double* const* U, const double* const* lam, double* const* out, hipStream_t s);
  long npts() const { return N_; }

 private:
  FieldsIn fin(int first) const;
  FieldsOut fout(int first) const;
  Box box_;
  Coef C_;
  double tinny_;
  long N_ = 0;
  double* buf_ = nullptr;  // all work arrays, kFields x N
  std::string err_;
};

}  // namespace amss
