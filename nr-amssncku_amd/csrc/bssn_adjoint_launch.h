// bssn_adjoint_launch.h -- launchers of the adjoint kernels (bssn_adjoint.hip).
#pragma once
#include "bssn_kernels.h"

namespace amss {

// RHS k at the interior points (no auxiliary outputs), GAUGE 0.
void adj_launch_rhs_k(const FieldsIn& in, const FieldsOut& out, const Box& box, const Coef& C, hipStream_t s);
// Reverse of the pointwise algebra: feature bars at interior points (feature
// f of point p at bars[f * stride + p]).  feat (kAdjInputs x N doubles): work
// array for the stencil values, computed by a separate kernel.
void adj_launch_point(const FieldsIn& in, const FieldsIn& lam, double* bars, long stride, const Box& box,
                      const Coef& C, hipStream_t s, double* feat);
// Transposed stencils: out (= or +=) dJ/d(in) at every point.  The bars
// arrays must be zero outside the interior and readable 3 planes beyond the
// array on either side; so must lam when lam_padded (else lam is read only
// inside the array, and only its interior values are used).
void adj_launch_gather(const FieldsIn& in, const FieldsIn& lam, bool lam_padded, const double* bars, long stride,
                       const FieldsOut& out, const Box& box, const Coef& C, bool accumulate, hipStream_t s);
void adj_launch_fwd_stage(const FieldsIn& Y0, const FieldsIn& k, const FieldsOut& acc, const FieldsOut& z,
                          const FieldsOut& Yn, const Box& box, int stage, double c, double tinny, hipStream_t s);
void adj_launch_final(const FieldsIn& lam, const FieldsIn& z3, const FieldsIn& Y0, const FieldsOut& lY0,
                      const FieldsOut* lk, const Box& box, double c, double tinny, hipStream_t s);
void adj_launch_stage(const FieldsIn& lY, const FieldsIn& z, const FieldsIn& Y0, const FieldsOut& lY0,
                      const FieldsOut& lk, const Box& box, double c, double tinny, hipStream_t s);
void adj_launch_open(const FieldsIn& U, const FieldsIn& lY0, const FieldsOut& lam, long N, hipStream_t s);
void adj_launch_copy_enforce(const FieldsIn& U, const FieldsOut& Y, long N, hipStream_t s);
void adj_launch_copy_fields(const FieldsIn& src, const FieldsOut& dst, long N, hipStream_t s);

}  // namespace amss
