// bssn_gauge0.hip -- instantiation of the point kernels for GAUGE 0.
#include "bssn_kernels_impl.h"

namespace amss {

void launch_rhs_g0(const FieldsIn& in, const FieldsIn& mat, const FieldsOut& out, const AuxOut& aux, const Box& box,
                   const Coef& C, int co, int* flag, bool edge, hipStream_t s) {
  launch_rhs_gauge<0>(in, mat, out, aux, box, C, co, flag, edge, s);
}

void launch_rk4_g0(const FieldsIn& in, const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                   const AuxOut& aux, const Box& box, const Coef& C, const Rk4Param& R, bool want_res, int* flag,
                   bool edge, hipStream_t s, bool interior, const SommerGeom& sg) {
  launch_rk4_gauge<0>(in, u0, acc, out, aux, box, C, R, want_res, flag, edge, s, interior, sg);
}

}  // namespace amss
