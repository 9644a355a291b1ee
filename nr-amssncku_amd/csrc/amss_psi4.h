// amss_psi4.h -- the Newman-Penrose Weyl scalar Psi4 on device arrays
// (getnp4.f90:10-452, tetradtype 2).
#pragma once
#include <hip/hip_runtime.h>

#include "bssn_kernels.h"

namespace amss {

// The physical Christoffels Gamma^i_jk (18, compute_rhs_bssn's order:
// i = x, y, z, then jk = xx xy xz yy yz zz) and Ricci R_ij (6) of the state,
// the coordinates of the array's points, and the two outputs.
struct Psi4Io {
  const double* __restrict__ gam[18];
  const double* __restrict__ ric[6];
  const double* __restrict__ X;
  const double* __restrict__ Y;
  const double* __restrict__ Z;
  double* __restrict__ rpsi;
  double* __restrict__ ipsi;
};

// Psi4 at every point of box.n (box.mlo: -3 on a symmetry-plane axis).
void launch_psi4(const FieldsIn& st, const Psi4Io& io, const Box& box, const Coef& C, hipStream_t s);

}  // namespace amss
