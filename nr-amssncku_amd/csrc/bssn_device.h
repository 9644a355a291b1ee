// bssn_device.h -- device-side building blocks of the BSSN right-hand side
// for gfx950: field layout, stencil environments and the pointwise algebra.
//
// The right-hand side is the one of compute_rhs_bssn (reference
// bssn_rhs.f90:5-1186, GAUGE 0 / GAUGE 2 branches), restated here in
// tensor-index form: every quantity is a small symmetric tensor evaluated
// per grid point in registers, and every stencil is applied to the stored
// input fields only, so one thread evaluates one grid point from the
// neighbourhood of the 24 inputs without any full-size temporaries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amss {

// Evolved fields, in the reference registry order (bssn_class.C:437-460).
enum Field : int {
  CHI = 0, TRK, DXX, GXY, GXZ, DYY, GYZ, DZZ, AXX, AXY, AXZ, AYY, AYZ, AZZ,
  GAMX, GAMY, GAMZ, LAP, BETAX, BETAY, BETAZ, DTSFX, DTSFY, DTSFZ, NFIELD
};
// Matter fields (rho, S_i, S_ij) in the reference argument order.
enum Matter : int { RHO = 0, SX, SY, SZ, SXX, SXY, SXZ, SYY, SYZ, SZZ, NMATTER };

// Symmetric 3x3 storage order xx xy xz yy yz zz (the reference's order for
// g_ij, A_ij, S_ij, R_ij and for the second-derivative outputs).
__host__ __device__ constexpr int sidx(int a, int b) {
  return a == b ? (a == 0 ? 0 : (a == 1 ? 3 : 5)) : (a + b == 1 ? 1 : (a + b == 2 ? 2 : 4));
}
__host__ __device__ constexpr bool sdiag(int s) { return s == 0 || s == 3 || s == 5; }

// Reflection parity of each field under x, y, z (bssn_class.C:183-361).
__host__ __device__ constexpr int parity(int v, int axis) {
  // bit set = odd under reflection of that axis
  // chi trK dxx gxy gxz dyy gyz dzz Axx Axy Axz Ayy Ayz Azz Gx Gy Gz Lap bx by bz Bx By Bz
  constexpr unsigned char odd[NFIELD] = {0, 0, 0, 3, 5, 0, 6, 0, 0, 3, 5, 0, 6, 0,
                                         1, 2, 4, 0, 1, 2, 4, 1, 2, 4};
  return (odd[v] >> axis) & 1 ? -1 : 1;
}

// The 24 fields of one time level as explicit pointers (kernel arguments).
struct FieldsIn {
  const double* __restrict__ p[NFIELD];
  __device__ __forceinline__ const double* ptr(int v) const { return p[v]; }
};
struct FieldsOut {
  double* __restrict__ p[NFIELD];
  __device__ __forceinline__ double* ptr(int v) const { return p[v]; }
};
// A double in global memory.  Pointers read from device tables are generic
// to the compiler, which then emits flat loads; flat operations count against
// the LDS wait counter too, so every LDS wait would also wait for them.
typedef __attribute__((address_space(1))) double gdouble;

// The 24 fields of one time level inside a Block's slab: field v starts at
// base + v * n (a Block keeps its lists as consecutive 24-field runs).
struct SlabFields {
  gdouble* base;
  long n;
  __device__ __forceinline__ gdouble* ptr(int v) const { return base + v * n; }
};

// Finite-difference coefficients of one patch (host-computed from
// dX = X[1]-X[0] etc., bssn_rhs.f90:142-144).
struct Coef {
  double h[3];
  double d1[3];     // 1/(12 h)          4th-order first derivative
  double d1lo[3];   // 1/(2 h)           2nd-order first derivative
  double d2[3];     // 1/(12 h^2)        4th-order second derivative
  double d2lo[3];   // 1/h^2
  double dm[3];     // 1/(144 h_a h_b)   4th-order mixed (xy, xz, yz)
  double dmlo[3];   // 1/(4 h_a h_b)
  double ko[3];     // eps/(64 h)        Kreiss-Oliger
};

// Geometry of one array: shape (with ghosts) and strides in doubles.
struct Shape {
  int n[3];
  long sy, sz;
};

}  // namespace amss
