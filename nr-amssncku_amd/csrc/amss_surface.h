// amss_surface.h -- Psi4's spin-weighted modes on an extraction sphere
// (surface_integral::surf_Wave, surface_integral.C:197-377).
#pragma once
#include <hip/hip_runtime.h>

namespace amss {

// One block's Psi4 and coordinates on the device.  cmin: the lowest block
// index global_interp may use per axis (1, or -2 on a symmetry plane).
struct SphereGrid {
  const double* __restrict__ rpsi;
  const double* __restrict__ ipsi;
  const double* __restrict__ X;
  const double* __restrict__ Y;
  const double* __restrict__ Z;
  int n[3];
  int cmin[3];
  double soa_r[3], soa_i[3];
};

// The sphere: n_theta Gauss-Legendre nodes in cos(theta) (upper hemisphere
// first) with their weights, n_phi midpoint nodes in phi, the reflections
// that unfold a hemisphere (lpsy 2) or an octant (8), the fields' parities.
struct SphereRule {
  const double* __restrict__ costheta;
  const double* __restrict__ weight;
  int n_theta, n_phi, lpsy;
  double dphi, rex;
  double soa_r[3], soa_i[3];
};

// vals[2n], vals[2n+1]: Rpsi4 and Ipsi4 at sphere point n.
void launch_sphere_interp(const SphereGrid& G, const SphereRule& R, double* vals, hipStream_t s);
// out[2k], out[2k+1]: real and imaginary part of mode k (m = mode_m[k]).
void launch_sphere_modes(const SphereRule& R, const double* vals, const double* thetap, const int* mode_m,
                         int nmodes, double* out, hipStream_t s);

}  // namespace amss
