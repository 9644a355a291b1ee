// amss_surface.hip -- the spin-weighted modes of Psi4 on an extraction sphere
// (surface_integral::surf_Wave, surface_integral.C:197-377, with the
// Gauss-Legendre rule in cos(theta) of :39-181): Psi4 interpolated at the
// sphere's points (global_interp, fmisc.f90:712-794) by one kernel, the
// weighted sums over the points by a second, one workgroup per mode.
#include "amss_interp.h"
#include "amss_surface.h"

namespace amss {

namespace {

// global_interp's 6^3 block (cell-centred, decide3d's mirror below a
// symmetry plane) of two fields at q, interpolated along z, y, x.
__device__ void interp2(const SphereGrid& G, const double q[3], double out[2]) {
  const double* C[3] = {G.X, G.Y, G.Z};
  int cxB[3];
  double cx[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const double d = C[m][1] - C[m][0];
    // idint((x - X(1)) / dX + 0.4) + 1 with the single-precision literal
    const int cxI = (int)((q[m] - C[m][0]) / d + (double)0.4f) + 1;
    int b = cxI - 2, t = b + 5;
    if (b < G.cmin[m]) {
      b = G.cmin[m];
      t = b + 5;
    }
    if (t > G.n[m]) {
      t = G.n[m];
      b = t - 5;
    }
    cxB[m] = b;
    cx[m] = b > 0 ? (q[m] - C[m][b - 1]) / d : (q[m] + C[m][-b]) / d;
  }
  for (int f = 0; f < 2; ++f) {
    const double* F = f == 0 ? G.rpsi : G.ipsi;
    const double* S = f == 0 ? G.soa_r : G.soa_i;
    double ym[6];
    for (int a = 0; a < 6; ++a) {
      double yz[6];
      for (int b = 0; b < 6; ++b) {
        double line[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const int g[3] = {cxB[0] + a, cxB[1] + b, cxB[2] + c};
          double s = 1.0;
          long idx = 0;
          const long st[3] = {1, G.n[0], (long)G.n[0] * G.n[1]};
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            int gi = g[m];
            if (gi < 1) {
              gi = 1 - gi;
              s *= S[m];
            }
            idx += (gi - 1) * st[m];
          }
          const double v = F[idx];
          line[c] = s == 1.0 ? v : v * s;
        }
        yz[b] = neville6(line, cx[2]);
      }
      ym[a] = neville6(yz, cx[1]);
    }
    out[f] = neville6(ym, cx[0]);
  }
}

__global__ void __launch_bounds__(256) sphere_interp_kernel(SphereGrid G, SphereRule R, double* __restrict__ vals) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= R.n_theta * R.n_phi) return;
  const int i = n / R.n_phi, j = n - i * R.n_phi;
  const double ct = R.costheta[i], st = sqrt(1.0 - ct * ct), ph = (j + 0.5) * R.dphi;
  const double q[3] = {R.rex * (st * cos(ph)), R.rex * (st * sin(ph)), R.rex * ct};
  double v[2];
  interp2(G, q, v);
  vals[2 * n] = v[0];
  vals[2 * n + 1] = v[1];
}

// One workgroup per mode k = (l, m): every thread sums its points (strided),
// then a fixed-order tree in LDS, so the result does not depend on timing.
// thetap holds sqrt((2l+1)/4pi) d^l_{m,s}(theta) per mode and theta node,
// for +cos(theta) (first n_theta) and -cos(theta) (next n_theta).
__global__ void __launch_bounds__(256) sphere_modes_kernel(SphereRule R, const double* __restrict__ vals,
                                                           const double* __restrict__ thetap,
                                                           const int* __restrict__ mode_m, double* __restrict__ out) {
  constexpr double PI = 3.14159265358979323846;
  const int k = blockIdx.x, pm = mode_m[k];
  const double* tp = thetap + (long)k * 2 * R.n_theta;
  double re = 0.0, im = 0.0;
  for (int n = threadIdx.x; n < R.n_theta * R.n_phi; n += 256) {
    const int i = n / R.n_phi, j = n - i * R.n_phi;
    const double w = R.weight[i];
    for (int lp = 0; lp < R.lpsy; ++lp) {
      // (theta or pi - theta) x (phi, 2 pi - phi, pi - phi, pi + phi) and the
      // parities of that reflection (surface_integral.C:265-326)
      const int flip = lp & 1, q = lp >> 1;
      double arg = q < 2 ? pm * (j + 0.5) * R.dphi
                         : (q == 2 ? pm * (PI - (j + 0.5) * R.dphi) : pm * (PI + (j + 0.5) * R.dphi));
      double sr = 1.0, si = 1.0;
      if (q == 1) {
        sr = R.soa_r[1];
        si = R.soa_i[1];
      } else if (q == 2) {
        sr = R.soa_r[0];
        si = R.soa_i[0];
      } else if (q == 3) {
        sr = R.soa_r[1] * R.soa_r[0];
        si = R.soa_i[1] * R.soa_i[0];
      }
      if (flip) {
        sr *= R.soa_r[2];
        si *= R.soa_i[2];
      }
      const double cm = cos(arg), sm = (q == 1 ? -1.0 : 1.0) * sin(arg);
      const double Rv = sr * vals[2 * n], Iv = si * vals[2 * n + 1];
      const double t = tp[i + (flip ? R.n_theta : 0)];
      re += t * (Rv * cm + Iv * sm) * w;
      im += t * (Iv * cm - Rv * sm) * w;
    }
  }
  __shared__ double sre[256], sim[256];
  sre[threadIdx.x] = re;
  sim[threadIdx.x] = im;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sre[threadIdx.x] += sre[threadIdx.x + h];
      sim[threadIdx.x] += sim[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * k] = sre[0] * R.rex * R.dphi;
    out[2 * k + 1] = sim[0] * R.rex * R.dphi;
  }
}

}  // namespace

void launch_sphere_interp(const SphereGrid& G, const SphereRule& R, double* vals, hipStream_t s) {
  const int n = R.n_theta * R.n_phi;
  if (n > 0) sphere_interp_kernel<<<(n + 255) / 256, 256, 0, s>>>(G, R, vals);
}

void launch_sphere_modes(const SphereRule& R, const double* vals, const double* thetap, const int* mode_m,
                         int nmodes, double* out, hipStream_t s) {
  if (nmodes > 0) sphere_modes_kernel<<<nmodes, 256, 0, s>>>(R, vals, thetap, mode_m, out);
}

}  // namespace amss
