// amss_xfer.hip -- inter-level transfer between 2:1 cell-centred grids: the
// 6-point, 5th-order prolongation (coarse -> fine) and restriction (fine ->
// coarse) of prolongrestrict_cell.f90 (Cell, ghost_width 3; prolong3
// :1915-2164, restrict3 :2330-2468).  One thread per target point; the
// 6x6x6 source block is contracted along z, then y, then x, with the
// reference's weights and summation order.
#include "amss_xfer.h"

#include <cmath>

namespace amss {

// Cell numbering of both grids from a common base point (:1958-1990): the
// reference rounds (x - base) / h + 0.4 toward zero, 0.4 being a
// single-precision literal.
static inline int cell_index(double x, double base, double h) {
  return (int)((x - base) / h + (double)0.4f);
}

int xfer_plan(const double* llbc, const double* uubc, const int* extc, const double* llbf, const double* uubf,
              const int* extf, const double* ll, const double* uu, bool prolong, XferPlan& P) {
  double CD[3], FD[3], base[3];
  for (int a = 0; a < 3; ++a) {
    if (extc[a] < 1 || extf[a] < 1) return 2;
    CD[a] = (uubc[a] - llbc[a]) / extc[a];
    FD[a] = (uubf[a] - llbf[a]) / extf[a];
    if (std::fabs(CD[a] - 2 * FD[a]) > 1e-10) return 2;  // not 2:1 (the reference stops)
    // a fine box starting inside a coarse cell shifts the base by half a
    // coarse cell when it starts on an odd fine cell (:1965-1977)
    if (llbc[a] <= llbf[a]) {
      base[a] = llbc[a];
    } else {
      const int j = cell_index(llbc[a], llbf[a], FD[a]);
      base[a] = (j % 2 == 0) ? llbf[a] : llbf[a] - CD[a] / 2;
    }
  }
  for (int a = 0; a < 3; ++a) {
    P.ext_c[a] = extc[a];
    P.ext_f[a] = extf[a];
    P.lbf[a] = cell_index(llbf[a], base[a], FD[a]) + 1;
    P.lbc[a] = cell_index(llbc[a], base[a], CD[a]) + 1;
    const double ht = prolong ? FD[a] : CD[a], hs = prolong ? CD[a] : FD[a];
    const int lt = cell_index(ll[a], base[a], ht) + 1, ut = cell_index(uu[a], base[a], ht);
    const int ls = cell_index(ll[a], base[a], hs) + 1, us = cell_index(uu[a], base[a], hs);
    const int tl = prolong ? P.lbf[a] : P.lbc[a], sl = prolong ? P.lbc[a] : P.lbf[a];
    const int nt = prolong ? extf[a] : extc[a], ns = prolong ? extc[a] : extf[a];
    // target range, 1-based; the reference's sanity test on both ranges
    // (:2001-2021, 2416-2427)
    P.lo[a] = lt - tl + 1;
    P.hi[a] = ut - tl + 1;
    if (P.lo[a] < 1 || ls - sl + 1 < 1 || P.hi[a] > nt || us - sl + 1 > ns - 2) return 1;
  }
  for (int a = 0; a < 3; ++a)
    if (P.hi[a] < P.lo[a]) return 0;  // an empty region: nothing to do
  for (int a = 0; a < 3; ++a) {
    const int ns = prolong ? extc[a] : extf[a];
    // every stencil inside the source and its mirrored halo (prolong:
    // funcc(-2:extc), restrict: funff(-1:extf)); the reference only
    // prints a warning there and reads outside its arrays
    const int c0 = prolong ? (P.lo[a] + P.lbf[a] - 1) / 2 - P.lbc[a] + 1 : 2 * (P.lo[a] + P.lbc[a] - 1) - P.lbf[a];
    const int c1 = prolong ? (P.hi[a] + P.lbf[a] - 1) / 2 - P.lbc[a] + 1 : 2 * (P.hi[a] + P.lbc[a] - 1) - P.lbf[a];
    if (c0 < (prolong ? 0 : 1) || c1 + 3 > ns) return 1;
  }
  return 0;
}

// The source value at 1-based index q (symmetry_bd's mirror below 1:
// q -> 1 - q with the parity of each mirrored axis, fmisc.f90:877-901).
struct XferSrc {
  const double* __restrict__ f;
  int n0, n1;
  double S[3];
  __device__ __forceinline__ double at(int i, int j, int k) const {
    double s = 1.0;
    if (i < 1) { i = 1 - i; s *= S[0]; }
    if (j < 1) { j = 1 - j; s *= S[1]; }
    if (k < 1) { k = 1 - k; s *= S[2]; }
    return f[(i - 1) + (long)n0 * ((j - 1) + (long)n1 * (k - 1))] * s;
  }
};

__global__ void __launch_bounds__(256) prolong3_kernel(XferSrc src, double* __restrict__ funf, XferPlan P) {
  const int e0 = P.hi[0] - P.lo[0] + 1, e1 = P.hi[1] - P.lo[1] + 1, e2 = P.hi[2] - P.lo[2] + 1;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)e0 * e1 * e2) return;
  const int p[3] = {P.lo[0] + (int)(t % e0), P.lo[1] + (int)((t / e0) % e1), P.lo[2] + (int)(t / ((long)e0 * e1))};
  // Lagrange weights at a quarter (even fine cell) or three quarters (odd)
  // of the coarse cell c toward c + 1, over coarse cells c-2 .. c+3
  constexpr double W[6] = {77.0 / 8192.0, -693.0 / 8192.0, 3465.0 / 4096.0,
                           1155.0 / 4096.0, -495.0 / 8192.0, 63.0 / 8192.0};
  int c[3];
  bool rev[3];
  for (int a = 0; a < 3; ++a) {
    const int g = p[a] + P.lbf[a] - 1;
    c[a] = g / 2 - P.lbc[a] + 1;
    rev[a] = (g % 2) != 0;
  }
  double row[6];
  for (int a = 0; a < 6; ++a) {
    double col = 0.0;
    for (int b = 0; b < 6; ++b) {
      double s = 0.0;
      for (int q = 0; q < 6; ++q) s += W[rev[2] ? 5 - q : q] * src.at(c[0] - 2 + a, c[1] - 2 + b, c[2] - 2 + q);
      col += W[rev[1] ? 5 - b : b] * s;
    }
    row[a] = col;
  }
  double r = 0.0;
  for (int a = 0; a < 6; ++a) r += W[rev[0] ? 5 - a : a] * row[a];
  funf[(p[0] - 1) + (long)P.ext_f[0] * ((p[1] - 1) + (long)P.ext_f[1] * (p[2] - 1))] = r;
}

__global__ void __launch_bounds__(256) restrict3_kernel(XferSrc src, double* __restrict__ func, XferPlan P) {
  const int e0 = P.hi[0] - P.lo[0] + 1, e1 = P.hi[1] - P.lo[1] + 1, e2 = P.hi[2] - P.lo[2] + 1;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)e0 * e1 * e2) return;
  const int p[3] = {P.lo[0] + (int)(t % e0), P.lo[1] + (int)((t / e0) % e1), P.lo[2] + (int)(t / ((long)e0 * e1))};
  // midpoint weights, symmetric: pairs (0,5), (1,4), (2,3) of fine cells c-2 .. c+3
  constexpr double W[3] = {3.0 / 256.0, -25.0 / 256.0, 75.0 / 128.0};
  int c[3];
  for (int a = 0; a < 3; ++a) c[a] = 2 * (p[a] + P.lbc[a] - 1) - P.lbf[a];
  double row[6];
  for (int a = 0; a < 6; ++a) {
    double cols[6];
    for (int b = 0; b < 6; ++b) {
      double v[6];
      for (int q = 0; q < 6; ++q) v[q] = src.at(c[0] - 2 + a, c[1] - 2 + b, c[2] - 2 + q);
      cols[b] = W[0] * (v[0] + v[5]) + W[1] * (v[1] + v[4]) + W[2] * (v[2] + v[3]);
    }
    row[a] = W[0] * (cols[0] + cols[5]) + W[1] * (cols[1] + cols[4]) + W[2] * (cols[2] + cols[3]);
  }
  func[(p[0] - 1) + (long)P.ext_c[0] * ((p[1] - 1) + (long)P.ext_c[1] * (p[2] - 1))] =
      W[0] * (row[0] + row[5]) + W[1] * (row[1] + row[4]) + W[2] * (row[2] + row[3]);
}

// target points of a plan; 0 when any axis is empty (nothing is launched)
static long plan_points(const XferPlan& P) {
  long n = 1;
  for (int a = 0; a < 3; ++a) n *= P.hi[a] >= P.lo[a] ? (long)(P.hi[a] - P.lo[a] + 1) : 0L;
  return n;
}

void launch_prolong3(const double* func, double* funf, const XferPlan& P, const double S[3], hipStream_t s) {
  const long n = plan_points(P);
  XferSrc src{func, P.ext_c[0], P.ext_c[1], {S[0], S[1], S[2]}};
  if (n > 0) prolong3_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, funf, P);
}

void launch_restrict3(const double* funf, double* func, const XferPlan& P, const double S[3], hipStream_t s) {
  const long n = plan_points(P);
  XferSrc src{funf, P.ext_f[0], P.ext_f[1], {S[0], S[1], S[2]}};
  if (n > 0) restrict3_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, func, P);
}

}  // namespace amss
