// amss_xfer.hip -- inter-level transfer between 2:1 cell-centred grids: the
// 6-point, 5th-order prolongation (coarse -> fine) and restriction (fine ->
// coarse) of prolongrestrict_cell.f90 (Cell, ghost_width 3; prolong3
// :1915-2164, restrict3 :2330-2468), on LDS-staged tiles of target cells.
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

// Both transfers work on tiles of target cells: the source block a tile
// needs is staged once in LDS and the 6x6x6 contraction is factorized (along
// z for every (y, x) of the block, then along y, then along x), which is
// the reference's summation order.  Cells past the end of the target range
// in a partial tile are computed from unloaded LDS and not stored; every LDS
// index stays inside the fixed block sizes (shown at each size below).
constexpr double kProW[6] = {77.0 / 8192.0, -693.0 / 8192.0, 3465.0 / 4096.0,
                             1155.0 / 4096.0, -495.0 / 8192.0, 63.0 / 8192.0};
constexpr double kResW[3] = {3.0 / 256.0, -25.0 / 256.0, 75.0 / 128.0};

// prolong: fine tile 16 x 8 x 8.  Fine cell p reads coarse cells
// c(p) - 2 .. c(p) + 3, c(p) = (p + lbf - 1) / 2 - lbc + 1; over a tile c
// advances by at most (F - 1 + 1) / 2 = F / 2, so a block of F / 2 + 6 coarse
// cells per axis covers it (index c(p) - c(p0) + q <= F / 2 + 5).
constexpr int kPFX = 16, kPFY = 8, kPFZ = 8;
constexpr int kPCX = kPFX / 2 + 6, kPCY = kPFY / 2 + 6, kPCZ = kPFZ / 2 + 6;

__global__ void __launch_bounds__(256) prolong3_kernel(XferSrc src, double* __restrict__ funf, XferPlan P) {
  __shared__ double B[kPCZ][kPCY][kPCX];
  __shared__ double T1[kPFZ][kPCY][kPCX];
  __shared__ double T2[kPFZ][kPFY][kPCX];
  const int F[3] = {kPFX, kPFY, kPFZ};
  int nt[3], p0[3], pe[3], c0[3], par0[3];
  for (int a = 0; a < 3; ++a) nt[a] = (P.hi[a] - P.lo[a] + F[a]) / F[a];
  const int b = blockIdx.x;
  const int tb[3] = {b % nt[0], (b / nt[0]) % nt[1], b / (nt[0] * nt[1])};
  for (int a = 0; a < 3; ++a) {
    p0[a] = P.lo[a] + tb[a] * F[a];
    pe[a] = min(p0[a] + F[a] - 1, P.hi[a]);
    c0[a] = (p0[a] + P.lbf[a] - 1) / 2 - P.lbc[a] + 1;
    par0[a] = p0[a] + P.lbf[a] - 1;  // fine global index of p0 (its parity picks the weights)
  }
  const int tid = threadIdx.x;
  // the coarse cells the tile's valid fine cells read, mirrored below 1
  int nc[3];
  for (int a = 0; a < 3; ++a) nc[a] = ((pe[a] + P.lbf[a] - 1) / 2 - P.lbc[a] + 1) + 3 - (c0[a] - 2) + 1;
  for (int e = tid; e < nc[0] * nc[1] * nc[2]; e += 256) {
    const int x = e % nc[0], y = (e / nc[0]) % nc[1], z = e / (nc[0] * nc[1]);
    B[z][y][x] = src.at(c0[0] - 2 + x, c0[1] - 2 + y, c0[2] - 2 + z);
  }
  __syncthreads();
  // offset of fine cell d of the tile in the block, and its weight order
  auto off = [&](int a, int d) { return (par0[a] + d) / 2 - par0[a] / 2; };
  auto rev = [&](int a, int d) { return ((par0[a] + d) & 1) != 0; };
  for (int e = tid; e < kPFZ * kPCY * kPCX; e += 256) {
    const int x = e % kPCX, y = (e / kPCX) % kPCY, fz = e / (kPCX * kPCY);
    const int o = off(2, fz);
    const bool r = rev(2, fz);
    double s = 0.0;
    for (int q = 0; q < 6; ++q) s += kProW[r ? 5 - q : q] * B[o + q][y][x];
    T1[fz][y][x] = s;
  }
  __syncthreads();
  for (int e = tid; e < kPFZ * kPFY * kPCX; e += 256) {
    const int x = e % kPCX, fy = (e / kPCX) % kPFY, fz = e / (kPCX * kPFY);
    const int o = off(1, fy);
    const bool r = rev(1, fy);
    double s = 0.0;
    for (int q = 0; q < 6; ++q) s += kProW[r ? 5 - q : q] * T1[fz][o + q][x];
    T2[fz][fy][x] = s;
  }
  __syncthreads();
  for (int e = tid; e < kPFZ * kPFY * kPFX; e += 256) {
    const int fx = e % kPFX, fy = (e / kPFX) % kPFY, fz = e / (kPFX * kPFY);
    const int o = off(0, fx);
    const bool r = rev(0, fx);
    double s = 0.0;
    for (int q = 0; q < 6; ++q) s += kProW[r ? 5 - q : q] * T2[fz][fy][o + q];
    const int i = p0[0] + fx, j = p0[1] + fy, k = p0[2] + fz;
    if (i <= pe[0] && j <= pe[1] && k <= pe[2])
      funf[(i - 1) + (long)P.ext_f[0] * ((j - 1) + (long)P.ext_f[1] * (k - 1))] = s;
  }
}

// restrict: coarse tile 8 x 8 x 4.  Coarse cell p reads fine cells
// f(p) - 2 .. f(p) + 3, f(p) = 2 (p + lbc - 1) - lbf; over a tile f advances
// by 2 per cell, so a block of 2 C + 4 fine cells per axis covers it
// (index 2 d + q <= 2 C + 3).
constexpr int kRCX = 8, kRCY = 8, kRCZ = 4;
constexpr int kRFX = 2 * kRCX + 4, kRFY = 2 * kRCY + 4, kRFZ = 2 * kRCZ + 4;

__global__ void __launch_bounds__(256) restrict3_kernel(XferSrc src, double* __restrict__ func, XferPlan P) {
  __shared__ double B[kRFZ][kRFY][kRFX];
  __shared__ double T1[kRCZ][kRFY][kRFX];
  __shared__ double T2[kRCZ][kRCY][kRFX];
  const int Ct[3] = {kRCX, kRCY, kRCZ};
  int nt[3], p0[3], pe[3], f0[3];
  for (int a = 0; a < 3; ++a) nt[a] = (P.hi[a] - P.lo[a] + Ct[a]) / Ct[a];
  const int b = blockIdx.x;
  const int tb[3] = {b % nt[0], (b / nt[0]) % nt[1], b / (nt[0] * nt[1])};
  for (int a = 0; a < 3; ++a) {
    p0[a] = P.lo[a] + tb[a] * Ct[a];
    pe[a] = min(p0[a] + Ct[a] - 1, P.hi[a]);
    f0[a] = 2 * (p0[a] + P.lbc[a] - 1) - P.lbf[a];
  }
  const int tid = threadIdx.x;
  int nf[3];
  for (int a = 0; a < 3; ++a) nf[a] = 2 * (pe[a] - p0[a]) + 6;
  for (int e = tid; e < nf[0] * nf[1] * nf[2]; e += 256) {
    const int x = e % nf[0], y = (e / nf[0]) % nf[1], z = e / (nf[0] * nf[1]);
    B[z][y][x] = src.at(f0[0] - 2 + x, f0[1] - 2 + y, f0[2] - 2 + z);
  }
  __syncthreads();
  for (int e = tid; e < kRCZ * kRFY * kRFX; e += 256) {
    const int x = e % kRFX, y = (e / kRFX) % kRFY, cz = e / (kRFX * kRFY);
    const double* v = &B[2 * cz][y][x];
    constexpr int S = kRFY * kRFX;
    T1[cz][y][x] = kResW[0] * (v[0] + v[5 * S]) + kResW[1] * (v[S] + v[4 * S]) + kResW[2] * (v[2 * S] + v[3 * S]);
  }
  __syncthreads();
  for (int e = tid; e < kRCZ * kRCY * kRFX; e += 256) {
    const int x = e % kRFX, cy = (e / kRFX) % kRCY, cz = e / (kRFX * kRCY);
    const double* v = &T1[cz][2 * cy][x];
    constexpr int S = kRFX;
    T2[cz][cy][x] = kResW[0] * (v[0] + v[5 * S]) + kResW[1] * (v[S] + v[4 * S]) + kResW[2] * (v[2 * S] + v[3 * S]);
  }
  __syncthreads();
  {
    const int cx = tid % kRCX, cy = (tid / kRCX) % kRCY, cz = tid / (kRCX * kRCY);
    const double* v = &T2[cz][cy][2 * cx];
    const double r = kResW[0] * (v[0] + v[5]) + kResW[1] * (v[1] + v[4]) + kResW[2] * (v[2] + v[3]);
    const int i = p0[0] + cx, j = p0[1] + cy, k = p0[2] + cz;
    if (i <= pe[0] && j <= pe[1] && k <= pe[2])
      func[(i - 1) + (long)P.ext_c[0] * ((j - 1) + (long)P.ext_c[1] * (k - 1))] = r;
  }
}

// target points of a plan; 0 when any axis is empty (nothing is launched)
static long plan_points(const XferPlan& P) {
  long n = 1;
  for (int a = 0; a < 3; ++a) n *= P.hi[a] >= P.lo[a] ? (long)(P.hi[a] - P.lo[a] + 1) : 0L;
  return n;
}

// tiles of size T covering the target range of a plan
static unsigned plan_tiles(const XferPlan& P, const int T[3]) {
  long n = 1;
  for (int a = 0; a < 3; ++a) n *= (P.hi[a] - P.lo[a] + T[a]) / T[a];
  return (unsigned)n;
}

void launch_prolong3(const double* func, double* funf, const XferPlan& P, const double S[3], hipStream_t s) {
  const int T[3] = {kPFX, kPFY, kPFZ};
  XferSrc src{func, P.ext_c[0], P.ext_c[1], {S[0], S[1], S[2]}};
  if (plan_points(P) > 0) prolong3_kernel<<<plan_tiles(P, T), 256, 0, s>>>(src, funf, P);
}

void launch_restrict3(const double* funf, double* func, const XferPlan& P, const double S[3], hipStream_t s) {
  const int T[3] = {kRCX, kRCY, kRCZ};
  XferSrc src{funf, P.ext_f[0], P.ext_f[1], {S[0], S[1], S[2]}};
  if (plan_points(P) > 0) restrict3_kernel<<<plan_tiles(P, T), 256, 0, s>>>(src, func, P);
}

}  // namespace amss
