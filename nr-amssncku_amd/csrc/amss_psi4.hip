// amss_psi4.hip -- the Newman-Penrose Weyl scalar Psi4 of the BSSN state
// (getnp4, getnp4.f90:10-452 with the radial tetrad of tetradtype 2): the
// electric and magnetic parts of the Weyl tensor, E_ij = R_ij - K_ik K^k_j
// + K K_ij and B_ij = sym(v^k (D_k K_ij - D_j K_ik)), contracted with
// m = (u - i w) / sqrt(2), where (v, u, w) is the Gram-Schmidt frame of the
// radial, azimuthal and polar directions in the physical metric:
// Psi4 = -(E - B)_ij m^i m^j.  One thread per point, the first derivatives
// with the reference's order rules at the edges (the boundary environment).
#include "amss_psi4.h"
#include "bssn_env.h"

namespace amss {

namespace {

// x^T M y for a symmetric M stored in sidx order
__device__ __forceinline__ double quad(const double M[6], const double x[3], const double y[3]) {
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) s += M[sidx(a, b)] * x[a] * y[b];
  return s;
}

// gradient of field v at the point with the centred order its room allows
// (diff_new.f90:1084-1096), mirror reads below a symmetry plane
__device__ void point_grad(const FieldsIn& F, const Coef& C, const Box& box, int i, int j, int k, int ord, int v,
                           double d[3]) {
  if (ord == 2) {
    EdgeEnv<2> E(F, F, C, i, j, k, box.n, box.mlo, box.sy, box.sz);
    E.grad(v, d);
  } else if (ord == 1) {
    EdgeEnv<1> E(F, F, C, i, j, k, box.n, box.mlo, box.sy, box.sz);
    E.grad(v, d);
  } else {
    d[0] = d[1] = d[2] = 0.0;
  }
}

}  // namespace

__global__ void __launch_bounds__(256) psi4_kernel(FieldsIn st, Psi4Io io, Box box, Coef C) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  const long p = blockIdx.x * 256L + threadIdx.x;
  if (p >= N) return;
  const int i = (int)(p % box.n[0]), j = (int)((p / box.n[0]) % box.n[1]), k = (int)(p / ((long)box.n[0] * box.n[1]));
  // the joint room test of fderivs: +-2 on every axis for fourth order, +-1 for second
  int room = 3;
  const int q[3] = {i, j, k};
#pragma unroll
  for (int a = 0; a < 3; ++a) room = min(room, min(q[a] - box.mlo[a], box.n[a] - 1 - q[a]));
  const int ord = room >= 2 ? 2 : (room >= 1 ? 1 : 0);

  double g[6], A[6], gu[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    g[s] = st.p[DXX + s][p] + (sdiag(s) ? 1.0 : 0.0);
    A[s] = st.p[AXX + s][p];
  }
  const double det = g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
                     g[2] * (g[1] * g[4] - g[3] * g[2]);
  gu[0] = (g[3] * g[5] - g[4] * g[4]) / det;
  gu[1] = (g[2] * g[4] - g[1] * g[5]) / det;
  gu[2] = (g[1] * g[4] - g[3] * g[2]) / det;
  gu[3] = (g[0] * g[5] - g[2] * g[2]) / det;
  gu[4] = (g[1] * g[2] - g[0] * g[4]) / det;
  gu[5] = (g[0] * g[3] - g[1] * g[1]) / det;
  const double f = 1.0 / (st.p[CHI][p] + 1.0);  // 1 / chi: gamma_ij = g_ij / chi
  const double K = st.p[TRK][p];

  // the frame: v radial (raised), u azimuthal, w polar, orthonormalized in
  // gamma in that order; a point on the z axis (or at the origin) takes the
  // reference's tiny stand-ins (getnp4.f90:259-341)
  constexpr double kTiny = 1.e-14;
  const double x = io.X[i], y = io.Y[j], z = io.Z[k];
  const bool axis = fabs(x) < kTiny && fabs(y) < kTiny;
  const double r0[3] = {axis && fabs(z) < kTiny ? kTiny : x, axis && fabs(z) < kTiny ? kTiny : y,
                        axis && fabs(z) < kTiny ? kTiny : z};
  double v[3], u[3], w[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) v[a] = gu[sidx(a, 0)] * r0[0] + gu[sidx(a, 1)] * r0[1] + gu[sidx(a, 2)] * r0[2];
  u[0] = axis ? -kTiny : -y;
  u[1] = axis ? kTiny : x;
  u[2] = 0.0;
  w[0] = axis ? kTiny * z : x * z;
  w[1] = axis ? kTiny * z : y * z;
  w[2] = axis ? -2.0 * kTiny * kTiny : -(x * x + y * y);
  auto normalize = [&](double e[3]) {
    const double n = sqrt(quad(g, e, e) * f);
#pragma unroll
    for (int a = 0; a < 3; ++a) e[a] /= n;
  };
  auto project_out = [&](double e[3], const double b[3]) {
    const double c = quad(g, b, e) * f;
#pragma unroll
    for (int a = 0; a < 3; ++a) e[a] -= c * b[a];
  };
  normalize(v);
  project_out(u, v);
  normalize(u);
  project_out(w, v);
  project_out(w, u);
  normalize(w);

  // T_ijk = chi D_k K_ij: d_k A_ij - Gamma^m_ki A_mj - Gamma^m_kj A_im
  //         - (d_k chi / chi) A_ij + g_ij d_k K / 3
  double dchi[3], dK[3];
  point_grad(st, C, box, i, j, k, ord, CHI, dchi);
  point_grad(st, C, box, i, j, k, ord, TRK, dK);
  double T[6][3];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    point_grad(st, C, box, i, j, k, ord, AXX + s, T[s]);
    const int a = s == 0 || s == 1 || s == 2 ? 0 : (s == 5 ? 2 : 1);
    const int b = s == 0 ? 0 : (s == 1 || s == 3 ? 1 : 2);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double gA = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m)
        gA += io.gam[6 * m + sidx(c, a)][p] * A[sidx(m, b)] + io.gam[6 * m + sidx(c, b)][p] * A[sidx(a, m)];
      T[s][c] = T[s][c] - gA - dchi[c] * f * A[s] + (1.0 / 3.0) * g[s] * dK[c];
    }
  }
  // E_ij - B_ij
  double EB[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int a = s == 0 || s == 1 || s == 2 ? 0 : (s == 5 ? 2 : 1);
    const int b = s == 0 ? 0 : (s == 1 || s == 3 ? 1 : 2);
    double B = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) B += v[c] * (T[s][c] - 0.5 * (T[sidx(a, c)][b] + T[sidx(b, c)][a]));
    // chi K_ij = A_ij + g_ij K / 3
    double KK = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e)
        KK += gu[sidx(c, e)] * (A[sidx(a, c)] + g[sidx(a, c)] * K / 3.0) * (A[sidx(b, e)] + g[sidx(b, e)] * K / 3.0);
    const double Kab = A[s] + g[s] * K / 3.0;
    EB[s] = io.ric[s][p] - (KK - Kab * K) * f - B * f;
  }
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      re += EB[sidx(a, b)] * (u[a] * u[b] - w[a] * w[b]);
      im += EB[sidx(a, b)] * (u[a] * w[b] + w[a] * u[b]);
    }
  io.rpsi[p] = -0.5 * re;
  io.ipsi[p] = -0.5 * im;
}

void launch_psi4(const FieldsIn& st, const Psi4Io& io, const Box& box, const Coef& C, hipStream_t s) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  if (N > 0) psi4_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(st, io, box, C);
}

}  // namespace amss
