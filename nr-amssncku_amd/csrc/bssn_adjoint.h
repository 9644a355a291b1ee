// bssn_adjoint.h -- reverse mode of the BSSN right-hand side at one point.
//
// The RHS at point p is a function of "features" of the input fields at p:
// centre values, first derivatives, second derivatives (stencils), plus the
// advection and KO terms, which enter linearly with weight one.  Given the
// adjoint lam_v(p) of the 24 outputs, bssn_rhs_adjoint_point recomputes the
// forward quantities of bssn_rhs_point (GAUGE 0, vacuum, no residuals) and
// accumulates, in reverse order, the adjoint of every feature:
//   value bars  (24)      dJ/d(value of field v at p), incl. the upwind
//                          stencil's dependence on beta^a(p),
//   gradient bars (15x3)  dJ/d(d_a f) for the 15 fields with derivatives,
//   Hessian bars  (11x6)  dJ/d(d_a d_b f) for the 11 fields with them.
// A second pass (the gather kernel) applies the transposed stencils to these
// bars and to lam (advection and KO) to give dJ/dU at every point.
//
#ifndef AMSS_ADJ_PIN
#define AMSS_ADJ_PIN 2
#endif
// Notation as in bssn_algebra.h; a trailing _b marks an adjoint.  Every loop
// is unrolled (arrays stay in registers) and sections are fenced so the
// scheduler does not stretch live ranges across them.
#pragma once
#include "bssn_lds.h"

namespace amss {

constexpr int kAdjGradFields = 15;  // betax..z, dxx..dzz, Gamx..z, chi, Lap, trK
constexpr int kAdjHessFields = 11;  // betax..z, dxx..dzz, chi, Lap
constexpr int kAdjFeatures = NFIELD + 3 * kAdjGradFields + 6 * kAdjHessFields;  // 135

__host__ __device__ constexpr int adj_grad_slot(int v) {
  return v >= BETAX && v <= BETAX + 2 ? v - BETAX
         : v >= DXX && v <= DZZ       ? 3 + (v - DXX)
         : v >= GAMX && v <= GAMZ     ? 9 + (v - GAMX)
         : v == CHI                   ? 12
         : v == LAP                   ? 13
         : v == TRK                   ? 14
                                      : -1;
}
__host__ __device__ constexpr int adj_hess_slot(int v) {
  return v >= BETAX && v <= BETAX + 2 ? v - BETAX
         : v >= DXX && v <= DZZ       ? 3 + (v - DXX)
         : v == CHI                   ? 9
         : v == LAP                   ? 10
                                      : -1;
}
__host__ __device__ constexpr int adj_value_index(int v) { return v; }
__host__ __device__ constexpr int adj_grad_index(int slot, int a) { return NFIELD + 3 * slot + a; }
__host__ __device__ constexpr int adj_hess_index(int slot, int t) { return NFIELD + 3 * kAdjGradFields + 6 * slot + t; }

// Where the feature bars of one point go (feature f of point p at base[f n + p]).
struct AdjSink {
  double* base;
  long n, p;
  __device__ __forceinline__ void put(int f, double x) const { base[f * n + p] = x; }
};

// d/d beta of beta_c (max(beta,0) up - min(beta,0) dn), with the one-sided
// rule at beta == 0 (both branches count), as torch's clamp differentiates.
__device__ __forceinline__ double upwind_dbeta(double b, double up, double dn) {
  return (b >= 0.0 ? up : 0.0) - (b <= 0.0 ? dn : 0.0);
}

// sum_v lam_v d/d beta^a (the upwind term of v), without the 1/(12 h) factor
template <class Env>
__device__ __forceinline__ void upwind_beta_sums(const Env& E, const double* lam, double ub[3]) {
  if constexpr (Env::kFeatures) {
    E.ubeta(ub);
  } else {
    double bp[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) bp[a] = E.val(BETAX + a), ub[a] = 0.0;
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) {
      double up[3], dn[3];
      E.updn(v, up, dn);
#pragma unroll
      for (int a = 0; a < 3; ++a) ub[a] += lam[v] * upwind_dbeta(bp[a], up[a], dn[a]);
    }
  }
}

// The stencil values bssn_rhs_adjoint_point reads, precomputed per point
// (feature f of point p at f * n + p): centre values (24), gradients (15 x 3,
// scaled), Hessians (11 x 6, scaled), upwind_beta_sums (3).
constexpr int kAdjInputs = NFIELD + 3 * kAdjGradFields + 6 * kAdjHessFields + 3;  // 138
struct FeatureEnv {
  static constexpr bool kFence = true, kFeatures = true;
  const double* f;
  long n, p;
  const Coef& C;
  __device__ __forceinline__ double val(int v) const { return f[v * n + p]; }
  __device__ __forceinline__ void grad(int v, double d[3]) const {
#pragma unroll
    for (int a = 0; a < 3; ++a) d[a] = f[adj_grad_index(adj_grad_slot(v), a) * n + p];
  }
  __device__ __forceinline__ void hess(int v, double h[6]) const {
#pragma unroll
    for (int t = 0; t < 6; ++t) h[t] = f[adj_hess_index(adj_hess_slot(v), t) * n + p];
  }
  __device__ __forceinline__ void ubeta(double ub[3]) const {
#pragma unroll
    for (int a = 0; a < 3; ++a) ub[a] = f[(kAdjFeatures + a) * n + p];
  }
};

// LDS stash of the reverse algebra: the Christoffel symbols of the first and
// second kind and their adjoints (4 x 18 doubles per point), lane-strided
// (element e of thread t at st[e * kAdjStride + t]: consecutive lanes hit
// consecutive banks).  Holding them in LDS instead of registers halves the
// kernel's spill traffic at its one wave per SIMD.
constexpr int kAdjStash = 72, kAdjStride = 256;  // doubles per point; threads per workgroup
template <class Env>
__device__ void bssn_rhs_adjoint_point(const Env& E, const double* lam, const AdjSink& O, lds_double* st) {
  constexpr double F1o3 = 1.0 / 3.0, F2o3 = 2.0 / 3.0;
  constexpr double W[6] = {1.0, 2.0, 2.0, 1.0, 2.0, 1.0};
  auto S2 = [](int i, int j) { return sidx(i, j); };

  // upwind stencils' dependence on beta^a(p), first: it leaves 3 numbers live
  double ub[3];
  upwind_beta_sums(E, lam, ub);

  AMSS_FENCE();
  // ======================= forward recomputation ===========================
  const double chi1 = E.val(CHI) + 1.0;
  const double ichi = 1.0 / chi1;
  const double K = E.val(TRK);
  const double alp = E.val(LAP) + 1.0;
  double g[6], gu[6], beta[3], A[6], B[3];
#pragma unroll
  for (int s = 0; s < 6; ++s) g[s] = E.val(DXX + s) + (sdiag(s) ? 1.0 : 0.0);
#pragma unroll
  for (int s = 0; s < 6; ++s) A[s] = E.val(AXX + s);
#pragma unroll
  for (int a = 0; a < 3; ++a) beta[a] = E.val(BETAX + a);
#pragma unroll
  for (int a = 0; a < 3; ++a) B[a] = E.val(DTSFX + a);
  gu[0] = g[3] * g[5] - g[4] * g[4];
  gu[1] = g[2] * g[4] - g[1] * g[5];
  gu[2] = g[1] * g[4] - g[2] * g[3];
  gu[3] = g[0] * g[5] - g[2] * g[2];
  gu[4] = g[1] * g[2] - g[0] * g[4];
  gu[5] = g[0] * g[3] - g[1] * g[1];
  {
    const double idet = 1.0 / (g[0] * gu[0] + g[1] * gu[1] + g[2] * gu[2]);
#pragma unroll
    for (int s = 0; s < 6; ++s) gu[s] *= idet;
  }
  auto tr = [&](const double* T) {
    return gu[0] * T[0] + gu[3] * T[3] + gu[5] * T[5] + 2.0 * (gu[1] * T[1] + gu[2] * T[2] + gu[4] * T[4]);
  };
  double AU[3][3], Aup[6];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) AU[i][j] = gu[S2(i, 0)] * A[S2(0, j)] + gu[S2(i, 1)] * A[S2(1, j)] + gu[S2(i, 2)] * A[S2(2, j)];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) Aup[S2(i, j)] = AU[i][0] * gu[S2(0, j)] + AU[i][1] * gu[S2(1, j)] + AU[i][2] * gu[S2(2, j)];

  AMSS_FENCE();
  double db[3][3], gdiv[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double h[6];
    E.grad(BETAX + a, db[a]);
    E.hess(BETAX + a, h);
#pragma unroll
    for (int j = 0; j < 3; ++j) gdiv[j] += h[S2(j, a)];
  }
  const double divb = db[0][0] + db[1][1] + db[2][2];

  AMSS_FENCE();
  auto G1 = [&](int k, int s) -> lds_double& { return st[(0 * 18 + k * 6 + s) * kAdjStride]; };
  double R[6];
  {
    double dg[3][6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      double d[3], h[6];
      E.grad(DXX + s, d);
#pragma unroll
      for (int k = 0; k < 3; ++k) dg[k][s] = d[k];
      E.hess(DXX + s, h);
      R[s] = -0.5 * tr(h);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) G1(k, S2(i, j)) = 0.5 * (dg[i][S2(j, k)] + dg[j][S2(i, k)] - dg[k][S2(i, j)]);
  }
  AMSS_FENCE();
  auto G2 = [&](int k, int s) -> lds_double& { return st[(1 * 18 + k * 6 + s) * kAdjStride]; };
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int s = 0; s < 6; ++s) G2(m, s) = gu[S2(m, 0)] * G1(0, s) + gu[S2(m, 1)] * G1(1, s) + gu[S2(m, 2)] * G1(2, s);
  AMSS_FENCE();
  double Gamxa[3], GA[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    Gamxa[m] = gu[0] * G2(m, 0) + gu[3] * G2(m, 3) + gu[5] * G2(m, 5) + 2.0 * (gu[1] * G2(m, 1) + gu[2] * G2(m, 2) + gu[4] * G2(m, 4));
    double t = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) t += W[s] * G2(m, s) * Aup[s];
    GA[m] = t;
  }
  AMSS_FENCE();
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double U[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        U[m][k] = gu[S2(0, m)] * G2(k, S2(0, i)) + gu[S2(1, m)] * G2(k, S2(1, i)) + gu[S2(2, m)] * G2(k, S2(2, i));
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double P = 0.0, H = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          P += U[m][k] * (G1(j, S2(k, m)) + G1(k, S2(m, j)));
          H += U[m][k] * G1(j, S2(k, m));
        }
      const double gg = 0.5 * (Gamxa[0] * G1(i, S2(j, 0)) + Gamxa[1] * G1(i, S2(j, 1)) + Gamxa[2] * G1(i, S2(j, 2)));
      if (j >= i) R[S2(i, j)] += P + gg;
      if (j <= i) R[S2(j, i)] += H + gg;
    }
  }
  AMSS_FENCE();
  double dGt[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    E.grad(GAMX + k, dGt[k]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) R[S2(i, j)] += 0.5 * (g[S2(k, i)] * dGt[k][j] + g[S2(k, j)] * dGt[k][i]);
  }
  AMSS_FENCE();
  double dchi[3], hh[6], fs = 0.0;
  {
    double h[6];
    E.grad(CHI, dchi);
    E.hess(CHI, h);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const int s = S2(i, j);
        hh[s] = h[s] - (G2(0, s) * dchi[0] + G2(1, s) * dchi[1] + G2(2, s) * dchi[2]);
        fs += W[s] * gu[s] * (hh[s] - 1.5 * ichi * dchi[i] * dchi[j]);
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const int s = S2(i, j);
        R[s] += 0.5 * ichi * (hh[s] - 0.5 * ichi * dchi[i] * dchi[j] + g[s] * fs);
      }
  }
  AMSS_FENCE();
  double u[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k] = ichi * (gu[S2(k, 0)] * dchi[0] + gu[S2(k, 1)] * dchi[1] + gu[S2(k, 2)] * dchi[2]);
  // physical Christoffels Gamma^k_s, s = (i, j), evaluated where used (not kept)
  auto P2 = [&](int k, int s) {
    const int i = s < 3 ? 0 : (s < 5 ? 1 : 2);
    const int j = s < 3 ? s : (s < 5 ? s - 2 : 2);
    double d = 0.5 * g[s] * u[k];
    if (k == i) d -= 0.5 * ichi * dchi[j];
    if (k == j) d -= 0.5 * ichi * dchi[i];
    return G2(k, s) + d;
  };
  AMSS_FENCE();
  double dalp[3], DDa[6];
  {
    double h[6];
    E.grad(LAP, dalp);
    E.hess(LAP, h);
#pragma unroll
    for (int s = 0; s < 6; ++s) DDa[s] = h[s] - (P2(0, s) * dalp[0] + P2(1, s) * dalp[1] + P2(2, s) * dalp[2]);
  }
  AMSS_FENCE();
  double dK[3];
  E.grad(TRK, dK);
  double Adchi[3], Ada[3], src[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Adchi[i] = Ada[i] = src[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Adchi[i] += Aup[S2(i, j)] * dchi[j];
      Ada[i] += Aup[S2(i, j)] * dalp[j];
      src[i] += gu[S2(i, j)] * (F2o3 * dK[j]);
    }
  }
  double trAA = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) trAA += AU[i][j] * AU[j][i];
  const double trDDa = tr(DDa);
  const double f0 = F2o3 * K * K - trAA;
  const double F = -F1o3 * (trDDa + alp * ichi * f0);

  // ======================= reverse sweep ===================================
  double chi1_b = 0.0, ichi_b = 0.0, K_b = 0.0, alp_b = 0.0;
  double g_b[6] = {}, A_b[6] = {}, B_b[3] = {}, gu_b[6] = {};
  double beta_b[3] = {E.C.d1[0] * ub[0], E.C.d1[1] * ub[1], E.C.d1[2] * ub[2]};
  double db_b[3][3] = {}, AU_b[3][3] = {}, Aup_b[6] = {};
  double R_b[6] = {}, DDa_b[6] = {}, F_b = 0.0, trDDa_b = 0.0, trAA_b = 0.0, divb_b = 0.0;
  double dchi_b[3] = {}, dalp_b[3] = {}, dK_b[3] = {}, Gamxa_b[3] = {}, GA_b[3] = {}, Lb_b[3] = {};
  auto G1_b = [&](int k, int s) -> lds_double& { return st[(2 * 18 + k * 6 + s) * kAdjStride]; };
  auto G2_b = [&](int k, int s) -> lds_double& { return st[(3 * 18 + k * 6 + s) * kAdjStride]; };
#pragma unroll
  for (int e = 0; e < 18; ++e) G1_b(e / 6, e % 6) = 0.0, G2_b(e / 6, e % 6) = 0.0;
  // the running sums of the most-fed bars are formed at every section end
  // (not deferred to the outputs, which would keep their terms live)
  auto pin_bars = [&]() {
    if constexpr (AMSS_ADJ_PIN) {
      amss_pin<Env>(g_b);
      amss_pin<Env>(gu_b);
      amss_pin<Env>(dchi_b);
      amss_pin1<Env>(ichi_b);
      amss_pin1<Env>(chi1_b);
      amss_pin1<Env>(alp_b);
      amss_pin1<Env>(K_b);
      if (AMSS_ADJ_PIN > 1) {
        amss_pin<Env>(A_b);
        amss_pin<Env>(Aup_b);
        amss_pin<Env>(dalp_b);
        amss_pin<Env>(R_b);
        amss_pin<Env>(DDa_b);
      }
    }
  };

  pin_bars();
  AMSS_FENCE();
  // A_ij rhs
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int i = s < 3 ? 0 : (s < 5 ? 1 : 2);
    const int j = s < 3 ? s : (s < 5 ? s - 2 : 2);
    const double l = lam[AXX + s];
    double AA = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) AA += A[S2(i, k)] * AU[k][j];
    chi1_b += l * (alp * R[s] - DDa[s] - g[s] * F);
    alp_b += l * (chi1 * R[s] + K * A[s] - 2.0 * AA);
    R_b[s] += l * chi1 * alp;
    DDa_b[s] -= l * chi1;
    g_b[s] -= l * chi1 * F;
    F_b -= l * chi1 * g[s];
    K_b += l * alp * A[s];
    A_b[s] += l * (alp * K - F2o3 * divb);
    divb_b -= F2o3 * A[s] * l;
    const double AA_b = -2.0 * alp * l;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      A_b[S2(i, k)] += AA_b * AU[k][j];
      AU_b[k][j] += AA_b * A[S2(i, k)];
      // Lie term A_ik d_j beta^k + A_jk d_i beta^k
      A_b[S2(i, k)] += l * db[k][j];
      db_b[k][j] += l * A[S2(i, k)];
      A_b[S2(j, k)] += l * db[k][i];
      db_b[k][i] += l * A[S2(j, k)];
    }
  }
  pin_bars();
  AMSS_FENCE();
  // F, f0
  trDDa_b -= F1o3 * F_b;
  alp_b -= F1o3 * F_b * ichi * f0;
  ichi_b -= F1o3 * F_b * alp * f0;
  {
    const double f0_b = -F1o3 * F_b * alp * ichi;
    K_b += 2.0 * F2o3 * K * f0_b;
    trAA_b -= f0_b;
  }
  pin_bars();
  AMSS_FENCE();
  // trK rhs
  {
    const double l = lam[TRK];
    chi1_b -= l * trDDa;
    trDDa_b -= l * chi1;
    alp_b += l * (F1o3 * K * K + trAA);
    K_b += l * alp * 2.0 * F1o3 * K;
    trAA_b += l * alp;
  }
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    gu_b[s] += trDDa_b * W[s] * DDa[s];
    DDa_b[s] += trDDa_b * W[s] * gu[s];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) AU_b[i][j] += 2.0 * trAA_b * AU[j][i];

  pin_bars();
  AMSS_FENCE();
  // Gamma-tilde and B rhs (GAUGE 0: eta = 2)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double gb = lam[GAMX + i] + lam[DTSFX + i];
    B_b[i] -= 2.0 * lam[DTSFX + i];
    const double Ada_b = -2.0 * gb;
    alp_b += 2.0 * gb * (GA[i] - 1.5 * ichi * Adchi[i] - src[i]);
    GA_b[i] += 2.0 * alp * gb;
    ichi_b -= 3.0 * alp * gb * Adchi[i];
    const double Adchi_b = -3.0 * alp * ichi * gb;
    const double src_b = -2.0 * alp * gb;
    Lb_b[i] += gb;
    Gamxa_b[i] += F2o3 * divb * gb;
    divb_b += F2o3 * Gamxa[i] * gb;
    const double ab_b = -gb;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Aup_b[S2(i, j)] += Adchi_b * dchi[j] + Ada_b * dalp[j];
      dchi_b[j] += Adchi_b * Aup[S2(i, j)];
      dalp_b[j] += Ada_b * Aup[S2(i, j)];
      gu_b[S2(i, j)] += src_b * F2o3 * dK[j];
      dK_b[j] += src_b * F2o3 * gu[S2(i, j)];
      Gamxa_b[j] += ab_b * db[i][j];
      db_b[i][j] += ab_b * Gamxa[j];
    }
  }
  pin_bars();
  AMSS_FENCE();
  // lapse rhs and D_i D_j alpha
  alp_b -= 2.0 * K * lam[LAP];
  K_b -= 2.0 * alp * lam[LAP];
  double ha_b[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) ha_b[s] = DDa_b[s];
  pin_bars();
  AMSS_FENCE();
  // physical Christoffels: their adjoint is -DDa_b[s] dalp[k]
  {
    double u_b[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
          const int s = S2(i, j);
          dalp_b[k] -= DDa_b[s] * P2(k, s);
          const double pb = -DDa_b[s] * dalp[k];
          G2_b(k, s) += pb;
          g_b[s] += 0.5 * u[k] * pb;
          u_b[k] += 0.5 * g[s] * pb;
          if (k == i) {
            ichi_b -= 0.5 * dchi[j] * pb;
            dchi_b[j] -= 0.5 * ichi * pb;
          }
          if (k == j) {
            ichi_b -= 0.5 * dchi[i] * pb;
            dchi_b[i] -= 0.5 * ichi * pb;
          }
        }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        ichi_b += u_b[k] * gu[S2(k, l)] * dchi[l];
        gu_b[S2(k, l)] += u_b[k] * ichi * dchi[l];
        dchi_b[l] += u_b[k] * ichi * gu[S2(k, l)];
      }
  }
  pin_bars();
  AMSS_FENCE();
  // chi rhs
  {
    const double l = lam[CHI];
    chi1_b += F2o3 * (alp * K - divb) * l;
    alp_b += F2o3 * chi1 * K * l;
    K_b += F2o3 * chi1 * alp * l;
    divb_b -= F2o3 * chi1 * l;
  }
  pin_bars();
  AMSS_FENCE();
  // chi part of Ricci
  double hc_b[6];
  {
    double hh_b[6], fs_b = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const int s = S2(i, j);
        const double T = hh[s] - 0.5 * ichi * dchi[i] * dchi[j] + g[s] * fs;
        ichi_b += 0.5 * T * R_b[s];
        const double T_b = 0.5 * ichi * R_b[s];
        hh_b[s] = T_b;
        ichi_b -= 0.5 * dchi[i] * dchi[j] * T_b;
        dchi_b[i] -= 0.5 * ichi * dchi[j] * T_b;
        dchi_b[j] -= 0.5 * ichi * dchi[i] * T_b;
        g_b[s] += fs * T_b;
        fs_b += g[s] * T_b;
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const int s = S2(i, j);
        const double w = fs_b * W[s];
        gu_b[s] += w * (hh[s] - 1.5 * ichi * dchi[i] * dchi[j]);
        hh_b[s] += w * gu[s];
        ichi_b -= w * gu[s] * 1.5 * dchi[i] * dchi[j];
        dchi_b[i] -= w * gu[s] * 1.5 * ichi * dchi[j];
        dchi_b[j] -= w * gu[s] * 1.5 * ichi * dchi[i];
      }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      hc_b[s] = hh_b[s];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        G2_b(m, s) -= hh_b[s] * dchi[m];
        dchi_b[m] -= hh_b[s] * G2(m, s);
      }
    }
  }
  pin_bars();
  AMSS_FENCE();
  // Gamma-tilde derivatives in Ricci
  double dGt_b[3][3] = {};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const double rb = 0.5 * R_b[S2(i, j)];
        g_b[S2(k, i)] += rb * dGt[k][j];
        dGt_b[k][j] += rb * g[S2(k, i)];
        g_b[S2(k, j)] += rb * dGt[k][i];
        dGt_b[k][i] += rb * g[S2(k, j)];
      }
  pin_bars();
  AMSS_FENCE();
  // quadratic Ricci
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double U[3][3], U_b[3][3] = {};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        U[m][k] = gu[S2(0, m)] * G2(k, S2(0, i)) + gu[S2(1, m)] * G2(k, S2(1, i)) + gu[S2(2, m)] * G2(k, S2(2, i));
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double P_b = j >= i ? R_b[S2(i, j)] : 0.0;
      const double H_b = j <= i ? R_b[S2(j, i)] : 0.0;
      const double gg_b = P_b + H_b;
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        Gamxa_b[l] += 0.5 * gg_b * G1(i, S2(j, l));
        G1_b(i, S2(j, l)) += 0.5 * gg_b * Gamxa[l];
      }
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          U_b[m][k] += P_b * (G1(j, S2(k, m)) + G1(k, S2(m, j))) + H_b * G1(j, S2(k, m));
          G1_b(j, S2(k, m)) += (P_b + H_b) * U[m][k];
          G1_b(k, S2(m, j)) += P_b * U[m][k];
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          gu_b[S2(l, m)] += U_b[m][k] * G2(k, S2(l, i));
          G2_b(k, S2(l, i)) += U_b[m][k] * gu[S2(l, m)];
        }
  }
  pin_bars();
  AMSS_FENCE();
  // Gamxa, GA
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      G2_b(m, s) += GA_b[m] * W[s] * Aup[s] + Gamxa_b[m] * W[s] * gu[s];
      Aup_b[s] += GA_b[m] * W[s] * G2(m, s);
      gu_b[s] += Gamxa_b[m] * W[s] * G2(m, s);
    }
  pin_bars();
  AMSS_FENCE();
  // G2 = gu G1
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        gu_b[S2(m, l)] += G2_b(m, s) * G1(l, s);
        G1_b(l, s) += G2_b(m, s) * gu[S2(m, l)];
      }
  pin_bars();
  AMSS_FENCE();
  // G1 from d g: metric gradient bars
  double dg_b[3][6] = {};
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const double b = 0.5 * G1_b(k, S2(i, j));
        dg_b[i][S2(j, k)] += b;
        dg_b[j][S2(i, k)] += b;
        dg_b[k][S2(i, j)] -= b;
      }
  pin_bars();
  AMSS_FENCE();
  // metric rhs
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int i = s < 3 ? 0 : (s < 5 ? 1 : 2);
    const int j = s < 3 ? s : (s < 5 ? s - 2 : 2);
    const double l = lam[DXX + s];
    alp_b -= 2.0 * A[s] * l;
    A_b[s] -= 2.0 * alp * l;
    g_b[s] -= F2o3 * divb * l;
    divb_b -= F2o3 * g[s] * l;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      g_b[S2(i, k)] += l * db[k][j];
      db_b[k][j] += l * g[S2(i, k)];
      g_b[S2(j, k)] += l * db[k][i];
      db_b[k][i] += l * g[S2(j, k)];
    }
  }
  pin_bars();
  AMSS_FENCE();
  // metric Hessians through R = -1/2 gu^kl d_k d_l g_ij: bars and gu_b
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    double h[6];
    E.hess(DXX + s, h);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      gu_b[t] -= 0.5 * R_b[s] * W[t] * h[t];
      O.put(adj_hess_index(adj_hess_slot(DXX + s), t), -0.5 * R_b[s] * W[t] * gu[t]);
    }
  }
  pin_bars();
  AMSS_FENCE();
  // shift: Lb = gu^jk d_j d_k beta^i + 1/3 gu^ij d_j div beta
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double h[6];
    E.hess(BETAX + a, h);
    double hb_b[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
      gu_b[t] += Lb_b[a] * W[t] * h[t];
      hb_b[t] += Lb_b[a] * W[t] * gu[t];
    }
    // gdiv[j] = sum_a h_a[(j,a)]; Lb_i += 1/3 gu_ij gdiv_j
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double gdiv_b = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) gdiv_b += F1o3 * Lb_b[i] * gu[S2(i, j)];
      hb_b[S2(j, a)] += gdiv_b;
    }
#pragma unroll
    for (int t = 0; t < 6; ++t) O.put(adj_hess_index(adj_hess_slot(BETAX + a), t), hb_b[t]);
    B_b[a] += 0.75 * lam[BETAX + a];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gu_b[S2(i, j)] += F1o3 * Lb_b[i] * gdiv[j];
#pragma unroll
  for (int a = 0; a < 3; ++a) db_b[a][a] += divb_b;
  pin_bars();
  AMSS_FENCE();
  // A^ij and A^i_j
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        AU_b[i][k] += Aup_b[S2(i, j)] * gu[S2(k, j)];
        gu_b[S2(k, j)] += Aup_b[S2(i, j)] * AU[i][k];
      }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gu_b[S2(i, k)] += AU_b[i][j] * A[S2(k, j)];
        A_b[S2(k, j)] += AU_b[i][j] * gu[S2(i, k)];
      }
  pin_bars();
  AMSS_FENCE();
  // inverse metric: d gu = -gu (d g) gu
  {
    double Ub[3][3], T[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Ub[i][j] = i == j ? gu_b[S2(i, j)] : 0.5 * gu_b[S2(i, j)];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) T[i][j] = Ub[i][0] * gu[S2(0, j)] + Ub[i][1] * gu[S2(1, j)] + Ub[i][2] * gu[S2(2, j)];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const double M = gu[S2(i, 0)] * T[0][j] + gu[S2(i, 1)] * T[1][j] + gu[S2(i, 2)] * T[2][j];
        g_b[S2(i, j)] -= i == j ? M : 2.0 * M;
      }
  }
  chi1_b -= ichi * ichi * ichi_b;

  // ======================= feature bars ====================================
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) {
    double x = 0.0;
    if (v == CHI) x = chi1_b;
    else if (v == TRK) x = K_b;
    else if (v == LAP) x = alp_b;
    else if (v >= DXX && v <= DZZ) x = g_b[v - DXX];
    else if (v >= AXX && v <= AZZ) x = A_b[v - AXX];
    else if (v >= BETAX && v <= BETAX + 2) x = beta_b[v - BETAX];
    else if (v >= DTSFX) x = B_b[v - DTSFX];
    O.put(adj_value_index(v), x);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) O.put(adj_grad_index(adj_grad_slot(BETAX + a), j), db_b[a][j]);
#pragma unroll
  for (int s = 0; s < 6; ++s)
#pragma unroll
    for (int k = 0; k < 3; ++k) O.put(adj_grad_index(adj_grad_slot(DXX + s), k), dg_b[k][s]);
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < 3; ++j) O.put(adj_grad_index(adj_grad_slot(GAMX + k), j), dGt_b[k][j]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    O.put(adj_grad_index(adj_grad_slot(CHI), a), dchi_b[a]);
    O.put(adj_grad_index(adj_grad_slot(LAP), a), dalp_b[a]);
    O.put(adj_grad_index(adj_grad_slot(TRK), a), dK_b[a]);
  }
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    O.put(adj_hess_index(adj_hess_slot(CHI), t), hc_b[t]);
    O.put(adj_hess_index(adj_hess_slot(LAP), t), ha_b[t]);
  }
}

}  // namespace amss
