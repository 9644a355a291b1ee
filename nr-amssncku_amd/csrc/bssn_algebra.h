// bssn_algebra.h -- the BSSN right-hand side at one grid point.
//
// Restates compute_rhs_bssn (reference bssn_rhs.f90:140-1134) for
// GAUGE 0 (B^i driver with eta = 2, :792-799) and GAUGE 2 (:809-820, the
// gauge of the reference's checked-in build configuration) in tensor-index
// form.  Notation: g = conformal metric (diagonal stored as offsets), gu its
// inverse, G1[k][ij] = Gamma_{k,ij} (first kind), G2[k][ij] = Gamma^k_{ij}
// (second kind); chi and alpha are stored as offsets.
//
// Register economy drives the order.  The stencils of one input field
// (gradient, Hessian, upwind advection, KO dissipation) are evaluated
// together in one "field group" and immediately reduced to the handful of
// numbers the algebra needs (e.g. a metric component's Hessian only enters
// through gu^kl d_k d_l g_ij), so only reduced quantities cross groups.
// AMSS_FENCE() between groups (where Env::kFence) stops the compiler from hoisting the loads of
// later groups (and from keeping loaded neighbours alive across groups).
//
// Env provides the stencils of the stored input fields at this point:
//   val(v), grad(v, d[3]), hess(v, d[6]), adv(v, offset, beta[3]), ko(v),
//   mat(m), and begin(v)/end() bracketing a field group (LDS staging hook).
// Sink receives put(v, k) once per evolved field and, when requested, the
// physical Christoffels (gam), physical Ricci (ric) and residuals (res).
#pragma once
#include <type_traits>
#include "bssn_device.h"

namespace amss {

constexpr double kPi = 3.141592653589793;  // acos(-1), bssn_rhs.f90:140

#ifndef AMSS_PIN
#define AMSS_PIN 2
#endif
// Group fence: the environment decides (Env::kFence) whether the compiler
// may move work across field groups.
template <class Env>
__device__ __forceinline__ void amss_fence() {
  if constexpr (Env::kFence) {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
  }
}
#define AMSS_FENCE() amss_fence<Env>()

// Pin values at a group boundary: sums finished here are not deferred to
// their use (which would keep every summand live until then).
template <class Env, int N>
__device__ __forceinline__ void amss_pin(double (&x)[N]) {
  if constexpr (Env::kFence && AMSS_PIN) {
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" : "+v"(x[i]));
  }
}
template <class Env>
__device__ __forceinline__ void amss_pin1(double& x) {
  if constexpr (Env::kFence && AMSS_PIN) asm volatile("" : "+v"(x));
}
// How many of the pins a sink wants (Sink::kPin if it says; measured per RK4
// sub-step: the first gains from fewer)
template <class S, class = void>
struct PinLevel {
  static constexpr int v = AMSS_PIN;
};
template <class S>
struct PinLevel<S, std::void_t<decltype(S::kPin)>> {
  static constexpr int v = S::kPin < AMSS_PIN ? S::kPin : AMSS_PIN;
};

template <int GAUGE, bool CO, bool GAMRIC, bool MATTER, class Env, class Sink>
__device__ __forceinline__ void bssn_rhs_point(Env& E, Sink& S) {
  constexpr double F1o3 = 1.0 / 3.0, F2o3 = 2.0 / 3.0;
  constexpr double W[6] = {1.0, 2.0, 2.0, 1.0, 2.0, 1.0};  // symmetric-sum weights

  // ---- centre values, inverse metric (:146-150, :194-201) -------------------
  const double chi1 = E.val(CHI) + 1.0;
  const double ichi = 1.0 / chi1;
  const double K = E.val(TRK);
  const double alp = E.val(LAP) + 1.0;
  double g[6], gu[6], beta[3];
#pragma unroll
  for (int s = 0; s < 6; ++s) g[s] = E.val(DXX + s) + (sdiag(s) ? 1.0 : 0.0);
#pragma unroll
  for (int a = 0; a < 3; ++a) beta[a] = E.val(BETAX + a);
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
  auto tr = [&](const double* T) {  // gu^ij T_ij
    return gu[0] * T[0] + gu[3] * T[3] + gu[5] * T[5] + 2.0 * (gu[1] * T[1] + gu[2] * T[2] + gu[4] * T[4]);
  };
  double Sv[3] = {0.0, 0.0, 0.0};
  if (MATTER) {
#pragma unroll
    for (int a = 0; a < 3; ++a) Sv[a] = E.mat(SX + a);
  }
  // A^ij (raised with gu), needed by the Gamma-tilde equation (:260-282)
  double Aup[6];
  {
    double A[6], AU[3][3];
#pragma unroll
    for (int s = 0; s < 6; ++s) A[s] = E.val(AXX + s);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        AU[i][j] = gu[sidx(i, 0)] * A[sidx(0, j)] + gu[sidx(i, 1)] * A[sidx(1, j)] + gu[sidx(i, 2)] * A[sidx(2, j)];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j)
        Aup[sidx(i, j)] = AU[i][0] * gu[sidx(0, j)] + AU[i][1] * gu[sidx(1, j)] + AU[i][2] * gu[sidx(2, j)];
  }
  AMSS_FENCE();

  // ---- shift group: d beta, gu^jk d_j d_k beta^i + gu^ij d_j div beta / 3 -----
  double db[3][3];  // db[i][j] = d_j beta^i
  double Lb[3];
  {
    double lapb[3], gdiv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      E.begin(BETAX + a);
      double h[6];
      E.grad(BETAX + a, db[a]);
      E.hess(BETAX + a, h);
      lapb[a] = tr(h);
#pragma unroll
      for (int j = 0; j < 3; ++j) gdiv[j] += h[sidx(j, a)];
      // beta^i rhs = 3/4 B^i (:793-795) + advection + KO
      S.put(BETAX + a, 0.75 * E.val(DTSFX + a) + E.adv(BETAX + a, 0.0, beta) + E.ko(BETAX + a));
      E.end();
      amss_pin<Env>(gdiv);
      amss_pin<Env>(db[a]);
      amss_pin1<Env>(lapb[a]);
      AMSS_FENCE();
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
      Lb[i] = lapb[i] + F1o3 * (gu[sidx(i, 0)] * gdiv[0] + gu[sidx(i, 1)] * gdiv[1] + gu[sidx(i, 2)] * gdiv[2]);
  }
  const double divb = db[0][0] + db[1][1] + db[2][2];

  // ---- metric group: d g, gu^kl d_k d_l g_ij, g_ij rhs (:162-191, :378-400) ----
  double G1[3][6];  // Gamma_{k,ij}
  double R[6];      // conformal Ricci, assembled below
  {
    double dg[3][6];  // dg[k][ij] = d_k g_ij
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      E.begin(DXX + s);
      double d[3], h[6];
      E.grad(DXX + s, d);
      dg[0][s] = d[0];
      dg[1][s] = d[1];
      dg[2][s] = d[2];
      E.hess(DXX + s, h);
      R[s] = -0.5 * tr(h);
      const int i = s < 3 ? 0 : (s < 5 ? 1 : 2);
      const int j = s < 3 ? s : (s < 5 ? s - 2 : 2);
      double lie = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) lie += g[sidx(i, k)] * db[k][j] + g[sidx(j, k)] * db[k][i];
      S.put(DXX + s, -2.0 * alp * E.val(AXX + s) + lie - F2o3 * g[s] * divb +
                         E.adv(DXX + s, sdiag(s) ? 1.0 : 0.0, beta) + E.ko(DXX + s));
      E.end();
      amss_pin1<Env>(dg[0][s]);
      amss_pin1<Env>(dg[1][s]);
      amss_pin1<Env>(dg[2][s]);
      amss_pin1<Env>(R[s]);
      AMSS_FENCE();
    }
    if (CO) {  // Gam^i_Res = Gam^i - gu^ik gu^jl d_j g_kl (:203-232)
      double t[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int l = 0; l < 3; ++l) acc += gu[sidx(j, l)] * dg[j][sidx(k, l)];
        t[k] = acc;
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
        S.res(4 + i, E.val(GAMX + i) - (gu[sidx(i, 0)] * t[0] + gu[sidx(i, 1)] * t[1] + gu[sidx(i, 2)] * t[2]));
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j)
          G1[k][sidx(i, j)] = 0.5 * (dg[i][sidx(j, k)] + dg[j][sidx(i, k)] - dg[k][sidx(i, j)]);
  }
  double G2[3][6];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int s = 0; s < 6; ++s)
      G2[m][s] = gu[sidx(m, 0)] * G1[0][s] + gu[sidx(m, 1)] * G1[1][s] + gu[sidx(m, 2)] * G1[2][s];
  double Gamxa[3], GA[3];  // gu^jk Gamma^i_jk ; Gamma^i_jk A^jk
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    Gamxa[m] = tr(G2[m]);
    double t = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) t += W[s] * G2[m][s] * Aup[s];
    GA[m] = t;
  }
  // Ricci (:402-601): Gamxa^k Gamma_(ij)k + gu^lm (2 G^k_l(i G_j)km + G^k_im G_klj)
  //   quadratic part = sum_mk U^mk_i (G1_j,km + G1_k,mj) + U^mk_j G1_i,km,
  //   U^mk_i = gu^lm Gamma^k_li
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double U[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        U[m][k] = gu[sidx(0, m)] * G2[k][sidx(0, i)] + gu[sidx(1, m)] * G2[k][sidx(1, i)] +
                  gu[sidx(2, m)] * G2[k][sidx(2, i)];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double P = 0.0, H = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          P += U[m][k] * (G1[j][sidx(k, m)] + G1[k][sidx(m, j)]);
          H += U[m][k] * G1[j][sidx(k, m)];
        }
      if (j >= i) R[sidx(i, j)] += P + 0.5 * (Gamxa[0] * G1[i][sidx(j, 0)] + Gamxa[1] * G1[i][sidx(j, 1)] +
                                              Gamxa[2] * G1[i][sidx(j, 2)]);
      if (j <= i) R[sidx(j, i)] += H + 0.5 * (Gamxa[0] * G1[i][sidx(j, 0)] + Gamxa[1] * G1[i][sidx(j, 1)] +
                                              Gamxa[2] * G1[i][sidx(j, 2)]);
    }
  }
  AMSS_FENCE();

  // ---- Gamma-tilde group: g_k(i d_j) Gamma-tilde^k into R; adv + KO kept ----
  double pGt[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    E.begin(GAMX + k);
    double d[3];
    E.grad(GAMX + k, d);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) R[sidx(i, j)] += 0.5 * (g[sidx(k, i)] * d[j] + g[sidx(k, j)] * d[i]);
    pGt[k] = E.adv(GAMX + k, 0.0, beta) + E.ko(GAMX + k);
    E.end();
    amss_pin<Env>(R);
    amss_pin1<Env>(pGt[k]);
    AMSS_FENCE();
  }

  // ---- chi group: chi rhs (:160), chi part of Ricci (:603-626) ----------------
  double dchi[3];
  {
    E.begin(CHI);
    double h[6];
    E.grad(CHI, dchi);
    E.hess(CHI, h);
    S.put(CHI, F2o3 * chi1 * (alp * K - divb) + E.adv(CHI, 0.0, beta) + E.ko(CHI));
    E.end();
    double fs = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const int s = sidx(i, j);
        h[s] -= G2[0][s] * dchi[0] + G2[1][s] * dchi[1] + G2[2][s] * dchi[2];
        fs += W[s] * gu[s] * (h[s] - 1.5 * ichi * dchi[i] * dchi[j]);
      }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = i; j < 3; ++j) {
        const int s = sidx(i, j);
        R[s] += 0.5 * ichi * (h[s] - 0.5 * ichi * dchi[i] * dchi[j] + g[s] * fs);
      }
  }
  // physical Christoffels (:632-650): G2 -> Gamma^k_ij of chi^-1 g
  {
    double u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      u[k] = ichi * (gu[sidx(k, 0)] * dchi[0] + gu[sidx(k, 1)] * dchi[1] + gu[sidx(k, 2)] * dchi[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
          const int s = sidx(i, j);
          double d = 0.5 * g[s] * u[k];
          if (k == i) d -= 0.5 * ichi * dchi[j];
          if (k == j) d -= 0.5 * ichi * dchi[i];
          G2[k][s] += d;
        }
  }
  if (PinLevel<Sink>::v > 1) {
    amss_pin<Env>(R);
    amss_pin<Env>(dchi);
  }
  AMSS_FENCE();

  // ---- lapse group: D_i D_j alpha (:629-664), alpha rhs (:791) ---------------
  double dalp[3], DDa[6];
  {
    E.begin(LAP);
    double h[6];
    E.grad(LAP, dalp);
    E.hess(LAP, h);
    S.put(LAP, -2.0 * alp * K + E.adv(LAP, 0.0, beta) + E.ko(LAP));
    E.end();
#pragma unroll
    for (int s = 0; s < 6; ++s) DDa[s] = h[s] - (G2[0][s] * dalp[0] + G2[1][s] * dalp[1] + G2[2][s] * dalp[2]);
  }
  if (PinLevel<Sink>::v > 1) {
    amss_pin<Env>(DDa);
    amss_pin<Env>(dalp);
  }
  if (GAMRIC) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int s = 0; s < 6; ++s) S.gam(k * 6 + s, G2[k][s]);
#pragma unroll
    for (int s = 0; s < 6; ++s) S.ric(s, R[s]);
  }
  AMSS_FENCE();

  // ---- trK group ---------------------------------------------------------------
  double dK[3], pK;
  {
    E.begin(TRK);
    E.grad(TRK, dK);
    pK = E.adv(TRK, 0.0, beta) + E.ko(TRK);
    E.end();
  }
  if (PinLevel<Sink>::v > 1) {
    amss_pin<Env>(dK);
    amss_pin1<Env>(pK);
  }
  AMSS_FENCE();

  // ---- Gamma-tilde rhs (:287-353) and the B^i driver (:792-820) -------------------
  {
    double gtr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double Adchi = 0.0, Ada = 0.0, src = 0.0, ab = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Adchi += Aup[sidx(i, j)] * dchi[j];
        Ada += Aup[sidx(i, j)] * dalp[j];
        src += gu[sidx(i, j)] * (F2o3 * dK[j] + 8.0 * kPi * Sv[j]);
        ab += Gamxa[j] * db[i][j];
      }
      gtr[i] = -2.0 * Ada + 2.0 * alp * (GA[i] - 1.5 * ichi * Adchi - src) + Lb[i] + F2o3 * Gamxa[i] * divb - ab;
      S.put(GAMX + i, gtr[i] + pGt[i]);
    }
    double eta_b = 2.0;
    if (GAUGE == 2) {
      double q = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) q += gu[sidx(i, j)] * dchi[i] * dchi[j];
      const double w = 1.0 - sqrt(chi1);
      eta_b = 0.655 * sqrt(q * ichi) / (w * w);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      E.begin(DTSFX + a);
      S.put(DTSFX + a, gtr[a] - eta_b * E.val(DTSFX + a) + E.adv(DTSFX + a, 0.0, beta) + E.ko(DTSFX + a));
      E.end();
      AMSS_FENCE();
    }
  }

  // ---- trK and A_ij rhs (:666-787) -------------------------------------------
  {
    double A[6], AU[3][3];
#pragma unroll
    for (int s = 0; s < 6; ++s) A[s] = E.val(AXX + s);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        AU[i][j] = gu[sidx(i, 0)] * A[sidx(0, j)] + gu[sidx(i, 1)] * A[sidx(1, j)] + gu[sidx(i, 2)] * A[sidx(2, j)];
    double trAA = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) trAA += AU[i][j] * AU[j][i];
    double rho = 0.0, Sij[6] = {0, 0, 0, 0, 0, 0};
    if (MATTER) {
      rho = E.mat(RHO);
#pragma unroll
      for (int s = 0; s < 6; ++s) Sij[s] = E.mat(SXX + s);
    }
    const double Str = chi1 * tr(Sij);
    const double trDDa = tr(DDa);
    S.put(TRK, -chi1 * trDDa + alp * (F1o3 * K * K + trAA + 4.0 * kPi * (rho + Str)) + pK);
    const double f0 = F2o3 * K * K - trAA - 16.0 * kPi * rho + 8.0 * kPi * Str;
    const double F = -F1o3 * (trDDa + alp * ichi * f0);
    double M[3] = {0.0, 0.0, 0.0};  // momentum residual accumulators (co == 0)
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int i = s < 3 ? 0 : (s < 5 ? 1 : 2);
      const int j = s < 3 ? s : (s < 5 ? s - 2 : 2);
      E.begin(AXX + s);
      double lie = 0.0, AA = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        lie += A[sidx(i, k)] * db[k][j] + A[sidx(j, k)] * db[k][i];
        AA += A[sidx(i, k)] * AU[k][j];
      }
      const double k = chi1 * (alp * (R[s] - 8.0 * kPi * Sij[s]) - DDa[s] - g[s] * F) +
                       alp * (K * A[s] - 2.0 * AA) + lie - F2o3 * A[s] * divb;
      S.put(AXX + s, k + E.adv(AXX + s, 0.0, beta) + E.ko(AXX + s));
      if (CO) {  // d_k A_ij contracted with gu^jk into M_i (and M_j)
        double d[3];
        E.grad(AXX + s, d);
        M[i] += gu[sidx(j, 0)] * d[0] + gu[sidx(j, 1)] * d[1] + gu[sidx(j, 2)] * d[2];
        if (i != j) M[j] += gu[sidx(i, 0)] * d[0] + gu[sidx(i, 1)] * d[1] + gu[sidx(i, 2)] * d[2];
      }
      E.end();
      AMSS_FENCE();
    }
    // ---- constraint residuals (:1044-1134) ----------------------------------
    if (CO) {
      S.res(0, chi1 * tr(R) + F2o3 * K * K - trAA - 16.0 * kPi * rho);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double m = M[i];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double D = dchi[k] * A[sidx(i, j)] * ichi;
#pragma unroll
            for (int l = 0; l < 3; ++l) D += G2[l][sidx(k, i)] * A[sidx(l, j)] + G2[l][sidx(k, j)] * A[sidx(i, l)];
            m -= gu[sidx(j, k)] * D;
          }
        S.res(1 + i, m - F2o3 * dK[i] - 8.0 * kPi * Sv[i]);
      }
    }
  }
}

}  // namespace amss
