// bssn_env.h -- stencil environments and result sinks for bssn_rhs_point.
#pragma once
#include "bssn_algebra.h"

namespace amss {

// ---------------------------------------------------------------------------
// Interior environment: the point has +-3 room along every axis, so every
// stencil is the full-order one (diff_new.f90:1084-1087, 1487-1507;
// lopsidediff.f90:324-372 first branches; kodiss.f90:166-207).  Reads go
// straight to memory through the vector L1/L2 as raw buffer loads: the field
// base lives in a scalar buffer descriptor and the (uniform) stencil offset
// in the scalar soffset, so every load of every field shares one per-lane
// offset register -- with flat addressing each field and each runtime
// stride would cost a 64-bit per-lane address pair.
// Requires 3 + 3 sy + 3 sz <= p and field size < 4 GiB.
struct InteriorEnv {
  static constexpr bool kFence = true, kFeatures = false;
  const FieldsIn& F;
  const FieldsIn& M;
  const Coef& C;
  long p, sy, sz;

  __device__ __forceinline__ double at(int v, long o) const {
    const __amdgpu_buffer_rsrc_t r =
        __builtin_amdgcn_make_buffer_rsrc((void*)F.p[v], (short)0, (int)0xFFFFFFFF, 0x00020000);
    const int q = 3 + 3 * (int)sy + 3 * (int)sz;  // most negative stencil offset, negated
    const auto t = __builtin_amdgcn_raw_buffer_load_b64(r, (unsigned)((p - q) * 8), (unsigned)((o + q) * 8), 0);
    return __builtin_bit_cast(double, t);
  }
  __device__ __forceinline__ double val(int v) const { return at(v, 0); }
  __device__ __forceinline__ double mat(int m) const { return M.p[m][p]; }
  __device__ __forceinline__ void begin(int) const {}
  __device__ __forceinline__ void end() const {}

  __device__ __forceinline__ double d1(int v, long s) const {
    return (at(v, -2 * s) - at(v, 2 * s)) + 8.0 * (at(v, s) - at(v, -s));
  }
  __device__ __forceinline__ void grad(int v, double d[3]) const {
    d[0] = C.d1[0] * d1(v, 1);
    d[1] = C.d1[1] * d1(v, sy);
    d[2] = C.d1[2] * d1(v, sz);
  }
  __device__ __forceinline__ double row(int v, long o, long s) const {
    return (at(v, o - 2 * s) - at(v, o + 2 * s)) + 8.0 * (at(v, o + s) - at(v, o - s));
  }
  __device__ __forceinline__ double mix(int v, long sa, long sb) const {
    return (row(v, -2 * sb, sa) - row(v, 2 * sb, sa)) + 8.0 * (row(v, sb, sa) - row(v, -sb, sa));
  }
  __device__ __forceinline__ void hess(int v, double d[6]) const {
    const double c = 30.0 * at(v, 0);
    d[0] = C.d2[0] * (16.0 * (at(v, -1) + at(v, 1)) - (at(v, -2) + at(v, 2)) - c);
    d[3] = C.d2[1] * (16.0 * (at(v, -sy) + at(v, sy)) - (at(v, -2 * sy) + at(v, 2 * sy)) - c);
    d[5] = C.d2[2] * (16.0 * (at(v, -sz) + at(v, sz)) - (at(v, -2 * sz) + at(v, 2 * sz)) - c);
    d[1] = C.dm[0] * mix(v, 1, sy);
    d[2] = C.dm[1] * mix(v, 1, sz);
    d[4] = C.dm[2] * mix(v, sy, sz);
  }
  // one upwind direction: beta>0 uses -3,-10,18,-6,1 on (-1..3), beta<0 the mirror
  __device__ __forceinline__ double up1(int v, long s, double b, double c) const {
    const double f0 = 10.0 * at(v, 0);
    const double up = (at(v, 3 * s) - 3.0 * at(v, -s)) + (18.0 * at(v, s) - 6.0 * at(v, 2 * s)) - f0;
    const double dn = (at(v, -3 * s) - 3.0 * at(v, s)) + (18.0 * at(v, -s) - 6.0 * at(v, -2 * s)) - f0;
    // branch-free: beta>0 -> beta c up ; beta<0 -> -beta c dn ; beta==0 -> 0
    return c * (fmax(b, 0.0) * up - fmin(b, 0.0) * dn);
  }
  __device__ __forceinline__ double adv(int v, double /*off*/, const double b[3]) const {
    return up1(v, 1, b[0], C.d1[0]) + up1(v, sy, b[1], C.d1[1]) + up1(v, sz, b[2], C.d1[2]);
  }
  // the two one-sided stencils of every direction, unscaled (adjoint use)
  __device__ __forceinline__ void updn(int v, double up[3], double dn[3]) const {
    const long st[3] = {1, sy, sz};
    const double f0 = 10.0 * at(v, 0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const long s = st[a];
      up[a] = (at(v, 3 * s) - 3.0 * at(v, -s)) + (18.0 * at(v, s) - 6.0 * at(v, 2 * s)) - f0;
      dn[a] = (at(v, -3 * s) - 3.0 * at(v, s)) + (18.0 * at(v, -s) - 6.0 * at(v, -2 * s)) - f0;
    }
  }
  __device__ __forceinline__ double ko1(int v, long s, double c20) const {
    return (at(v, -3 * s) + at(v, 3 * s)) - 6.0 * (at(v, -2 * s) + at(v, 2 * s)) +
           15.0 * (at(v, -s) + at(v, s)) - c20;
  }
  __device__ __forceinline__ double ko(int v) const {
    const double c20 = 20.0 * at(v, 0);
    return C.ko[0] * ko1(v, 1, c20) + C.ko[1] * ko1(v, sy, c20) + C.ko[2] * ko1(v, sz, c20);
  }
};

// ---------------------------------------------------------------------------
// The reference's stencils at a boundary point (Symmetry = 0): the
// derivative order drops for all three directions together when any axis
// lacks +-2 (then +-1) room, the last index of any axis gets zero
// derivatives and no advection, the upwind stencil falls back per direction,
// KO needs +-3 room on every axis, and the halo index -1 mirrors index 0
// with the field's parity (symmetry_bd, fmisc.f90:877-901).
// ORD: the order of the centred stencils at the point (2, 1 or 0), fixed per
// launch by the shell the point lies in, so that each launch compiles only
// its own stencils.  D supplies rel(v, di, dj, dk, off): field v (plus off)
// at the point's neighbour (i + di, j + dj, k + dk), which every stencil
// here keeps within [lo, n - 1] per axis and within +-3 of the point.
template <int ORD, class D>
struct EdgeStencils {
  static constexpr int ord = ORD;
  const Coef& C;
  int i, j, k;
  int n[3];
  int lo[3];  // lowest readable index per axis (0, or -3 on a symmetry plane)
  bool last, kook;

  __device__ EdgeStencils(const Coef& C_, int i_, int j_, int k_, const int* n_, const int* lo_)
      : C(C_), i(i_), j(j_), k(k_) {
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      n[a] = n_[a];
      lo[a] = lo_[a];
    }
    last = (i == n[0] - 1) || (j == n[1] - 1) || (k == n[2] - 1);
    // the joint three-axis room test of kodis (kodiss.f90:162-212); the
    // stencil order's test (diff_new.f90:996-1100, 1397-1524) chose ORD
    auto rm = [&](int r) {
      return i - r >= lo[0] && i + r <= n[0] - 1 && j - r >= lo[1] && j + r <= n[1] - 1 && k - r >= lo[2] &&
             k + r <= n[2] - 1;
    };
    kook = ORD == 2 && rm(3);
  }
  __device__ __forceinline__ const D& self() const { return *static_cast<const D*>(this); }
  __device__ double at3(int v, int di, int dj, int dk) const { return self().rel(v, di, dj, dk, 0.0); }

  __device__ double ax(int v, int axis, int m, double off = 0.0) const {
    return axis == 0 ? self().rel(v, m, 0, 0, off) : (axis == 1 ? self().rel(v, 0, m, 0, off) : self().rel(v, 0, 0, m, off));
  }
  __device__ void grad(int v, double d[3]) const {
    #pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (ord == 2)
        d[a] = C.d1[a] * ((ax(v, a, -2) - ax(v, a, 2)) + 8.0 * (ax(v, a, 1) - ax(v, a, -1)));
      else if (ord == 1)
        d[a] = C.d1lo[a] * (ax(v, a, 1) - ax(v, a, -1));
      else
        d[a] = 0.0;
    }
  }
  __device__ void hess(int v, double d[6]) const {
    #pragma unroll
    for (int s = 0; s < 6; ++s) d[s] = 0.0;
    if (ord == 0) return;
    const double c = at3(v, 0, 0, 0);
    if (ord == 2) {
      #pragma unroll
      for (int a = 0; a < 3; ++a)
        d[sidx(a, a)] = C.d2[a] * (16.0 * (ax(v, a, -1) + ax(v, a, 1)) - (ax(v, a, -2) + ax(v, a, 2)) - 30.0 * c);
      const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
      #pragma unroll
      for (int q = 0; q < 3; ++q) {
        double acc = 0.0;
        const double w[5] = {1.0, -8.0, 0.0, 8.0, -1.0};
        #pragma unroll
        for (int bb = -2; bb <= 2; ++bb) {
          if (bb == 0) continue;
          double r = 0.0;
          #pragma unroll
          for (int aa = -2; aa <= 2; ++aa) {
            if (aa == 0) continue;
            int o[3] = {0, 0, 0};
            o[pa[q]] = aa;
            o[pb[q]] = bb;
            r += w[aa + 2] * at3(v, o[0], o[1], o[2]);
          }
          acc += w[bb + 2] * r;
        }
        d[sidx(pa[q], pb[q])] = C.dm[q] * acc;
      }
    } else {
      #pragma unroll
      for (int a = 0; a < 3; ++a)
        d[sidx(a, a)] = C.d2lo[a] * (ax(v, a, -1) + ax(v, a, 1) - 2.0 * c);
      const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
      #pragma unroll
      for (int q = 0; q < 3; ++q) {
        int o[3] = {0, 0, 0};
        auto f = [&](int a, int b) {
          o[0] = o[1] = o[2] = 0;
          o[pa[q]] = a;
          o[pb[q]] = b;
          return at3(v, o[0], o[1], o[2]);
        };
        d[sidx(pa[q], pb[q])] = C.dmlo[q] * ((f(-1, -1) - f(1, -1)) - (f(-1, 1) - f(1, 1)));
      }
    }
  }
  // lopsidediff.f90:321-481, one direction
  __device__ double adv1(int v, int a, double b, double off) const {
    const int pos = a == 0 ? i : (a == 1 ? j : k);
    const int nn = n[a];
    auto up = [&]() {
      return -3.0 * ax(v, a, -1, off) - 10.0 * ax(v, a, 0, off) + 18.0 * ax(v, a, 1, off) - 6.0 * ax(v, a, 2, off) +
             ax(v, a, 3, off);
    };
    auto dn = [&]() {
      return -3.0 * ax(v, a, 1, off) - 10.0 * ax(v, a, 0, off) + 18.0 * ax(v, a, -1, off) - 6.0 * ax(v, a, -2, off) +
             ax(v, a, -3, off);
    };
    auto ce = [&]() { return ax(v, a, -2, off) - 8.0 * ax(v, a, -1, off) + 8.0 * ax(v, a, 1, off) - ax(v, a, 2, off); };
    const double c = C.d1[a];
    if (b > 0.0) {
      if (pos + 3 <= nn - 1) return b * c * up();
      if (pos + 2 <= nn - 1) return b * c * ce();
      if (pos + 1 <= nn - 1) return -b * c * dn();
    } else if (b < 0.0) {
      if (pos - 3 >= lo[a]) return -b * c * dn();
      if (pos - 2 >= lo[a]) return b * c * ce();
      if (pos - 1 >= lo[a]) return b * c * up();
    }
    return 0.0;
  }
  __device__ double adv(int v, double off, const double b[3]) const {
    if (last) return 0.0;
    return adv1(v, 0, b[0], off) + adv1(v, 1, b[1], off) + adv1(v, 2, b[2], off);
  }
  __device__ double ko(int v) const {
    if (!kook) return 0.0;
    const double c20 = 20.0 * at3(v, 0, 0, 0);
    double r = 0.0;
    #pragma unroll
    for (int a = 0; a < 3; ++a)
      r += C.ko[a] * ((ax(v, a, -3) + ax(v, a, 3)) - 6.0 * (ax(v, a, -2) + ax(v, a, 2)) +
                      15.0 * (ax(v, a, -1) + ax(v, a, 1)) - c20);
    return r;
  }
};

// Edge environment: the boundary stencils on the arrays in global memory.
template <int ORD>
struct EdgeEnv : EdgeStencils<ORD, EdgeEnv<ORD>> {
  static constexpr bool kFence = true;
  const FieldsIn& F;
  const FieldsIn& M;
  long sy, sz;

  __device__ EdgeEnv(const FieldsIn& F_, const FieldsIn& M_, const Coef& C_, int i_, int j_, int k_,
                     const int* n_, const int* lo_, long sy_, long sz_)
      : EdgeStencils<ORD, EdgeEnv<ORD>>(C_, i_, j_, k_, n_, lo_), F(F_), M(M_), sy(sy_), sz(sz_) {}
  __device__ double rel(int v, int di, int dj, int dk, double off) const {
    int a = this->i + di, b = this->j + dj, c = this->k + dk;
    double s = 1.0;
    if (a < 0) { a = -a - 1; s *= parity(v, 0); }
    if (b < 0) { b = -b - 1; s *= parity(v, 1); }
    if (c < 0) { c = -c - 1; s *= parity(v, 2); }
    return (F.p[v][a + b * sy + c * sz] + off) * s;
  }
  __device__ double val(int v) const { return F.p[v][this->i + this->j * sy + this->k * sz]; }
  __device__ double mat(int m) const { return M.p[m][this->i + this->j * sy + this->k * sz]; }
  __device__ void begin(int) const {}
  __device__ void end() const {}
};

// ---------------------------------------------------------------------------
// Extra outputs of the reference ABI: physical Christoffels (18), physical
// Ricci (6) and constraint residuals (7: ham, mom x3, Gam x3).
struct AuxOut {
  double* __restrict__ gam[18];
  double* __restrict__ ric[6];
  double* __restrict__ res[7];
  __device__ __forceinline__ double* resp(int a) const { return res[a]; }
};
// The 7 residual fields inside a Block's slab: residual a at base + a * n.
struct SlabAux {
  gdouble* base;
  long n;
  __device__ __forceinline__ gdouble* resp(int a) const { return base + a * n; }
};

// Plain RHS sink: k -> rhs arrays.
struct RhsSink {
  const FieldsOut& O;
  const AuxOut& X;
  long p;
  __device__ __forceinline__ void put(int v, double k) { O.p[v][p] = k; }
  __device__ __forceinline__ void gam(int a, double x) { X.gam[a][p] = x; }
  __device__ __forceinline__ void ric(int a, double x) { X.ric[a][p] = x; }
  __device__ __forceinline__ void res(int a, double x) { X.res[a][p] = x; }
  __device__ __forceinline__ void finish() {}
};

// enforce_ga (enforce_algebra.f90:72-127) on one point, offsets in place:
// rescale g to unit determinant, then remove the trace of A taken with the
// cofactors of the rescaled metric.
__device__ __forceinline__ void enforce_point(double d[6], double a[6]) {
  double g[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) g[s] = d[s] + (sdiag(s) ? 1.0 : 0.0);
  const double det = g[0] * g[3] * g[5] + g[1] * g[4] * g[2] + g[2] * g[1] * g[4] - g[2] * g[3] * g[2] -
                     g[1] * g[1] * g[5] - g[0] * g[4] * g[4];
  const double sc = 1.0 / pow(det, 1.0 / 3.0);
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    g[s] *= sc;
    d[s] = sdiag(s) ? g[s] - 1.0 : g[s];
  }
  const double cxx = g[3] * g[5] - g[4] * g[4], cxy = g[2] * g[4] - g[1] * g[5], cxz = g[1] * g[4] - g[3] * g[2];
  const double cyy = g[0] * g[5] - g[2] * g[2], cyz = g[1] * g[2] - g[0] * g[4], czz = g[0] * g[3] - g[1] * g[1];
  const double trA = cxx * a[0] + cyy * a[3] + czz * a[5] + 2.0 * (cxy * a[1] + cxz * a[2] + cyz * a[4]);
#pragma unroll
  for (int s = 0; s < 6; ++s) a[s] -= (1.0 / 3.0) * g[s] * trA;
}

// RK4 sub-step sink (rungekutta4_rout.f90:110-128 fused into the RHS
// epilogue), followed by the lower bound on chi (fmisc.f90:1644-1657) and,
// for sub-steps 0-2, the enforce_ga that the next RHS call applies to its
// input (bssn_class.C:3415-3428).  Stage 3 leaves the new state un-enforced
// exactly as the reference does.
struct Rk4Sink {
  const FieldsIn& U0;
  const FieldsOut& ACC;
  const FieldsOut& OUT;
  const AuxOut& X;
  long p;
  int stage;
  double c;      // 0.5 dt, 0.5 dt, dt, dt/6
  double tinny;  // chitiny - 1
  double gA[12];

  __device__ __forceinline__ void put(int v, double k) {
    double kk = k;
    if (stage == 0)
      ACC.p[v][p] = k;
    else if (stage < 3)
      ACC.p[v][p] = ACC.p[v][p] + 2.0 * k;
    else
      kk = k + ACC.p[v][p];
    double u = U0.p[v][p] + c * kk;
    if (v == CHI) u = u < tinny ? tinny : u;
    if (stage < 3 && v >= DXX && v <= AZZ)
      gA[v - DXX] = u;
    else
      OUT.p[v][p] = u;
  }
  __device__ __forceinline__ void gam(int, double) {}
  __device__ __forceinline__ void ric(int, double) {}
  __device__ __forceinline__ void res(int a, double x) { X.res[a][p] = x; }
  __device__ __forceinline__ void finish() {
    if (stage < 3) {
      enforce_point(gA, gA + 6);
#pragma unroll
      for (int s = 0; s < 12; ++s) OUT.p[DXX + s][p] = gA[s];
    }
  }
};

// ---------------------------------------------------------------------------
// BAM Sommerfeld outer boundary (sommerfeld_rout.f90:223-449): on the
// outermost point layer of each physical face the RHS of field v becomes
// -s_v (x.grad f + f) / r, f the stage input, s_v the field's propagation
// speed, with second-order differences taken from the side an outgoing wave
// arrives from (x_a > 0: backward), centred or reversed where the array
// leaves no room.
struct SommerGeom {
  double x0[3] = {0.0, 0.0, 0.0};  // coordinate of index 0 along each axis
  double h[3] = {1.0, 1.0, 1.0};
  int faces = 0;           // bit a: the upper face of axis a is physical; bit 3 + a: the lower one
  int lo[3] = {0, 0, 0};   // lowest readable index: -1 below a symmetry plane (one mirrored layer)
  // faces (same bits) whose outermost layer the one-layer copy overwrites
  // right after the update (sommerfeld_rout's copy branch): its points need
  // neither the RHS nor the update
  int copy = 0;
};

// point (i, j, k) lies on the outermost layer of a face in the bit mask
__host__ __device__ __forceinline__ bool on_faces(int mask, const int n[3], int i, int j, int k) {
  const int p[3] = {i, j, k};
  bool on = false;
  for (int a = 0; a < 3; ++a)
    on = on || (((mask >> a) & 1) && p[a] == n[a] - 1) || (((mask >> (3 + a)) & 1) && p[a] == 0);
  return on;
}
__host__ __device__ __forceinline__ bool sommer_point(const SommerGeom& G, const int n[3], int i, int j, int k) {
  return on_faces(G.faces, n, i, j, k);
}

// Propagation speed: sqrt(2) for chi, trK and the lapse (the 1+log gauge
// speed, bssn_class.C:390-410), 1 for the other fields (var.C:30).
__host__ __device__ __forceinline__ double prop_speed(int v) {
  return (v == CHI || v == TRK || v == LAP) ? 1.4142135623730951 : 1.0;
}

// The boundary RHS at point p (sommerfeld_rout.f90:376-440).  rd(a, m): the
// stage input m points along axis a (|m| <= 2; index -1 is the mirror of 0).
// Where w = s x_a / r is 0 the reference leaves the derivative unset; x_a = 0
// there, so its term vanishes, and here it is 0.  Every neighbour the three
// difference forms may use is read before the form is chosen (one the array
// lacks room for reads the point itself and goes unused: with 5 or more
// points per axis the chosen form always has its room), so the reads do not
// wait on the choice.
template <class Rd>
__device__ __forceinline__ double sommer_rhs(const SommerGeom& G, const int n[3], const int p[3], double vel,
                                             Rd rd) {
  double x[3];
  for (int a = 0; a < 3; ++a) x[a] = G.x0[a] + p[a] * G.h[a];
  const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double f = rd(0, 0);
  double s = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double w = vel * x[a] / r, d2 = 0.5 / G.h[a];
    const bool room_dn2 = p[a] - 2 >= G.lo[a], room_dn1 = p[a] - 1 >= G.lo[a];
    const bool room_up2 = p[a] + 2 <= n[a] - 1, room_up1 = p[a] + 1 <= n[a] - 1;
    const double m2 = rd(a, room_dn2 ? -2 : 0), m1 = rd(a, room_dn1 ? -1 : 0);
    const double p1 = rd(a, room_up1 ? 1 : 0), p2 = rd(a, room_up2 ? 2 : 0);
    // 0: backward, 1: centred, 2: forward.  All three are formed and one
    // is selected (the empty asm keeps the compiler from sinking each into
    // a branch of its own), so the reads have no branch to wait behind.
    const bool up = w > 0.0, dn = w < 0.0;
    const bool use_b = (up && room_dn2) || (dn && !room_up2 && !room_up1);
    const bool use_c = (up && !room_dn2 && room_dn1) || (dn && !room_up2 && room_up1);
    const bool use_f = (up && !room_dn2 && !room_dn1) || (dn && room_up2);
    double db = d2 * (3.0 * f - 4.0 * m1 + m2);
    double dc = d2 * (-m1 + p1);
    double df = d2 * (-p2 + 4.0 * p1 - 3.0 * f);
    asm volatile("" : "+v"(db), "+v"(dc), "+v"(df));
    double d = 0.0;
    d = use_b ? db : d;
    d = use_c ? dc : d;
    d = use_f ? df : d;
    s += d * x[a];
  }
  return -(vel * (s + f)) / r;
}

// RK4 epilogue of a boundary-kernel point: on a Sommerfeld face layer each
// field's RHS is replaced by the boundary condition's before the update, as
// the reference applies sommerfeld_routbam to the stage's RHS before
// rungekutta4_rout (bssn_class.C:3106-3117, 3476-3484).
template <class Env, class Inner>
struct SommerRk4Sink {
  Inner S;
  const Env& E;
  const SommerGeom& G;
  bool on;
  __device__ __forceinline__ void put(int v, double k) {
    if (on) {
      const int p[3] = {E.i, E.j, E.k};
      k = sommer_rhs(G, E.n, p, prop_speed(v), [&](int a, int m) { return E.ax(v, a, m); });
    }
    S.put(v, k);
  }
  __device__ __forceinline__ void gam(int a, double x) { S.gam(a, x); }
  __device__ __forceinline__ void ric(int a, double x) { S.ric(a, x); }
  __device__ __forceinline__ void res(int a, double x) { S.res(a, x); }
  __device__ __forceinline__ void finish() { S.finish(); }
};

// The update's independent work items of one point: before sub-step 3 the
// metric and A fields go together (enforce_ga couples them; item 0) and each
// other field alone (items 1..12); in sub-step 3 every field alone.
template <int STAGE>
struct UpdateItems {
  static constexpr int kCount = STAGE < 3 ? 1 + (NFIELD - 12) : NFIELD;
  // the field of single-field item y (y >= 1 before sub-step 3)
  __host__ __device__ static constexpr int field(int y) {
    if (STAGE == 3) return y;
    constexpr signed char f[NFIELD - 12] = {CHI, TRK, GAMX, GAMY, GAMZ, LAP, BETAX, BETAY, BETAZ, DTSFX, DTSFY, DTSFZ};
    return f[y - 1];
  }
};

// Same epilogue with the sub-step fixed at compile time: the interior kernels
// use this form, since a runtime stage makes the compiler duplicate the whole
// point computation per stage branch (and blows the register budget).
template <int STAGE, class FI = FieldsIn, class FO = FieldsOut, class AX = AuxOut>
struct Rk4SinkT {
  static constexpr int kPin = STAGE == 0 ? 1 : 2;
  const FI& U0;
  const FO& ACC;
  const FO& OUT;
  const AX& X;
  long p;
  double c, tinny;
  double gA[12];
  // U0 and ACC at p of the fields of the LDS environment's open pair of
  // groups (p*) and of the next pair (n*), loaded one pair ahead: the
  // barrier's full wait has settled them before they are used, so the
  // epilogue does not wait for memory while the next halo tiles stream in.
  // (The Gamma-tilde and trK rhs, stored in later pairs, load theirs then:
  // keeping them live longer costs more in spills than it saves.)
  int pv[2] = {-1, -1}, nv[2] = {-1, -1};
  double pu[2], pa[2], nu[2], na[2];

  __device__ __forceinline__ void rotate() {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      pv[h] = nv[h];
      pu[h] = nu[h];
      pa[h] = na[h];
      nv[h] = -1;
    }
  }
  // v0, v1: the fields of the pair whose values these are
  __device__ __forceinline__ void prefetch(int v0, int v1) {
    const int v[2] = {v0, v1};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if ((v[h] >= GAMX && v[h] <= GAMZ) || v[h] == TRK) continue;
      nv[h] = v[h];
      nu[h] = U0.ptr(v[h])[p];
      na[h] = STAGE > 0 ? (double)ACC.ptr(v[h])[p] : 0.0;
    }
  }

  // U0 and (STAGE > 0) ACC of field v at p: prefetched, or loaded now
  __device__ __forceinline__ void u0_acc(int v, double& u0v, double& av) const {
    av = 0.0;
    if (v == pv[0]) {
      u0v = pu[0], av = pa[0];
    } else if (v == pv[1]) {
      u0v = pu[1], av = pa[1];
    } else {
      u0v = U0.ptr(v)[p];
      if (STAGE > 0) av = ACC.ptr(v)[p];
    }
  }
  // rungekutta4_rout at p for field v from its rhs k: the accumulator's new
  // value (sub-steps 0-2) and the stage output (chi bounded below)
  __device__ __forceinline__ void rk4_values(int v, double k, double u0v, double av, double& accv, double& u) const {
    double kk = k;
    accv = 0.0;
    if (STAGE == 0)
      accv = k;
    else if (STAGE < 3)
      accv = av + 2.0 * k;
    else
      kk = k + av;
    u = u0v + c * kk;
    if (v == CHI) u = u < tinny ? tinny : u;
  }
  __device__ __forceinline__ void put(int v, double k) {
    double u0v, av;
    u0_acc(v, u0v, av);
    put_loaded(v, k, u0v, av);
  }
  // the update of field v from its rhs k, U0 value u0v and (STAGE > 0)
  // accumulator value av
  __device__ __forceinline__ void put_loaded(int v, double k, double u0v, double av) {
    double accv, u;
    rk4_values(v, k, u0v, av, accv, u);
    if (STAGE < 3) ACC.ptr(v)[p] = accv;
    if (STAGE < 3 && v >= DXX && v <= AZZ)
      gA[v - DXX] = u;
    else
      OUT.ptr(v)[p] = u;
  }
  __device__ __forceinline__ void gam(int, double) {}
  __device__ __forceinline__ void ric(int, double) {}
  __device__ __forceinline__ void res(int a, double x) { X.resp(a)[p] = x; }
  __device__ __forceinline__ void finish() {
    if (STAGE < 3) {
      enforce_point(gA, gA + 6);
#pragma unroll
      for (int s = 0; s < 12; ++s) OUT.ptr(DXX + s)[p] = gA[s];
    }
  }
};

}  // namespace amss
