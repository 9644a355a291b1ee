// bssn_lds.h -- LDS-staged stencil environment for the fused RHS kernels.
//
// A 256-thread workgroup owns a TX x TY x TZ tile of interior points (one
// point per thread).  For every field group of the algebra the tile plus a
// 3-point halo of that field, (TX+6)(TY+6)(TZ+6) doubles, sits in LDS; every
// stencil of the group (gradient, Hessian incl. mixed, upwind, KO) reads it
// from there with compile-time offsets.  Two tile buffers alternate: while
// group g computes from one, the halo tile of group g+1 streams into the
// other with gfx950's direct global->LDS loads (global_load_lds_dwordx4:
// each lane moves 16 bytes, a wave fills a contiguous 1 KiB run starting at
// the wave-uniform LDS base), so staging costs no VGPRs.  The centre value of
// the open group's field comes from its tile; the other centre values a pair
// reads (and the RK4 epilogue's U0 and accumulator values of its fields) are
// loaded one pair ahead, so the barrier's full wait settles them and no
// later wait for them stalls on the halo loads in flight.
#pragma once
#include <type_traits>

#include "bssn_env.h"

namespace amss {

// The order in which bssn_rhs_point opens its field groups.
__host__ __device__ constexpr int group_field(int g) {
  constexpr signed char seq[NFIELD] = {BETAX, BETAY, BETAZ, DXX, GXY, GXZ, DYY, GYZ,
                                       DZZ,   GAMX,  GAMY,  GAMZ, CHI, LAP, TRK, DTSFX,
                                       DTSFY, DTSFZ, AXX,   AXY,  AXZ, AYY, AYZ, AZZ};
  return seq[g];
}

// The centre value of another field that the point algebra reads inside a
// group (the shift rhs reads B^i, the metric rhs A_ij), or -1.
__host__ __device__ constexpr int group_extra(int v) {
  return v >= BETAX && v <= BETAX + 2 ? DTSFX + (v - BETAX) : (v >= DXX && v <= DZZ ? AXX + (v - DXX) : -1);
}

// Pointers into LDS (32-bit local addresses; generic pointers would make
// every stencil offset a 64-bit flat address).
typedef __attribute__((address_space(3))) double lds_double;

#ifndef AMSS_ROWPERM
#define AMSS_ROWPERM 1
#endif
// Tile row of a thread's row index r (of TY): with 8 rows a 16-lane LDS
// access group can hold rows y and y + 4, whose 14-double pitch puts them
// 112 dwords apart (16 mod 32 banks).
__host__ __device__ constexpr int tile_row(int r, int TY) {
  return (AMSS_ROWPERM && TY == 8) ? ((r & 1) ? 4 + (r >> 1) : (r >> 1)) : r;
}

template <int TX, int TY, int TZ>
struct LdsTile {
  static constexpr int LX = TX + 6, LY = TY + 6, LZ = TZ + 6;
  static constexpr int NE = LX * LY * LZ;  // doubles per field tile
  static constexpr int NCH = NE / 2;       // 16-byte chunks
  static constexpr int NT = TX * TY * TZ;  // threads
  static constexpr int ROUNDS = (NCH + NT - 1) / NT;
  static constexpr int NPAD = 2 * ROUNDS * NT;  // doubles per buffer: whole rounds, no bounds branches
  static_assert(LX % 2 == 0, "tile rows must hold whole 16-byte chunks");
  static_assert(NT % 64 == 0, "whole waves");
};

#ifndef AMSS_LDS_FENCE
#define AMSS_LDS_FENCE 1
#endif
#ifndef AMSS_LDS_PAIRS
#define AMSS_LDS_PAIRS 1
#endif
// LDS tile buffers per workgroup: two pairs (two field groups per barrier)
// or two single buffers.
constexpr int kLdsBuffers = AMSS_LDS_PAIRS ? 4 : 2;

#ifndef AMSS_TOP
#define AMSS_TOP 1
#endif
#ifndef AMSS_DMA_AT
#define AMSS_DMA_AT 1
#endif
// Self: a type deriving from this environment and from the RK4 epilogue,
// whose prefetch(v0, v1) the environment calls for each pair (void: none).
template <int TX, int TY, int TZ, class FI = FieldsIn, class Self = void>
struct LdsEnv {
  static constexpr bool kFence = AMSS_LDS_FENCE;
  using T = LdsTile<TX, TY, TZ>;
  static constexpr long SY = T::LX, SZ = (long)T::LX * T::LY;

  const FI& F;
  const FI& M;
  const Coef& C;
  lds_double* buf0;   // two pairs of LDS tile buffers (2 x NPAD doubles each)
  lds_double* buf1;
  long p;         // global index of this thread's point (clamped into the array)
  int lt;         // LDS index of this thread's point
  int gbase;      // global index of tile element (0,0,0)
  int sy, sz;     // global strides
  int gmax;       // number of points of one field
  int tid;        // thread index in the workgroup
  int g = 0;      // next group to open
  bool bad = false;  // a NaN among the centre values of the first 21 fields (bssn_rhs.f90:110-113)
  double top[NFIELD];  // centre values the algebra reads before its first group (in_top)
  int xf[2] = {-1, -1}, nxf[2] = {-1, -1};  // group_extra of the open (next) pair's fields
  double xv[2], nxv[2];                      // ... and their centre values, loaded a pair ahead

  // Stream the halo tile of field v into buffer b: every lane issues ROUNDS
  // 16-byte direct global->LDS loads.  The buffer is padded to whole rounds,
  // so chunks past the tile land in the padding (reading a clamped, valid
  // address) and the loop has no bounds branches -- branches here split the
  // point kernel into many blocks and wreck its register allocation.
  __device__ __forceinline__ void issue(int v, lds_double* b) const {
    const auto base = F.ptr(v);
    const int wave = tid >> 6;
#pragma unroll
    for (int r = 0; r < T::ROUNDS; ++r) {
      const int c = r * T::NT + tid;
      const int e = 2 * c;
      const int x = e % T::LX, y = (e / T::LX) % T::LY, z = e / (T::LX * T::LY);
      int gi = gbase + x + y * sy + z * sz;  // < 2^29 points per field
      gi = gi < 0 ? 0 : (gi > gmax - 2 ? gmax - 2 : gi);
      __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(base + gi),
                                       (__attribute__((address_space(3))) void*)(b + 2 * (r * T::NT + wave * 64)),
                                       16, 0, 0);
    }
  }
  // Groups are staged two at a time: groups 2q and 2q+1 share pair q & 1,
  // so a tile needs one barrier per two field groups.
  __device__ __forceinline__ lds_double* slot(int grp) const {
    if (!AMSS_LDS_PAIRS) return (grp & 1) ? buf1 : buf0;
    return (((grp >> 1) & 1) ? buf1 : buf0) + (grp & 1) * T::NPAD;
  }
  __device__ __forceinline__ lds_double* cur() const { return slot(g - 1); }  // g already advanced

  // Centre values a pair needs from global memory, loaded one pair ahead
  // (into the n* slots; rotate makes them current at the pair's barrier).
  __device__ __forceinline__ void ahead(int g0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      nxf[h] = group_extra(group_field(g0 + h));
      if (nxf[h] >= 0) nxv[h] = F.ptr(nxf[h])[p];
    }
    if constexpr (!std::is_void_v<Self>) static_cast<Self*>(this)->prefetch(group_field(g0), group_field(g0 + 1));
  }
  __device__ __forceinline__ void rotate() {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      xf[h] = nxf[h];
      xv[h] = nxv[h];
      nxf[h] = -1;
    }
    if constexpr (!std::is_void_v<Self>) static_cast<Self*>(this)->sink_rotate();
  }

  // the centre values bssn_rhs_point reads before it opens a group
  __device__ __forceinline__ static constexpr bool in_top(int v) {
    return v == CHI || v == TRK || v == LAP || (v >= DXX && v <= AZZ) || (v >= BETAX && v <= BETAX + 2);
  }
  __device__ __forceinline__ void prologue() {
    if (AMSS_TOP) {
#pragma unroll
      for (int v = 0; v < NFIELD; ++v)
        if (in_top(v)) top[v] = F.ptr(v)[p];
    }
    if (AMSS_LDS_PAIRS) ahead(0);
    __builtin_amdgcn_sched_barrier(0);
    issue(group_field(0), slot(0));
    if (AMSS_LDS_PAIRS) issue(group_field(1), slot(1));
  }
  __device__ __forceinline__ void begin(int v) {
    if (!AMSS_LDS_PAIRS) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (g + 1 < NFIELD) issue(group_field(g + 1), slot(g + 1));
    } else if ((g & 1) == 0) {
      __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA of groups g, g+1 (and everything else) landed
      __syncthreads();                // ... for every wave; and the readers of groups g-2, g-1 are done
      rotate();
      if (g + 2 < NFIELD) ahead(g + 2);
      __builtin_amdgcn_sched_barrier(0);
      if (g + 2 < NFIELD) {
        if (AMSS_DMA_AT < 2) issue(group_field(g + 2), slot(g + 2));
        if (AMSS_DMA_AT < 1) issue(group_field(g + 3), slot(g + 3));
      }
    } else if (AMSS_DMA_AT > 0 && g + 1 < NFIELD) {
      // the next pair's halo loads start with the pair's second group
      __builtin_amdgcn_sched_barrier(0);
      if (AMSS_DMA_AT > 1) issue(group_field(g + 1), slot(g + 1));
      issue(group_field(g + 2), slot(g + 2));
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (!std::is_void_v<Self>) static_cast<Self*>(this)->sink_flush();
    ++g;
    if (v < 21) {
      const double x = cur()[lt];
      bad |= !(x == x);
    }
  }
  __device__ __forceinline__ void end() const {}

  __device__ __forceinline__ double at(int, long o) const { return cur()[lt + o]; }
  // the centre value of the open group's field is in its halo tile; the
  // pair's extra centre values were loaded at its barrier
  __device__ __forceinline__ double val(int v) const {
    if (AMSS_TOP && g == 0 && in_top(v)) return top[v];
    if (g > 0 && v == group_field(g - 1)) return cur()[lt];
    if (v == xf[0]) return xv[0];
    if (v == xf[1]) return xv[1];
    return F.ptr(v)[p];
  }
  __device__ __forceinline__ double mat(int m) const { return M.ptr(m)[p]; }

  __device__ __forceinline__ double d1(int v, long s) const {
    return (at(v, -2 * s) - at(v, 2 * s)) + 8.0 * (at(v, s) - at(v, -s));
  }
  __device__ __forceinline__ void grad(int v, double d[3]) const {
    d[0] = C.d1[0] * d1(v, 1);
    d[1] = C.d1[1] * d1(v, SY);
    d[2] = C.d1[2] * d1(v, SZ);
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
    d[3] = C.d2[1] * (16.0 * (at(v, -SY) + at(v, SY)) - (at(v, -2 * SY) + at(v, 2 * SY)) - c);
    d[5] = C.d2[2] * (16.0 * (at(v, -SZ) + at(v, SZ)) - (at(v, -2 * SZ) + at(v, 2 * SZ)) - c);
    d[1] = C.dm[0] * mix(v, 1, SY);
    d[2] = C.dm[1] * mix(v, 1, SZ);
    d[4] = C.dm[2] * mix(v, SY, SZ);
  }
  __device__ __forceinline__ double up1(int v, long s, double b, double c) const {
    const double f0 = 10.0 * at(v, 0);
    const double up = (at(v, 3 * s) - 3.0 * at(v, -s)) + (18.0 * at(v, s) - 6.0 * at(v, 2 * s)) - f0;
    const double dn = (at(v, -3 * s) - 3.0 * at(v, s)) + (18.0 * at(v, -s) - 6.0 * at(v, -2 * s)) - f0;
    return c * (fmax(b, 0.0) * up - fmin(b, 0.0) * dn);
  }
  // the two one-sided stencils of every direction, unscaled (adjoint use)
  __device__ __forceinline__ void updn(int v, double up[3], double dn[3]) const {
    const long st[3] = {1, SY, SZ};
    const double f0 = 10.0 * at(v, 0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const long s = st[a];
      up[a] = (at(v, 3 * s) - 3.0 * at(v, -s)) + (18.0 * at(v, s) - 6.0 * at(v, 2 * s)) - f0;
      dn[a] = (at(v, -3 * s) - 3.0 * at(v, s)) + (18.0 * at(v, -s) - 6.0 * at(v, -2 * s)) - f0;
    }
  }
  __device__ __forceinline__ double adv(int v, double, const double b[3]) const {
    return up1(v, 1, b[0], C.d1[0]) + up1(v, SY, b[1], C.d1[1]) + up1(v, SZ, b[2], C.d1[2]);
  }
  __device__ __forceinline__ double ko1(int v, long s, double c20) const {
    return (at(v, -3 * s) + at(v, 3 * s)) - 6.0 * (at(v, -2 * s) + at(v, 2 * s)) +
           15.0 * (at(v, -s) + at(v, s)) - c20;
  }
  __device__ __forceinline__ double ko(int v) const {
    const double c20 = 20.0 * at(v, 0);
    return C.ko[0] * ko1(v, 1, c20) + C.ko[1] * ko1(v, SY, c20) + C.ko[2] * ko1(v, SZ, c20);
  }
};

// The LDS environment and the RK4 epilogue as one object (passed to the point
// algebra as both), so the environment can hand the epilogue its prefetches.
// DEFER: the last put's stores wait for the next group boundary, where the
// environment calls sink_flush after its barrier -- a store issued just
// before a barrier would be waited for there (its s_waitcnt(0) waits for
// every memory operation), one issued after it has a group's computation to
// land.  Measured on one box against the immediate stores: the single-patch
// kernel +0.9%, the tile-list kernel -1.4%, so only the former defers; and
// sub-step 3, one store per field, gains nothing.
template <int STAGE, int TX, int TY, int TZ, class FI = FieldsIn, class FO = FieldsOut, class AX = AuxOut,
          bool DEFER = false>
struct LdsRk4 : LdsEnv<TX, TY, TZ, FI, LdsRk4<STAGE, TX, TY, TZ, FI, FO, AX, DEFER>>, Rk4SinkT<STAGE, FI, FO, AX> {
  using Sink = Rk4SinkT<STAGE, FI, FO, AX>;
  static constexpr bool kDefer = DEFER && STAGE < 3;
  int dv = -1;
  double dacc, dout;
  __device__ __forceinline__ void sink_rotate() { Sink::rotate(); }
  __device__ __forceinline__ void prefetch(int v0, int v1) { Sink::prefetch(v0, v1); }
  __device__ __forceinline__ void sink_flush() {
    if (!kDefer || dv < 0) return;
    if (STAGE < 3) Sink::ACC.ptr(dv)[Sink::p] = dacc;
    if (!(STAGE < 3 && dv >= DXX && dv <= AZZ)) Sink::OUT.ptr(dv)[Sink::p] = dout;
    dv = -1;
  }
  __device__ __forceinline__ void put(int v, double k) {
    if (!kDefer) {
      Sink::put(v, k);
      return;
    }
    double u0v, av, accv, u;
    Sink::u0_acc(v, u0v, av);
    Sink::rk4_values(v, k, u0v, av, accv, u);
    sink_flush();  // an earlier put since the last boundary
    dv = v;
    dacc = accv;
    if (STAGE < 3 && v >= DXX && v <= AZZ)
      Sink::gA[v - DXX] = u;
    else
      dout = u;
  }
  __device__ __forceinline__ void finish() {
    sink_flush();
    Sink::finish();
  }
};

}  // namespace amss
