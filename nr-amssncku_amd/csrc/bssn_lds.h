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
// the wave-uniform LDS base), so staging costs no VGPRs.  Centre values that
// the algebra reads outside a group come from global memory.
#pragma once
#include "bssn_env.h"

namespace amss {

// The order in which bssn_rhs_point opens its field groups.
__host__ __device__ constexpr int group_field(int g) {
  constexpr signed char seq[NFIELD] = {BETAX, BETAY, BETAZ, DXX, GXY, GXZ, DYY, GYZ,
                                       DZZ,   GAMX,  GAMY,  GAMZ, CHI, LAP, TRK, DTSFX,
                                       DTSFY, DTSFZ, AXX,   AXY,  AXZ, AYY, AYZ, AZZ};
  return seq[g];
}

// Pointers into LDS (32-bit local addresses; generic pointers would make
// every stencil offset a 64-bit flat address).
typedef __attribute__((address_space(3))) double lds_double;

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

template <int TX, int TY, int TZ, class FI = FieldsIn>
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

  __device__ __forceinline__ void prologue() {
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
      if (g + 2 < NFIELD) {
        issue(group_field(g + 2), slot(g + 2));
        issue(group_field(g + 3), slot(g + 3));
      }
    }
    ++g;
    if (v < 21) {
      const double x = cur()[lt];
      bad |= !(x == x);
    }
  }
  __device__ __forceinline__ void end() const {}

  __device__ __forceinline__ double at(int, long o) const { return cur()[lt + o]; }
  __device__ __forceinline__ double val(int v) const { return F.ptr(v)[p]; }
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

}  // namespace amss
