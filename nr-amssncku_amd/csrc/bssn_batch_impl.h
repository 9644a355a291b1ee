// bssn_batch_impl.h -- the tile-list kernel of the batched level step,
// instantiated per gauge and tile shape in bssn_batch_g*_s*.hip.
#pragma once
#include "bssn_batch.h"
#include "bssn_kernels_impl.h"

namespace amss {

// One tile of the list: the same LDS-staged point computation as
// rk4_lds_kernel, with the fields of the tile's patch addressed through the
// slab and the role slots.
template <int GAUGE, int STAGE, bool CO, int TX, int TY, int TZ>
__device__ __forceinline__ void tile_body(const PatchDesc* __restrict__ P, const TileRef& r, const Roles& m,
                                          const Coef& C, const Rk4Param& R, int* flag, lds_double* b0, int tid) {
  using Tl = LdsTile<TX, TY, TZ>;
  const int tx = tid % TX, ty = tile_row((tid / TX) % TY, TY), tz = tid / (TX * TY);
  const int ltv = (tx + 3) + Tl::LX * ((ty + 3) + Tl::LY * (tz + 3));
  const PatchDesc d = P[r.patch];
  const long p = (r.i0 + tx) + (long)(r.j0 + ty) * d.sy + (long)(r.k0 + tz) * d.sz;
  const long ln = (long)NFIELD * d.n;
  gdouble* slab = (gdouble*)d.slab;
  const SlabFields in{slab + m.in * ln, d.n}, u0{slab + m.u0 * ln, d.n};
  const SlabFields acc{slab + m.acc * ln, d.n}, out{slab + m.out * ln, d.n};
  const SlabAux aux{slab + kResidualBase * d.n, d.n};
  LdsRk4<STAGE, TX, TY, TZ, SlabFields, SlabFields, SlabAux> X{
      {in, in, C, b0, b0 + (kLdsBuffers / 2) * Tl::NPAD, p, ltv, (r.i0 - 3) + (r.j0 - 3) * d.sy + (r.k0 - 3) * d.sz,
       d.sy, d.sz, (int)d.n, tid},
      {u0, acc, out, aux, p, R.c, R.tinny, {}}};
  X.prologue();
  bssn_rhs_point<GAUGE, CO, false, false>(X, X);
  X.finish();
  if (X.bad) *flag = 1;
}

// PERSIST: a grid of resident workgroups strides over the tile list.  The
// LDS buffers need no extra barrier between tiles: the next tile's prologue
// fills pair 0, last read by groups 20 and 21, which every wave finished
// before the barrier that opened groups 22 and 23.  Otherwise one workgroup
// per tile (no loop: a loop makes the compiler keep more values across
// iterations and spill more).
template <int GAUGE, int STAGE, bool CO, int TX, int TY, int TZ, bool PERSIST>
__global__ void __launch_bounds__(TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(2, 2)))
rk4_tiles_kernel(const PatchDesc* __restrict__ P, const TileRef* __restrict__ T, int t0, int ntiles, Roles m,
                 Coef C, Rk4Param R, int remap, int* flag, FillWork fw) {
  using Tl = LdsTile<TX, TY, TZ>;
  __shared__ __attribute__((aligned(16))) double tiles[kLdsBuffers * Tl::NPAD];
  lds_double* b0 = (lds_double*)tiles;
  if (!PERSIST && (int)blockIdx.x >= ntiles) {
    // fill work: TX TY TZ / 256 items of 256 points per workgroup
    constexpr int kItems = TX * TY * TZ / 256;
    const int item = ((int)blockIdx.x - ntiles) * kItems + (int)threadIdx.x / 256;
    if (item < fw.n) fill_ghost_point(P, fw.B, fw.W[item], threadIdx.x % 256, m, STAGE, R.tinny, flag);
    return;
  }
  if (PERSIST) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
      // keep the compiler from hoisting per-thread LDS addresses out of the
      // loop (hundreds of loop-invariant addresses would occupy VGPRs)
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      const TileRef r = T[t0 + (remap ? xcd_spread(t, ntiles) : t)];
      tile_body<GAUGE, STAGE, CO, TX, TY, TZ>(P, r, m, C, R, flag, b0, tid);
    }
  } else {
    const int t = blockIdx.x;
    const TileRef r = T[t0 + (remap ? xcd_spread(t, ntiles) : t)];
    tile_body<GAUGE, STAGE, CO, TX, TY, TZ>(P, r, m, C, R, flag, b0, (int)threadIdx.x);
  }
}

template <int G, int TX, int TY, int TZ, bool PERSIST>
void launch_tiles_p(unsigned g, const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m,
                    const Coef& C, const Rk4Param& R, bool want_res, int remap, int* flag, hipStream_t s,
                    const FillWork& fw) {
  const int nt = TX * TY * TZ;
#define AMSS_TILES(ST, CO) \
  rk4_tiles_kernel<G, ST, CO, TX, TY, TZ, PERSIST><<<g, nt, 0, s>>>(P, T, t0, ntiles, m, C, R, remap, flag, fw)
  switch (R.stage) {
    case 0:
      if (want_res) AMSS_TILES(0, true);
      else AMSS_TILES(0, false);
      break;
    case 1: AMSS_TILES(1, false); break;
    case 2: AMSS_TILES(2, false); break;
    default: AMSS_TILES(3, false); break;
  }
#undef AMSS_TILES
}

// Fill work rides along only with one workgroup per tile (the caller runs it
// separately otherwise: launch_rk4_tiles_fuses_fill).
template <int G, int TX, int TY, int TZ>
void launch_rk4_tiles_shape(const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m,
                            const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune, int* flag,
                            hipStream_t s, const FillWork& fw) {
  if (ntiles <= 0) return;
  constexpr int kItems = TX * TY * TZ / 256;
  if (tune.grid > 0 && tune.grid < ntiles) {
    launch_tiles_p<G, TX, TY, TZ, true>((unsigned)tune.grid, P, T, t0, ntiles, m, C, R, want_res, tune.remap, flag, s,
                                        FillWork{});
  } else {
    const unsigned extra = (unsigned)((fw.n + kItems - 1) / kItems);
    launch_tiles_p<G, TX, TY, TZ, false>((unsigned)ntiles + extra, P, T, t0, ntiles, m, C, R, want_res, tune.remap,
                                         flag, s, fw);
  }
}

// One translation unit per gauge and tile shape (they compile in parallel).
#define AMSS_BATCH_TU(G, S)                                                                                  \
  void launch_rk4_tiles_g##G##_s##S(const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m, \
                                    const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune,  \
                                    int* flag, hipStream_t s, const FillWork& fw) {                          \
    launch_rk4_tiles_shape<G, tile_dim(S, 0), tile_dim(S, 1), tile_dim(S, 2)>(P, T, t0, ntiles, m, C, R,      \
                                                                              want_res, tune, flag, s, fw);  \
  }

}  // namespace amss
