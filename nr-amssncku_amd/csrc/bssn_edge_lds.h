// bssn_edge_lds.h -- the boundary stencils on LDS windows: the RK4 sub-step
// on the shells of stencil order 2 and 1 (the points between the interior
// box and the outermost layer).
//
// The stencils, their fallbacks and the mirror below a symmetry plane are
// EdgeStencils' (the reference's fderivs/fdderivs/lopsided/kodis at the
// array edge); only the reads change: every field group of the algebra is
// staged, as the tile's window, into one of two LDS buffers by the
// direct global->LDS loads, the next group streaming into the other buffer
// while this one computes.  The many short branches of the boundary
// stencils then wait on LDS, not on memory.
//
// The sub-step is two launches over all tiles of both shells.  The stencil
// kernel stores each field's right-hand side tile-compact (field v of tile b
// at K[(b * NFIELD + v) * 256 + thread], coalesced whatever the slab's
// orientation); the update kernel then applies rungekutta4_rout, the chi
// floor and enforce_ga, per point and field.  Fused, the per-point RK4 loads and
// stores of an x-face -- one cache line per point -- would sit inside the
// stencil kernel's barriers, which wait for every store before the next
// field group.
#pragma once
#include "bssn_edge_tiles.h"
#include "bssn_kernels_impl.h"

namespace amss {

template <int ORD, int NRM>
struct EdgeLdsEnv : EdgeStencils<ORD, EdgeLdsEnv<ORD, NRM>> {
  static constexpr bool kFence = true;
  using Base = EdgeStencils<ORD, EdgeLdsEnv<ORD, NRM>>;
  using W = EdgeWindow<NRM>;
  const FieldsIn& F;
  const FieldsIn& M;
  lds_double* buf;  // two window buffers of W::PAD doubles
  long p;           // global index of the point (clamped into the slab where !live)
  int gbase;        // global index of window element (0, 0, 0)
  int sy, sz, gmax;
  int tid;
  int wlo[3];       // array index of window element (0, 0, 0)
  int mir;          // bit a: the window reaches below index 0 on a symmetry plane of axis a
  int lt;           // LDS index of the point
  // stencil reads address one base, the point's lowest neighbour (-3, -3,
  // -3), so every offset is a non-negative immediate, in both buffers
  static constexpr int kLow = 3 + 3 * W::P0 + 3 * W::P01;
  static_assert((W::PAD + 2 * kLow) * 8 < 65536, "one base reaches both buffers by 16-bit offsets");
  bool live;
  int g = 0;           // next group to open
  bool bad = false;    // a NaN among the centre values of the first 21 fields (bssn_rhs.f90:110-113)
  double top[NFIELD];  // centre values the algebra reads before its first group

  __device__ __forceinline__ EdgeLdsEnv(const FieldsIn& F_, const FieldsIn& M_, const Coef& C_, const int q[3],
                                        const Box& box, const EdgeTile& t, lds_double* buf_, int tid_, bool live_)
      : Base(C_, q[0], q[1], q[2], box.n, box.mlo), F(F_), M(M_), buf(buf_), tid(tid_), live(live_) {
    sy = (int)box.sy;
    sz = (int)box.sz;
    gmax = box.n[0] * box.n[1] * box.n[2];
    p = q[0] + q[1] * box.sy + q[2] * box.sz;
    mir = 0;
    for (int a = 0; a < 3; ++a) {
      wlo[a] = t.wlo[a];
      if (box.mlo[a] < 0 && wlo[a] < 0) mir |= 1 << a;
    }
    gbase = wlo[0] + wlo[1] * sy + wlo[2] * sz;
    lt = (q[0] - wlo[0]) + (q[1] - wlo[1]) * W::P0 + (q[2] - wlo[2]) * W::P01;
  }

  // Stream field v's window into buffer b: ROUNDS 16-byte direct loads per
  // lane.  Chunks past the window, the rows' padding, and window points
  // outside the array (which no stencil reads) load from a clamped valid
  // address.
  __device__ __forceinline__ void issue(int v, lds_double* b) const {
    const auto base = F.ptr(v);
    const int wave = tid >> 6;
#pragma unroll
    for (int r = 0; r < W::ROUNDS; ++r) {
      const int e = 2 * (r * kEdgeThreads + tid);
      const int x = e % W::P0, y = (e / W::P0) % W::W1, z = e / W::P01;
      int gi = gbase + x + y * sy + z * sz;
      gi = gi < 0 ? 0 : (gi > gmax - 2 ? gmax - 2 : gi);
      __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)(base + gi),
                                       (__attribute__((address_space(3))) void*)(b + 2 * (r * kEdgeThreads + wave * 64)),
                                       16, 0, 0);
    }
  }
  // Below a symmetry plane the window holds the mirror image of the field
  // with its parity (symmetry_bd, fmisc.f90:877-901): index -m-1 of axis a
  // is index m times the parity of v under reflection of a.  The image
  // cells read only cells at or above 0 on the mirrored axes, which this
  // pass does not write.
  __device__ __forceinline__ void mirror(int v, lds_double* b) const {
    for (int c = tid; c < W::NE; c += kEdgeThreads) {
      const int x[3] = {c % W::P0, (c / W::P0) % W::W1, c / W::P01};
      int src = 0;
      double s = 1.0;
      bool img = false;
      const int pitch[3] = {1, W::P0, W::P01};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        int m = wlo[a] + x[a];
        if (((mir >> a) & 1) && m < 0) {
          m = -m - 1;
          s *= parity(v, a);
          img = true;
        }
        src += (m - wlo[a]) * pitch[a];
      }
      if (img) b[c] = s * b[src];
    }
  }
  __device__ __forceinline__ lds_double* slot(int grp) const { return buf + (grp & 1) * W::PAD; }
  __device__ __forceinline__ lds_double* cur() const { return slot(g - 1); }  // g already advanced

  __device__ __forceinline__ static constexpr bool in_top(int v) {
    return v == CHI || v == TRK || v == LAP || (v >= DXX && v <= AZZ) || (v >= BETAX && v <= BETAX + 2);
  }
  __device__ __forceinline__ void prologue() {
#pragma unroll
    for (int v = 0; v < NFIELD; ++v)
      if (in_top(v)) top[v] = F.ptr(v)[p];
    __builtin_amdgcn_sched_barrier(0);
    issue(group_field(0), slot(0));
  }
  // one barrier per group: the group's window has landed for every wave,
  // and every wave is done with the previous group's buffer, which the next
  // group's loads then fill (a tile on a symmetry plane writes its mirror
  // cells first, behind a second barrier)
  __device__ __forceinline__ void begin(int v) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (g + 1 < NFIELD) issue(group_field(g + 1), slot(g + 1));
    if (mir) {
      mirror(v, slot(g));
      __syncthreads();
    }
    ++g;
    if (v < 21 && live) {
      const double x = cur()[lt];
      bad |= !(x == x);
    }
  }
  __device__ __forceinline__ void end() const {}

  // every neighbour the stencils read lies in the window (mirror cells included)
  __device__ __forceinline__ double rel(int, int di, int dj, int dk, double off) const {
    return (cur() + (lt - kLow))[kLow + di + dj * W::P0 + dk * W::P01] + off;
  }
  __device__ __forceinline__ double val(int v) const {
    if (g == 0 && in_top(v)) return top[v];
    if (g > 0 && v == group_field(g - 1)) return cur()[lt];
    return F.ptr(v)[p];
  }
  __device__ __forceinline__ double mat(int m) const { return M.ptr(m)[p]; }
};

// The stencil kernel's sink: each field's right-hand side to the tile-compact
// buffer, the constraint residuals (predictor only) to their arrays.
struct CompactSink {
  double* __restrict__ k;  // this tile's NFIELD x 256 slots
  const AuxOut& X;
  long p;
  int tid;
  bool live;
  __device__ __forceinline__ void put(int v, double x) { k[v * kEdgeThreads + tid] = x; }
  __device__ __forceinline__ void gam(int, double) {}
  __device__ __forceinline__ void ric(int, double) {}
  __device__ __forceinline__ void res(int a, double x) {
    if (live) X.res[a][p] = x;
  }
  __device__ __forceinline__ void finish() {}
};

#ifndef AMSS_EDGE_LDS_WAVES
#define AMSS_EDGE_LDS_WAVES 2
#endif
// LDS per workgroup: two window buffers of the largest window.
constexpr int kEdgeLdsDoubles = 2 * (EdgeWindow<0>::PAD > EdgeWindow<1>::PAD
                                         ? (EdgeWindow<0>::PAD > EdgeWindow<2>::PAD ? EdgeWindow<0>::PAD : EdgeWindow<2>::PAD)
                                         : (EdgeWindow<1>::PAD > EdgeWindow<2>::PAD ? EdgeWindow<1>::PAD : EdgeWindow<2>::PAD));

// The right-hand side on tile lb of a shell of stencil order ORD (1 or 2,
// whose points lie on no Sommerfeld or copy face: those are outermost-layer
// points) whose slabs have normal axis NRM; k: the tile's compact slots.
template <int GAUGE, bool CO, int ORD, int NRM>
__device__ __forceinline__ void edge_rhs_tile(const FieldsIn& in, const AuxOut& aux, const Box& box,
                                              const EdgeTiles& T, int lb, const Coef& C, double* k, int* flag,
                                              lds_double* win) {
  const EdgeTile t = edge_tile(T, lb);
  int q[3];
  bool live;
  edge_tile_point(t, (int)threadIdx.x, q, live);
  EdgeLdsEnv<ORD, NRM> E(in, in, C, q, box, t, win, (int)threadIdx.x, live);
  CompactSink S{k, aux, E.p, (int)threadIdx.x, live};
  E.prologue();
  bssn_rhs_point<GAUGE, CO, false, false>(E, S);
  if (E.bad) *flag = 1;
}

// The right-hand side on every tile of the order-1 and order-2 shells, one
// launch: the workgroup's block index picks the tile set, and with it the
// stencil order and window layout the code is compiled for, so all tiles
// progress together (each tile's time is a chain of one field-group load
// and barrier after another; separate launches would queue these chains).
template <int GAUGE, bool CO>
__global__ void __launch_bounds__(kEdgeThreads) __attribute__((amdgpu_waves_per_eu(AMSS_EDGE_LDS_WAVES, AMSS_EDGE_LDS_WAVES)))
rhs_edge_lds_kernel(FieldsIn in, AuxOut aux, Box box, EdgeTileSet T, Coef C, double* K, int* flag) {
  __shared__ __attribute__((aligned(16))) double win[kEdgeLdsDoubles];
  const int b = (int)blockIdx.x, s = T.set_of(b), lb = b - T.first[s];
  double* k = K + (long)b * NFIELD * kEdgeThreads;
  lds_double* w = (lds_double*)win;
  switch (s) {
    case 0: edge_rhs_tile<GAUGE, CO, 1, 0>(in, aux, box, T.t[0], lb, C, k, flag, w); break;
    case 1: edge_rhs_tile<GAUGE, CO, 1, 1>(in, aux, box, T.t[1], lb, C, k, flag, w); break;
    case 2: edge_rhs_tile<GAUGE, CO, 1, 2>(in, aux, box, T.t[2], lb, C, k, flag, w); break;
    case 3: edge_rhs_tile<GAUGE, CO, 2, 0>(in, aux, box, T.t[3], lb, C, k, flag, w); break;
    case 4: edge_rhs_tile<GAUGE, CO, 2, 1>(in, aux, box, T.t[4], lb, C, k, flag, w); break;
    default: edge_rhs_tile<GAUGE, CO, 2, 2>(in, aux, box, T.t[5], lb, C, k, flag, w); break;
  }
}

// The tile of block b of the set.
__device__ __forceinline__ EdgeTile edge_set_tile(const EdgeTileSet& T, int b) {
  const int s = T.set_of(b), lb = b - T.first[s];
  switch (s) {
    case 0: return edge_tile(T.t[0], lb);
    case 1: return edge_tile(T.t[1], lb);
    case 2: return edge_tile(T.t[2], lb);
    case 3: return edge_tile(T.t[3], lb);
    case 4: return edge_tile(T.t[4], lb);
    default: return edge_tile(T.t[5], lb);
  }
}

// rungekutta4_rout (rungekutta4_rout.f90:110-128), the chi floor and, before
// sub-step 3, enforce_ga on the points of the tiles, from their compact
// right-hand sides: the epilogue of the fused interior kernel, per point and
// field.  blockIdx.y is the point's work item (UpdateItems): the fields a
// thread updates alone cost it one memory round trip, and the points of an
// x-face, one cache line per point and field, get many threads in flight.
template <int STAGE>
__global__ void __launch_bounds__(kEdgeThreads) edge_update_kernel(FieldsIn u0, FieldsOut acc, FieldsOut out,
                                                                   AuxOut aux, Box box, EdgeTileSet T,
                                                                   const double* K, Rk4Param R) {
  const EdgeTile t = edge_set_tile(T, (int)blockIdx.x);
  int q[3];
  bool live;
  edge_tile_point(t, (int)threadIdx.x, q, live);
  if (!live) return;
  const long p = q[0] + q[1] * box.sy + q[2] * box.sz;
  const double* k = K + (long)blockIdx.x * NFIELD * kEdgeThreads + threadIdx.x;
  Rk4SinkT<STAGE> S{u0, acc, out, aux, p, R.c, R.tinny, {}};
  const int y = (int)blockIdx.y;
  if (STAGE < 3 && y == 0) {
    // the 12 fields enforce_ga couples, every load first
    double kv[12], uv[12], av[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) {
      kv[s] = k[(DXX + s) * kEdgeThreads];
      uv[s] = u0.p[DXX + s][p];
      av[s] = STAGE > 0 ? acc.p[DXX + s][p] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 12; ++s) S.put_loaded(DXX + s, kv[s], uv[s], av[s]);
    S.finish();
  } else {
    const int v = UpdateItems<STAGE>::field(y);
    S.put_loaded(v, k[v * kEdgeThreads], u0.p[v][p], STAGE > 0 ? acc.p[v][p] : 0.0);
  }
}

template <int G, bool CO>
void launch_edge_lds_rhs(const FieldsIn& in, const AuxOut& aux, const Box& box, const EdgeTileSet& T,
                         const Coef& C, double* K, int* flag, hipStream_t s) {
  rhs_edge_lds_kernel<G, CO><<<(unsigned)T.tiles(), kEdgeThreads, 0, s>>>(in, aux, box, T, C, K, flag);
}

}  // namespace amss
