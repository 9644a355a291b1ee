// bssn_kernels_impl.h -- kernel templates of the BSSN RHS + RK4 path.
// Instantiated once per gauge in bssn_gauge0.hip / bssn_gauge2.hip so the
// heavy point kernels compile in parallel translation units.
#pragma once
#include "bssn_batch.h"
#include "bssn_edge_tiles.h"
#include "bssn_lds.h"

namespace amss {

#define AMSS_POINT_KERNEL __global__ void __launch_bounds__(kRhsThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))

template <class FI>
__device__ __forceinline__ bool bad_input(const FI& F, long p) {
  // the reference's NaN test covers the first 21 fields (bssn_rhs.f90:110-113)
  bool b = false;
#pragma unroll
  for (int v = 0; v < 21; ++v) {
    const double x = F.ptr(v)[p];
    b |= !(x == x);
  }
  return b;
}

// The frozen-ghost fill of point p after an RK4 sub-step: the sub-step's
// base state U0, the NaN test on it, enforce_ga the next RHS applies to its
// input (sub-steps 0-2) and the lower bound on chi -- the sequence rk4 ->
// copy -> lowerboundset -> enforce_ga the reference runs on such points
// (bssn_class.C:3119-3156, 3415-3525).
__device__ __forceinline__ void ghost_fill_point(const FieldsIn& u0, const FieldsOut& out, long p, int stage,
                                                 double tinny, int* flag) {
  double u[NFIELD];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) u[v] = u0.p[v][p];
  bool bad = false;
#pragma unroll
  for (int v = 0; v < 21; ++v) bad |= !(u[v] == u[v]);  // bssn_rhs.f90:110-113
  if (bad) *flag = 1;
  if (stage < 3) enforce_point(u + DXX, u + AXX);
  u[CHI] = u[CHI] < tinny ? tinny : u[CHI];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) out.p[v][p] = u[v];
}

__device__ __forceinline__ bool interior_index(const Box& box, int& i, int& j, int& k) {
  i = box.lo[0] + blockIdx.x * kTileX + (threadIdx.x % kTileX);
  j = box.lo[1] + blockIdx.y * kTileY + (threadIdx.x / kTileX);
  k = box.lo[2] + blockIdx.z;
  return i < box.hi[0] && j < box.hi[1];
}
inline EdgeShells edge_shells(const Box& b) { return edge_shells_of(b.n, b.lo, b.hi, b.mlo); }

// Every point of the shell lies on the outermost layer of a face in mask
// (the on_faces bits): each of its slabs is one point deep along its axis,
// at index 0 of a lower face or n - 1 of an upper one in mask.
inline bool shell_on_faces(const Shell& sh, const int n[3], int mask) {
  // slab s: axis 2 - s / 2 (z, z, y, y, x, x), lower slab for even s
  for (int s = 0; s < 6; ++s) {
    if (sh.cnt[s] == 0) continue;
    const int a = 2 - s / 2;
    const int lo = (s & 1) ? sh.IH[a] : sh.OL[a], hi = (s & 1) ? sh.OH[a] : sh.IL[a];
    if (hi - lo != 1) return false;
    const bool on = (s & 1) ? (lo == n[a] - 1 && ((mask >> a) & 1)) : (lo == 0 && ((mask >> (3 + a)) & 1));
    if (!on) return false;
  }
  return true;
}

// Whether the LDS boundary kernels may take the order-1 and order-2 shells
// of this box: an even row pitch (the direct loads move 16 bytes from even
// indices), no Sommerfeld or copy face on them (a physical lower face has
// no mirrored halo, so its index 0 is an outermost-layer point; upper faces
// are at n - 1), and room for their tiles in the right-hand-side buffer.
inline bool edge_lds_ok(const Box& box, const SommerGeom& sg, const EdgeTileSet& T, const Rk4Param& R) {
  if (!R.edge_k || (box.sy % 2) != 0 || (box.sz % 2) != 0 || T.tiles() > R.edge_k_tiles) return false;
  const int mask = sg.faces | sg.copy;
  for (int a = 0; a < 3; ++a)
    if (((mask >> (3 + a)) & 1) && box.mlo[a] != 0) return false;
  return true;
}

// The right-hand side on the tiles of the order-1 and order-2 shells
// (bssn_edge_lds.h; instantiated per gauge and residual flag in
// bssn_edge_g<G>[_res].hip), and the RK4 update of their points.
template <int G, bool CO>
void launch_edge_lds_rhs(const FieldsIn& in, const AuxOut& aux, const Box& box, const EdgeTileSet& T,
                         const Coef& C, double* K, int* flag, hipStream_t s);
void launch_edge_lds_update(const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out, const AuxOut& aux,
                            const Box& box, const EdgeTileSet& T, const double* K, const Rk4Param& R,
                            hipStream_t s);

__device__ __forceinline__ bool edge_index(const Shell& sh, int& i, int& j, int& k) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= sh.size()) return false;
  sh.point(t, i, j, k);
  return true;
}

// ---- plain RHS (the compute_rhs_bssn outputs) ------------------------------
template <int GAUGE, bool CO>
AMSS_POINT_KERNEL rhs_interior_kernel(FieldsIn in, FieldsIn mat, FieldsOut out, AuxOut aux, Box box, Coef C,
                                      int* flag) {
  int i, j, k;
  if (!interior_index(box, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  if (bad_input(in, p)) *flag = 1;
  InteriorEnv E{in, mat, C, p, box.sy, box.sz};
  RhsSink S{out, aux, p};
  bssn_rhs_point<GAUGE, CO, true, true>(E, S);
}

template <int GAUGE, bool CO, int ORD>
__global__ void __launch_bounds__(256) rhs_edge_kernel(FieldsIn in, FieldsIn mat, FieldsOut out, AuxOut aux,
                                                       Box box, Shell sh, Coef C, int* flag) {
  int i, j, k;
  if (!edge_index(sh, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  if (bad_input(in, p)) *flag = 1;
  EdgeEnv<ORD> E(in, mat, C, i, j, k, box.n, box.mlo, box.sy, box.sz);
  RhsSink S{out, aux, p};
  bssn_rhs_point<GAUGE, CO, true, true>(E, S);
}

// ---- RHS fused with the RK4 sub-step ----------------------------------------
template <int GAUGE, int STAGE, bool CO>
AMSS_POINT_KERNEL rk4_interior_kernel(FieldsIn in, FieldsIn u0, FieldsOut acc, FieldsOut out, AuxOut aux, Box box,
                                      Coef C, Rk4Param R, int* flag) {
  int i, j, k;
  if (!interior_index(box, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  if (bad_input(in, p)) *flag = 1;
  InteriorEnv E{in, in, C, p, box.sy, box.sz};
  Rk4SinkT<STAGE> S{u0, acc, out, aux, p, R.c, R.tinny, {}};
  bssn_rhs_point<GAUGE, CO, false, false>(E, S);
  S.finish();
}

// LDS-staged variant: 8x8xTZ-point tiles (TZ = 8 where the box holds whole
// 8-deep tiles, else 4), halo tiles of every field group in LDS
// (double-buffered, filled by direct global->LDS loads).  A 1-D grid of
// tiles in XCD-contiguous order: consecutive tiles, whose halos overlap, run
// behind the same L2.  The 8-deep tile (512 threads, one workgroup per CU)
// re-reads 14/8 instead of 10/4 z-planes of halo per interior plane.
constexpr int kLdsTX = 8, kLdsTY = 8, kLdsTZ = 4;
#ifndef AMSS_LDS_WAVES
#define AMSS_LDS_WAVES 2
#endif
template <int GAUGE, int STAGE, bool CO, int TZ>
__global__ void __launch_bounds__(kLdsTX * kLdsTY * TZ) __attribute__((amdgpu_waves_per_eu(AMSS_LDS_WAVES, AMSS_LDS_WAVES)))
rk4_lds_kernel(FieldsIn in, FieldsIn u0, FieldsOut acc, FieldsOut out, AuxOut aux, Box box,
                                 Coef C, Rk4Param R, int* flag, long nfill) {
  using Tl = LdsTile<kLdsTX, kLdsTY, TZ>;
  __shared__ __attribute__((aligned(16))) double tiles[kLdsBuffers * Tl::NPAD];
  const int gx = (box.hi[0] - box.lo[0]) / kLdsTX, gy = (box.hi[1] - box.lo[1]) / kLdsTY;
  const int ntiles = gx * gy * ((box.hi[2] - box.lo[2]) / TZ);
  if ((int)blockIdx.x >= ntiles) {
    // the frozen-ghost fill of the nfill points outside [3, n - 3) of the
    // array (nfill is 0 when the caller fills them), in the workgroups after
    // the tiles (they start as the last tiles finish); its points lie
    // outside every tile
    if constexpr (STAGE < 2) {  // rk4_fuses_fill: sub-steps 0 and 1
      const long t = (long)(blockIdx.x - ntiles) * blockDim.x + threadIdx.x;
      if (t < nfill) {
        int i, j, k;
        ghost3_point(t, box.n[0], box.n[1], box.n[2], i, j, k);
        ghost_fill_point(u0, out, i + j * box.sy + k * box.sz, STAGE, R.tinny, flag);
      }
    }
    return;
  }
  const int tx = threadIdx.x % kLdsTX, ty = tile_row((threadIdx.x / kLdsTX) % kLdsTY, kLdsTY),
            tz = threadIdx.x / (kLdsTX * kLdsTY);
  const int b = xcd_spread(blockIdx.x, ntiles);
  const int i0 = box.lo[0] + (b % gx) * kLdsTX, j0 = box.lo[1] + ((b / gx) % gy) * kLdsTY;
  const int k0 = box.lo[2] + (b / (gx * gy)) * TZ;
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  // the launch covers whole tiles only (remainder slabs go to the direct kernel)
  const long p = i + j * box.sy + k * box.sz;
  const int N = box.n[0] * box.n[1] * box.n[2];
  lds_double* t0 = (lds_double*)tiles;
  LdsRk4<STAGE, kLdsTX, kLdsTY, TZ, FieldsIn, FieldsOut, AuxOut, true> X{
      {in, in, C, t0, t0 + (kLdsBuffers / 2) * Tl::NPAD, p, (tx + 3) + Tl::LX * ((ty + 3) + Tl::LY * (tz + 3)),
       (int)((i0 - 3) + (j0 - 3) * box.sy + (k0 - 3) * box.sz), (int)box.sy, (int)box.sz, N, (int)threadIdx.x},
      {u0, acc, out, aux, p, R.c, R.tinny, {}}};
  X.prologue();
  bssn_rhs_point<GAUGE, CO, false, false>(X, X);
  X.finish();
  if (X.bad) *flag = 1;
}

// sg: the Sommerfeld faces of the array (none when sg.faces == 0).  The
// sub-step is a template argument, as in the interior kernels (a runtime
// stage costs registers and spills).
#ifndef AMSS_EDGE_WAVES
#define AMSS_EDGE_WAVES 2
#endif
template <int GAUGE, int STAGE, bool CO, int ORD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AMSS_EDGE_WAVES, AMSS_EDGE_WAVES)))
rk4_edge_kernel(FieldsIn in, FieldsIn u0, FieldsOut acc, FieldsOut out, AuxOut aux, Box box, Shell sh, Coef C,
                Rk4Param R, int* flag, SommerGeom sg) {
  int i, j, k;
  if (!edge_index(sh, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  if (bad_input(in, p)) *flag = 1;
  // without the residuals, the points of the copy faces need nothing and
  // those of the Sommerfeld faces take sommer_face_kernel
  if (ORD == 0 && !CO && on_faces(sg.copy | sg.faces, box.n, i, j, k)) return;
  EdgeEnv<ORD> E(in, in, C, i, j, k, box.n, box.mlo, box.sy, box.sz);
  SommerRk4Sink<EdgeEnv<ORD>, Rk4SinkT<STAGE>> S{
      {u0, acc, out, aux, p, R.c, R.tinny, {}}, E, sg, sommer_point(sg, box.n, i, j, k)};
  bssn_rhs_point<GAUGE, CO, false, false>(E, S);
  S.finish();
}

// The RK4 sub-step on the Sommerfeld-face points of the outermost layer
// when no residuals are wanted: every field's RHS there is the boundary
// condition's (sommerfeld_rout.f90:376-440, applied to the stage's RHS
// before rungekutta4_rout, bssn_class.C:3106-3117), so the point algebra is
// skipped.  blockIdx.y is the point's work item (UpdateItems): one field
// alone, or before sub-step 3 the 12 fields enforce_ga couples, whose
// boundary RHS are all formed and whose U0 and accumulator values are all
// loaded before the first store.  With test_nan the kernel also applies the
// reference's NaN test to the shell's points (the launcher then runs no
// other kernel on them).
template <int STAGE>
__global__ void __launch_bounds__(256) sommer_face_kernel(FieldsIn in, FieldsIn u0, FieldsOut acc, FieldsOut out,
                                                          AuxOut aux, Box box, Shell sh, Rk4Param R, SommerGeom sg,
                                                          int* flag, bool test_nan) {
  int i, j, k;
  if (!edge_index(sh, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  const int y = (int)blockIdx.y;
  if (y == 0 && test_nan && bad_input(in, p)) *flag = 1;
  if (on_faces(sg.copy, box.n, i, j, k) || !on_faces(sg.faces, box.n, i, j, k)) return;
  const int q[3] = {i, j, k};
  // the boundary RHS of field v; index -1 is index 0 times the field's
  // parity (the one mirrored layer below a symmetry plane), taken without a
  // branch so that all the reads can be in flight together
  auto rhs = [&](int v) {
    auto rd = [&](int a, int m) {
      int c[3] = {q[0], q[1], q[2]};
      const int t = c[a] + m, neg = t >> 31;
      c[a] = t ^ neg;
      const double x = in.p[v][c[0] + c[1] * box.sy + c[2] * box.sz];
      return (x + 0.0) * (neg ? (double)parity(v, a) : 1.0);
    };
    return sommer_rhs(sg, box.n, q, prop_speed(v), rd);
  };
  Rk4SinkT<STAGE> S{u0, acc, out, aux, p, R.c, R.tinny, {}};
  if (STAGE < 3 && y == 0) {
    double kv[12], uv[12], av[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) kv[s] = rhs(DXX + s);
#pragma unroll
    for (int s = 0; s < 12; ++s) {
      uv[s] = u0.p[DXX + s][p];
      av[s] = STAGE > 0 ? acc.p[DXX + s][p] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < 12; ++s) S.put_loaded(DXX + s, kv[s], uv[s], av[s]);
    S.finish();
  } else {
    const int v = UpdateItems<STAGE>::field(y);
    const double kv = rhs(v), uv = u0.p[v][p], av = STAGE > 0 ? acc.p[v][p] : 0.0;
    S.put_loaded(v, kv, uv, av);
  }
}

inline dim3 interior_grid(const Box& b) {
  return dim3((b.hi[0] - b.lo[0] + kTileX - 1) / kTileX, (b.hi[1] - b.lo[1] + kTileY - 1) / kTileY,
              b.hi[2] - b.lo[2]);
}
inline bool box_empty(const Box& b) { return b.hi[0] <= b.lo[0] || b.hi[1] <= b.lo[1] || b.hi[2] <= b.lo[2]; }
inline unsigned blocks_for(long n, int t) { return (unsigned)((n + t - 1) / t); }

template <int G>
void launch_rhs_gauge(const FieldsIn& in, const FieldsIn& mat, const FieldsOut& out, const AuxOut& aux,
                      const Box& box, const Coef& C, int co, int* flag, bool edge, hipStream_t s) {
  if (co == 0) {
    if (!box_empty(box)) rhs_interior_kernel<G, true><<<interior_grid(box), kRhsThreads, 0, s>>>(in, mat, out, aux, box, C, flag);
  } else {
    if (!box_empty(box)) rhs_interior_kernel<G, false><<<interior_grid(box), kRhsThreads, 0, s>>>(in, mat, out, aux, box, C, flag);
  }
  if (!edge) return;
  const EdgeShells es = edge_shells(box);
#define AMSS_REDGE(O)                                                                                       \
  if (es.s[O].size() > 0) {                                                                                 \
    const unsigned gr = blocks_for(es.s[O].size(), 256);                                                    \
    if (co == 0) rhs_edge_kernel<G, true, O><<<gr, 256, 0, s>>>(in, mat, out, aux, box, es.s[O], C, flag);  \
    else rhs_edge_kernel<G, false, O><<<gr, 256, 0, s>>>(in, mat, out, aux, box, es.s[O], C, flag);         \
  }
  AMSS_REDGE(0)
  AMSS_REDGE(1)
  AMSS_REDGE(2)
#undef AMSS_REDGE
}

template <int G>
void launch_rk4_gauge(const FieldsIn& in, const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                      const AuxOut& aux, const Box& box, const Coef& C, const Rk4Param& R, bool want_res, int* flag,
                      bool edge, hipStream_t s, bool interior, const SommerGeom& sg) {
  // The LDS kernel's 16-byte direct loads start at even x of a row, which is
  // 16-byte aligned only when the row pitch is even; odd pitches go direct.
  if (interior && !box_empty(box) && R.variant == 1 && (box.sy % 2) == 0) {
    // LDS kernel on the largest tile-aligned sub-box, the direct kernel on the rest
    Box core = box;
    const int tz = box.hi[2] - box.lo[2] >= 8 ? 8 : kLdsTZ;
    const int T[3] = {kLdsTX, kLdsTY, tz};
    for (int a = 0; a < 3; ++a) core.hi[a] = box.lo[a] + (box.hi[a] - box.lo[a]) / T[a] * T[a];
    if (!box_empty(core)) {
      const unsigned ntile = (unsigned)((core.hi[0] - core.lo[0]) / kLdsTX) * ((core.hi[1] - core.lo[1]) / kLdsTY) *
                             ((core.hi[2] - core.lo[2]) / tz);
      const int nt = kLdsTX * kLdsTY * tz;
      // the fused frozen-ghost fill (none unless rk4_fuses_fill)
      const long nfill = rk4_fuses_fill(box, R) ? ghost3_count(box.n[0], box.n[1], box.n[2]) : 0;
      const unsigned gr = ntile + (unsigned)((nfill + nt - 1) / nt);
#define AMSS_LDS(TZ)                                                                                          \
  switch (R.stage) {                                                                                          \
    case 0:                                                                                                   \
      if (want_res) rk4_lds_kernel<G, 0, true, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag, nfill); \
      else rk4_lds_kernel<G, 0, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag, nfill);         \
      break;                                                                                                  \
    case 1: rk4_lds_kernel<G, 1, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag, nfill); break; \
    case 2: rk4_lds_kernel<G, 2, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag, nfill); break; \
    default: rk4_lds_kernel<G, 3, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag, nfill); break; \
  }
      if (tz == 8) {
        AMSS_LDS(8)
      } else {
        AMSS_LDS(4)
      }
#undef AMSS_LDS
    }
    Rk4Param Rd = R;
    Rd.variant = 0;
    for (int a = 0; a < 3; ++a) {  // remainder slab [core.hi[a], box.hi[a]) along axis a
      Box rem = box;
      for (int c = 0; c < a; ++c) rem.hi[c] = core.hi[c];
      rem.lo[a] = core.hi[a];
      if (!box_empty(rem)) launch_rk4_gauge<G>(in, u0, acc, out, aux, rem, C, Rd, want_res, flag, false, s, true, sg);
    }
  } else if (interior && !box_empty(box)) {
    const dim3 gr = interior_grid(box);  // also variant 1 on an odd row pitch
    switch (R.stage) {
      case 0:
        if (want_res) rk4_interior_kernel<G, 0, true><<<gr, kRhsThreads, 0, s>>>(in, u0, acc, out, aux, box, C, R, flag);
        else rk4_interior_kernel<G, 0, false><<<gr, kRhsThreads, 0, s>>>(in, u0, acc, out, aux, box, C, R, flag);
        break;
      case 1: rk4_interior_kernel<G, 1, false><<<gr, kRhsThreads, 0, s>>>(in, u0, acc, out, aux, box, C, R, flag); break;
      case 2: rk4_interior_kernel<G, 2, false><<<gr, kRhsThreads, 0, s>>>(in, u0, acc, out, aux, box, C, R, flag); break;
      default: rk4_interior_kernel<G, 3, false><<<gr, kRhsThreads, 0, s>>>(in, u0, acc, out, aux, box, C, R, flag); break;
    }
  }
  if (edge) {
    const EdgeShells es = edge_shells(box);
    // the order-1 and order-2 shells in one LDS launch pair where it may
    // take them (run with edge_order 1; edge_order 2 then has nothing left)
    const EdgeTileSet ts = edge_tile_set(es);
    const bool lds = R.variant == 1 && edge_lds_ok(box, sg, ts, R);
    for (int o = 0; o < 3; ++o) {
      const Shell& sh = es.s[o];
      if (sh.size() <= 0 || (R.edge_order >= 0 && o != R.edge_order)) continue;
      if (lds && o >= 1) {
        if (o == 1 && ts.tiles() > 0) {
          if (want_res) launch_edge_lds_rhs<G, true>(in, aux, box, ts, C, R.edge_k, flag, s);
          else launch_edge_lds_rhs<G, false>(in, aux, box, ts, C, R.edge_k, flag, s);
          launch_edge_lds_update(u0, acc, out, aux, box, ts, R.edge_k, R, s);
        }
        continue;
      }
      const unsigned gr = blocks_for(sh.size(), 256);
      // Without the residuals, an outermost layer made of Sommerfeld and
      // copy faces only needs the face kernel.
      const bool faces_only = o == 0 && !want_res && (sg.faces | sg.copy) != 0 &&
                              shell_on_faces(sh, box.n, sg.faces | sg.copy);
      if (!faces_only) {
#define AMSS_EDGE(ST, CO)                                                                                      \
  do {                                                                                                         \
    if (o == 0) rk4_edge_kernel<G, ST, CO, 0><<<gr, 256, 0, s>>>(in, u0, acc, out, aux, box, sh, C, R, flag, sg); \
    else if (o == 1) rk4_edge_kernel<G, ST, CO, 1><<<gr, 256, 0, s>>>(in, u0, acc, out, aux, box, sh, C, R, flag, sg); \
    else rk4_edge_kernel<G, ST, CO, 2><<<gr, 256, 0, s>>>(in, u0, acc, out, aux, box, sh, C, R, flag, sg);       \
  } while (0)
        switch (R.stage) {
          case 0:
            if (want_res) AMSS_EDGE(0, true);
            else AMSS_EDGE(0, false);
            break;
          case 1: AMSS_EDGE(1, false); break;
          case 2: AMSS_EDGE(2, false); break;
          default: AMSS_EDGE(3, false); break;
        }
#undef AMSS_EDGE
      }
      if (o == 0 && !want_res && (faces_only || sg.faces != 0)) {
        const bool nan = faces_only;
#define AMSS_FACE(ST)                                                                                      \
  sommer_face_kernel<ST><<<dim3(gr, UpdateItems<ST>::kCount), 256, 0, s>>>(in, u0, acc, out, aux, box, sh, R, sg, \
                                                                          flag, nan)
        switch (R.stage) {
          case 0: AMSS_FACE(0); break;
          case 1: AMSS_FACE(1); break;
          case 2: AMSS_FACE(2); break;
          default: AMSS_FACE(3); break;
        }
#undef AMSS_FACE
      }
    }
  }
}

}  // namespace amss
