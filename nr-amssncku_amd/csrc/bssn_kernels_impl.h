// bssn_kernels_impl.h -- kernel templates of the BSSN RHS + RK4 path.
// Instantiated once per gauge in bssn_gauge0.hip / bssn_gauge2.hip so the
// heavy point kernels compile in parallel translation units.
#pragma once
#include "bssn_batch.h"
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

__device__ __forceinline__ bool interior_index(const Box& box, int& i, int& j, int& k) {
  i = box.lo[0] + blockIdx.x * kTileX + (threadIdx.x % kTileX);
  j = box.lo[1] + blockIdx.y * kTileY + (threadIdx.x / kTileX);
  k = box.lo[2] + blockIdx.z;
  return i < box.hi[0] && j < box.hi[1];
}
// The points of an outer box [OL, OH) outside an inner box [IL, IH) (clamped
// into it), enumerated as six disjoint slabs (z faces whole, then y faces
// over the inner z range, then x faces over the inner y and z ranges), so
// that every lane of a wave has a point: walking the whole array instead
// leaves an x-face row's points alone in a wave.  An inner box empty along
// any axis holds no point, and the slabs then cover the whole outer box.
struct Shell {
  int OL[3], OH[3], IL[3], IH[3];
  long cnt[6];
  __host__ __device__ Shell(const int ol[3], const int oh[3], const int il[3], const int ih[3]) {
    bool empty = false;
    for (int a = 0; a < 3; ++a) {
      OL[a] = ol[a];
      OH[a] = oh[a] < ol[a] ? ol[a] : oh[a];
      IL[a] = il[a] < OL[a] ? OL[a] : (il[a] > OH[a] ? OH[a] : il[a]);
      IH[a] = ih[a] < IL[a] ? IL[a] : (ih[a] > OH[a] ? OH[a] : ih[a]);
      empty = empty || IH[a] == IL[a];
    }
    if (empty)
      for (int a = 0; a < 3; ++a) IH[a] = IL[a];
    const long ex = OH[0] - OL[0], ey = OH[1] - OL[1];
    const long mk = IH[2] - IL[2], mj = IH[1] - IL[1];
    cnt[0] = ex * ey * (IL[2] - OL[2]);
    cnt[1] = ex * ey * (OH[2] - IH[2]);
    cnt[2] = ex * (IL[1] - OL[1]) * mk;
    cnt[3] = ex * (OH[1] - IH[1]) * mk;
    cnt[4] = (long)(IL[0] - OL[0]) * mj * mk;
    cnt[5] = (long)(OH[0] - IH[0]) * mj * mk;
  }
  __host__ __device__ long size() const { return cnt[0] + cnt[1] + cnt[2] + cnt[3] + cnt[4] + cnt[5]; }
  // point t of the shell (t < size())
  __device__ void point(long t, int& i, int& j, int& k) const {
    int s = 0;
    while (s < 5 && t >= cnt[s]) t -= cnt[s++];
    const int x0 = s < 4 ? OL[0] : (s == 4 ? OL[0] : IH[0]);
    const int ex = s < 4 ? OH[0] - OL[0] : (s == 4 ? IL[0] - OL[0] : OH[0] - IH[0]);
    const int y0 = s < 2 ? OL[1] : (s == 2 ? OL[1] : (s == 3 ? IH[1] : IL[1]));
    const int ey = s < 2 ? OH[1] - OL[1] : (s == 2 ? IL[1] - OL[1] : (s == 3 ? OH[1] - IH[1] : IH[1] - IL[1]));
    const int z0 = s == 0 ? OL[2] : (s == 1 ? IH[2] : IL[2]);
    i = x0 + (int)(t % ex);
    j = y0 + (int)((t / ex) % ey);
    k = z0 + (int)(t / ((long)ex * ey));
  }
};

// The boundary points of an array outside the interior box, split by the
// order of their centred stencils (the joint three-axis room test of
// fderivs/fdderivs, diff_new.f90:996-1100, with the lowest readable index
// mlo): order 2 on [max(0, mlo + 2), n - 2), order 1 on the next layer out to
// [max(0, mlo + 1), n - 1), order 0 (and the last index) on the rest.
struct EdgeShells {
  Shell s[3];  // by stencil order 0, 1, 2
};
inline EdgeShells edge_shells(const Box& b) {
  int z[3] = {0, 0, 0}, L1[3], H1[3], L2[3], H2[3];
  for (int a = 0; a < 3; ++a) {
    L1[a] = b.mlo[a] + 1 > 0 ? b.mlo[a] + 1 : 0;
    H1[a] = b.n[a] - 1;
    L2[a] = b.mlo[a] + 2 > 0 ? b.mlo[a] + 2 : 0;
    H2[a] = b.n[a] - 2;
    if (H2[a] < L2[a]) H2[a] = L2[a];
    if (H1[a] < L1[a]) H1[a] = L1[a];
  }
  int blo[3], bhi[3];
  bool empty = false;
  for (int a = 0; a < 3; ++a) {
    blo[a] = b.lo[a];
    bhi[a] = b.hi[a];
    empty = empty || b.hi[a] <= b.lo[a];
  }
  if (empty)  // no interior launch: the order-2 shell takes its box whole
    for (int a = 0; a < 3; ++a) bhi[a] = blo[a] = L2[a];
  return EdgeShells{{Shell(z, b.n, L1, H1), Shell(L1, H1, L2, H2), Shell(L2, H2, blo, bhi)}};
}

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
                                 Coef C, Rk4Param R, int* flag) {
  using Tl = LdsTile<kLdsTX, kLdsTY, TZ>;
  __shared__ __attribute__((aligned(16))) double tiles[kLdsBuffers * Tl::NPAD];
  const int tx = threadIdx.x % kLdsTX, ty = tile_row((threadIdx.x / kLdsTX) % kLdsTY, kLdsTY),
            tz = threadIdx.x / (kLdsTX * kLdsTY);
  const int gx = (box.hi[0] - box.lo[0]) / kLdsTX, gy = (box.hi[1] - box.lo[1]) / kLdsTY;
  const int b = xcd_spread(blockIdx.x, gridDim.x);
  const int i0 = box.lo[0] + (b % gx) * kLdsTX, j0 = box.lo[1] + ((b / gx) % gy) * kLdsTY;
  const int k0 = box.lo[2] + (b / (gx * gy)) * TZ;
  const int i = i0 + tx, j = j0 + ty, k = k0 + tz;
  // the launch covers whole tiles only (remainder slabs go to the direct kernel)
  const long p = i + j * box.sy + k * box.sz;
  const int N = box.n[0] * box.n[1] * box.n[2];
  lds_double* t0 = (lds_double*)tiles;
  LdsRk4<STAGE, kLdsTX, kLdsTY, TZ> X{
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
  if (ORD == 0 && !CO && on_faces(sg.copy, box.n, i, j, k)) return;
  EdgeEnv<ORD> E(in, in, C, i, j, k, box.n, box.mlo, box.sy, box.sz);
  SommerRk4Sink<EdgeEnv<ORD>, Rk4SinkT<STAGE>> S{
      {u0, acc, out, aux, p, R.c, R.tinny, {}}, E, sg, sommer_point(sg, box.n, i, j, k)};
  // On a Sommerfeld face every field's RHS is replaced before the update, so
  // unless the residuals are wanted the RHS algebra there is skipped (face
  // points lie on the outermost layer: stencil order 0).
  if (ORD == 0 && !CO && S.on) {
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) S.put(v, 0.0);
  } else {
    bssn_rhs_point<GAUGE, CO, false, false>(E, S);
  }
  S.finish();
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
      const unsigned gr = (unsigned)((core.hi[0] - core.lo[0]) / kLdsTX) * ((core.hi[1] - core.lo[1]) / kLdsTY) *
                          ((core.hi[2] - core.lo[2]) / tz);
      const int nt = kLdsTX * kLdsTY * tz;
#define AMSS_LDS(TZ)                                                                                          \
  switch (R.stage) {                                                                                          \
    case 0:                                                                                                   \
      if (want_res) rk4_lds_kernel<G, 0, true, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag); \
      else rk4_lds_kernel<G, 0, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag);         \
      break;                                                                                                  \
    case 1: rk4_lds_kernel<G, 1, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag); break; \
    case 2: rk4_lds_kernel<G, 2, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag); break; \
    default: rk4_lds_kernel<G, 3, false, TZ><<<gr, nt, 0, s>>>(in, u0, acc, out, aux, core, C, R, flag); break; \
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
    for (int o = 0; o < 3; ++o) {
      const Shell& sh = es.s[o];
      if (sh.size() <= 0 || (R.edge_order >= 0 && o != R.edge_order)) continue;
      const unsigned gr = blocks_for(sh.size(), 256);
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
  }
}

}  // namespace amss
