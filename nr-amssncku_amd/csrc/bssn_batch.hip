// bssn_batch.hip -- box-table kernels of the batched level step (ghost fill,
// inter-patch exchange, enforce_ga over all patches) and the dispatch of the
// tile-list kernel by gauge and tile shape.
#include "bssn_batch.h"

namespace amss {

namespace {

template <class FI>
__device__ __forceinline__ bool bad_point(const FI& F, long p) {
  // the reference's NaN test covers the first 21 fields (bssn_rhs.f90:110-113)
  bool b = false;
#pragma unroll
  for (int v = 0; v < 21; ++v) {
    const double x = F.ptr(v)[p];
    b |= !(x == x);
  }
  return b;
}

// Boundary points after a sub-step: the state's values (frozen ghost zones or
// the one-layer copy of sommerfeld_rout precor = 1), with chi bounded below
// and, for sub-steps 0-2, enforce_ga applied -- the same per-point work as the
// per-patch ghost fill.  blockIdx.y selects the box.
__global__ void __launch_bounds__(256) fill_boxes_kernel(const PatchDesc* __restrict__ P,
                                                         const FillBox* __restrict__ B, Roles m, int stage,
                                                         double tinny, int* flag) {
  const FillBox b = B[blockIdx.y];
  const PatchDesc d = P[b.patch];
  const long ln = (long)NFIELD * d.n;
  const SlabFields u0{(gdouble*)d.slab + m.u0 * ln, d.n}, out{(gdouble*)d.slab + m.out * ln, d.n};
  const long npts = (long)b.ext[0] * b.ext[1] * b.ext[2];
  for (long t = blockIdx.x * 256L + threadIdx.x; t < npts; t += gridDim.x * 256L) {
    const int i = b.lo[0] + (int)(t % b.ext[0]);
    const int j = b.lo[1] + (int)((t / b.ext[0]) % b.ext[1]);
    const int k = b.lo[2] + (int)(t / ((long)b.ext[0] * b.ext[1]));
    const long p = i + (long)j * d.sy + (long)k * d.sz;
    if (bad_point(u0, p)) *flag = 1;
    double ga[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) ga[s] = u0.ptr(DXX + s)[p];
    if (stage < 3) enforce_point(ga, ga + 6);
#pragma unroll
    for (int s = 0; s < 12; ++s) out.ptr(DXX + s)[p] = ga[s];
    const double c = u0.ptr(CHI)[p];
    out.ptr(CHI)[p] = c < tinny ? tinny : c;
    out.ptr(TRK)[p] = u0.ptr(TRK)[p];
#pragma unroll
    for (int v = GAMX; v < NFIELD; ++v) out.ptr(v)[p] = u0.ptr(v)[p];
  }
}

// Ghost exchange between patches: all 24 fields of list slot `slot`,
// source interior box -> destination ghost box.  blockIdx.y selects the box.
__global__ void __launch_bounds__(256) seg_boxes_kernel(const PatchDesc* __restrict__ P,
                                                        const SegRef* __restrict__ S, int slot) {
  const SegRef g = S[blockIdx.y];
  const PatchDesc a = P[g.src], b = P[g.dst];
  const gdouble* src = (gdouble*)a.slab + (long)slot * NFIELD * a.n;
  gdouble* dst = (gdouble*)b.slab + (long)slot * NFIELD * b.n;
  const long npts = (long)g.ext[0] * g.ext[1] * g.ext[2];
  for (long t = blockIdx.x * 256L + threadIdx.x; t < npts; t += gridDim.x * 256L) {
    const int x = (int)(t % g.ext[0]), y = (int)((t / g.ext[0]) % g.ext[1]);
    const int z = (int)(t / ((long)g.ext[0] * g.ext[1]));
    const long ps = (g.slo[0] + x) + (long)(g.slo[1] + y) * a.sy + (long)(g.slo[2] + z) * a.sz;
    const long pd = (g.dlo[0] + x) + (long)(g.dlo[1] + y) * b.sy + (long)(g.dlo[2] + z) * b.sz;
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) dst[v * b.n + pd] = src[v * a.n + ps];
  }
}

__global__ void __launch_bounds__(256) seg_boxes_adjoint_kernel(const PatchDesc* __restrict__ P,
                                                                const SegRef* __restrict__ S) {
  const SegRef g = S[blockIdx.y];
  const PatchDesc a = P[g.src], b = P[g.dst];
  const long npts = (long)g.ext[0] * g.ext[1] * g.ext[2];
  for (long t = blockIdx.x * 256L + threadIdx.x; t < npts; t += gridDim.x * 256L) {
    const int x = (int)(t % g.ext[0]), y = (int)((t / g.ext[0]) % g.ext[1]);
    const int z = (int)(t / ((long)g.ext[0] * g.ext[1]));
    const long ps = (g.slo[0] + x) + (long)(g.slo[1] + y) * a.sy + (long)(g.slo[2] + z) * a.sz;
    const long pd = (g.dlo[0] + x) + (long)(g.dlo[1] + y) * b.sy + (long)(g.dlo[2] + z) * b.sz;
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) {
      double* d = b.slab + v * b.n + pd;
      const double l = *d;
      *d = 0.0;
      atomicAdd(a.slab + v * a.n + ps, l);
    }
  }
}

// enforce_ga (enforce_algebra.f90:72-127) on every point of every patch.
__global__ void __launch_bounds__(256) enforce_patches_kernel(const PatchDesc* __restrict__ P, int slot) {
  const PatchDesc d = P[blockIdx.y];
  const SlabFields f{(gdouble*)d.slab + (long)slot * NFIELD * d.n, d.n};
  for (long p = blockIdx.x * 256L + threadIdx.x; p < d.n; p += gridDim.x * 256L) {
    double ga[12];
#pragma unroll
    for (int s = 0; s < 12; ++s) ga[s] = f.ptr(DXX + s)[p];
    enforce_point(ga, ga + 6);
#pragma unroll
    for (int s = 0; s < 12; ++s) f.ptr(DXX + s)[p] = ga[s];
  }
}

unsigned xblocks(long maxpts) {
  const long b = (maxpts + 255) / 256;
  return (unsigned)(b < 1 ? 1 : b > 2048 ? 2048 : b);
}

}  // namespace

void launch_rk4_tiles(int gauge, const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m,
                      const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune, int* flag,
                      hipStream_t s) {
#define AMSS_CASE(G, S)                                                                      \
  case S:                                                                                    \
    launch_rk4_tiles_g##G##_s##S(P, T, t0, ntiles, m, C, R, want_res, tune, flag, s); \
    break;
  if (gauge == 2) {
    switch (tune.shape) { AMSS_CASE(2, 0) AMSS_CASE(2, 1) AMSS_CASE(2, 2) AMSS_CASE(2, 3) }
  } else {
    switch (tune.shape) { AMSS_CASE(0, 0) AMSS_CASE(0, 1) AMSS_CASE(0, 2) AMSS_CASE(0, 3) }
  }
#undef AMSS_CASE
}

void launch_fill_boxes(const PatchDesc* P, const FillBox* B, int nbox, int maxpts, const Roles& m, int stage,
                       double tinny, int* flag, hipStream_t s) {
  if (nbox > 0) fill_boxes_kernel<<<dim3(xblocks(maxpts), nbox), 256, 0, s>>>(P, B, m, stage, tinny, flag);
}

void launch_seg_boxes(const PatchDesc* P, const SegRef* S, int nseg, int maxpts, int slot, hipStream_t s) {
  if (nseg > 0) seg_boxes_kernel<<<dim3(xblocks(maxpts), nseg), 256, 0, s>>>(P, S, slot);
}

void launch_seg_boxes_adjoint(const PatchDesc* P, const SegRef* S, int nseg, int maxpts, hipStream_t s) {
  if (nseg > 0) seg_boxes_adjoint_kernel<<<dim3(xblocks(maxpts), nseg), 256, 0, s>>>(P, S);
}

void launch_enforce_patches(const PatchDesc* P, int npatch, long maxn, int slot, hipStream_t s) {
  if (npatch > 0) enforce_patches_kernel<<<dim3(xblocks(maxn), npatch), 256, 0, s>>>(P, slot);
}

}  // namespace amss
