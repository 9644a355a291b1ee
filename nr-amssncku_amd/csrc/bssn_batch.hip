// bssn_batch.hip -- box-table kernels of the batched level step (ghost fill
// and inter-patch exchange, enforce_ga over all patches) and the dispatch of the
// tile-list kernel by gauge and tile shape.
#include "bssn_batch.h"

namespace amss {

namespace {

// Boundary fill and exchange of one sub-step; blockIdx.x is one work entry
// (256 points of one box, so the box kind is uniform per workgroup).  Every
// field is loaded before any is stored, so a wave keeps all its loads in
// flight.  Fill points take the state's values with chi bounded below and,
// for sub-steps 0-2, enforce_ga -- the per-point work of the per-patch ghost
// fill; exchange points take the source patch's sub-step output.
__global__ void __launch_bounds__(256) ghost_boxes_kernel(const PatchDesc* __restrict__ P,
                                                          const GhostBox* __restrict__ B,
                                                          const GhostWork* __restrict__ W, Roles m, int stage,
                                                          double tinny, int check_u0, int* flag) {
  const GhostWork w = W[blockIdx.x];
  const GhostBox b = B[w.box];
  if (b.kind == 0) {
    fill_ghost_point(P, B, w, threadIdx.x, m, stage, tinny, flag);
    return;
  }
  const unsigned t = (unsigned)w.start + threadIdx.x;
  const unsigned e01 = (unsigned)b.ext[0] * b.ext[1];
  if (t >= e01 * b.ext[2]) return;
  const int x = (int)(t % b.ext[0]), y = (int)((t / b.ext[0]) % b.ext[1]), z = (int)(t / e01);
  const PatchDesc d = P[b.dst];
  const long pd = (b.dlo[0] + x) + (long)(b.dlo[1] + y) * d.sy + (long)(b.dlo[2] + z) * d.sz;
  const long ln = (long)NFIELD * d.n;
  double u[NFIELD];
  const PatchDesc a = P[b.src];
  const long ps = (b.slo[0] + x) + (long)(b.slo[1] + y) * a.sy + (long)(b.slo[2] + z) * a.sz;
  const gdouble* src = (const gdouble*)a.slab + m.out * (long)NFIELD * a.n;
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) u[v] = src[v * a.n + ps];
  if (check_u0) {
    const gdouble* u0 = (const gdouble*)d.slab + m.u0 * ln;
    bool bad = false;
#pragma unroll
    for (int v = 0; v < 21; ++v) {
      const double c = u0[v * d.n + pd];
      bad |= !(c == c);
    }
    if (bad) *flag = 1;
  }
  gdouble* out = (gdouble*)d.slab + m.out * ln;
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) out[v * d.n + pd] = u[v];
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
    double u[NFIELD];
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) u[v] = src[v * a.n + ps];
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) dst[v * b.n + pd] = u[v];
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
                      hipStream_t s, const FillWork& fw) {
#define AMSS_CASE(G, S)                                                                      \
  case S:                                                                                    \
    launch_rk4_tiles_g##G##_s##S(P, T, t0, ntiles, m, C, R, want_res, tune, flag, s, fw); \
    break;
  if (gauge == 2) {
    switch (tune.shape) { AMSS_CASE(2, 0) AMSS_CASE(2, 1) AMSS_CASE(2, 2) AMSS_CASE(2, 3) }
  } else {
    switch (tune.shape) { AMSS_CASE(0, 0) AMSS_CASE(0, 1) AMSS_CASE(0, 2) AMSS_CASE(0, 3) }
  }
#undef AMSS_CASE
}

void launch_ghost_boxes(const PatchDesc* P, const GhostBox* B, const GhostWork* W, int nwork, const Roles& m,
                        int stage, double tinny, bool check_u0, int* flag, hipStream_t s) {
  if (nwork > 0) ghost_boxes_kernel<<<nwork, 256, 0, s>>>(P, B, W, m, stage, tinny, check_u0 ? 1 : 0, flag);
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
