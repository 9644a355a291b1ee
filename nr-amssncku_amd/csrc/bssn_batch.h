// bssn_batch.h -- one refinement level advanced as a whole: every patch's
// interior is cut into LDS tiles and one launch per RK4 sub-step walks the
// tile list of all patches (persistent workgroups striding over the list), so
// small and uneven patches fill the GPU together.  Ghost fill, inter-patch
// ghost exchange and enforce_ga run as one launch each over box tables.
//
// Field addressing: a Block keeps its 5 time levels as consecutive runs of 24
// fields in one slab (list slot s, field v at slab + (24 s + v) n) and the 7
// residuals after them; Block::swapList only permutes which slot a list
// names, so a launch carries the slot of each role instead of 96 pointers.
#pragma once
#include "bssn_kernels.h"

namespace amss {

// Device view of one Block.
struct PatchDesc {
  double* slab;  // (5 * 24 + 7) * n doubles
  long n;        // points per grid function
  int sy, sz;    // strides (shape[0], shape[0] * shape[1])
  int shape[3];
};

// One tile of the interior: patch index and the array index of its first point.
struct TileRef {
  unsigned short patch, i0, j0, k0;
};

// List slot (0..4) of each role of a fused sub-step.
struct Roles {
  int in, u0, acc, out;
};
constexpr int kResidualBase = 5 * NFIELD;  // first residual grid function of a slab

// A copy of a box between patches (the level adjoint's exchange).
struct SegRef {
  int src, dst;
  int slo[3], dlo[3], ext[3];
};

// Ghost work of the batched step, as disjoint boxes of destination points:
// kind 0 sets points of patch dst from its state u0 (the frozen or copied
// boundary: chi's lower bound, enforce_ga for sub-steps 0-2), kind 1 copies
// patch src's box at slo of the sub-step output into patch dst's box at dlo
// (the exchange).  Fill boxes exclude the points an exchange box overwrites.
struct GhostBox {
  int kind, src, dst;
  int slo[3], dlo[3], ext[3];
};
// Points [start, start + 256) of box `box` (one workgroup each).
struct GhostWork {
  int box, start;
};

// One fill point of work item w (box kind 0): the state's values with chi
// bounded below and, for sub-steps 0-2, enforce_ga; every field is loaded
// before any is stored (all loads of a wave in flight together).
__device__ __forceinline__ void fill_ghost_point(const PatchDesc* __restrict__ P, const GhostBox* __restrict__ B,
                                                 const GhostWork& w, unsigned lane, const Roles& m, int stage,
                                                 double tinny, int* flag) {
  const GhostBox b = B[w.box];
  const unsigned t = (unsigned)w.start + lane;
  const unsigned e01 = (unsigned)b.ext[0] * b.ext[1];
  if (t >= e01 * b.ext[2]) return;
  const int x = (int)(t % b.ext[0]), y = (int)((t / b.ext[0]) % b.ext[1]), z = (int)(t / e01);
  const PatchDesc d = P[b.dst];
  const long pd = (b.dlo[0] + x) + (long)(b.dlo[1] + y) * d.sy + (long)(b.dlo[2] + z) * d.sz;
  const long ln = (long)NFIELD * d.n;
  const gdouble* u0 = (const gdouble*)d.slab + m.u0 * ln;
  gdouble* out = (gdouble*)d.slab + m.out * ln;
  double u[NFIELD];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) u[v] = u0[v * d.n + pd];
  bool bad = false;
#pragma unroll
  for (int v = 0; v < 21; ++v) bad |= !(u[v] == u[v]);  // bssn_rhs.f90:110-113
  if (bad) *flag = 1;
  if (stage < 3) enforce_point(u + DXX, u + AXX);
  u[CHI] = u[CHI] < tinny ? tinny : u[CHI];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) out[v * d.n + pd] = u[v];
}

// Fill work that a tile launch runs in extra workgroups after its tiles
// (frozen ghosts: the fill reads only the state and writes only ghost points
// no tile writes, so it overlaps the latency-bound tiles and their tail).
struct FillWork {
  const GhostBox* B = nullptr;
  const GhostWork* W = nullptr;
  int n = 0;  // 256-point work items
};

// Tile shapes of the batched kernel (points per tile along x, y, z).
constexpr int kNumTileShapes = 4;
__host__ __device__ constexpr int tile_dim(int shape, int a) {
  // 0: 8x8x4, 1: 4x8x8, 2: 8x4x8 (256 threads); 3: 8x8x8 (512 threads)
  return shape == 0 ? (a == 2 ? 4 : 8) : shape == 1 ? (a == 0 ? 4 : 8) : shape == 2 ? (a == 1 ? 4 : 8) : 8;
}

// Workgroup t runs on XCD t % 8 (round-robin dispatch over the 8 XCDs of an
// MI355X).  Giving XCD x the contiguous run of tiles [start(x), start(x+1))
// keeps neighbouring tiles, whose halos overlap, behind the same L2.
__host__ __device__ __forceinline__ int xcd_spread(int t, int n) {
  const int x = t & 7, l = t >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + l;
}

struct BatchTune {
  int shape;   // tile shape index
  int grid;    // workgroups per launch (persistent); 0 = one per tile
  int remap;   // 1 = XCD-contiguous tile order
};

// Fused RHS + RK4 sub-step over tiles [t0, t0 + ntiles) of the tile list.
// fw: fill work run in extra workgroups (only with one workgroup per tile:
// launch_rk4_tiles_fuses_fill).
void launch_rk4_tiles(int gauge, const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m,
                      const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune, int* flag,
                      hipStream_t s, const FillWork& fw = FillWork{});
inline bool launch_rk4_tiles_fuses_fill(const BatchTune& tune, int ntiles) {
  return !(tune.grid > 0 && tune.grid < ntiles) && ntiles > 0;
}
#define AMSS_DECL_TILES(G, S)                                                                             \
  void launch_rk4_tiles_g##G##_s##S(const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m, \
                                    const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune,  \
                                    int* flag, hipStream_t s, const FillWork& fw);
AMSS_DECL_TILES(0, 0)
AMSS_DECL_TILES(0, 1)
AMSS_DECL_TILES(0, 2)
AMSS_DECL_TILES(0, 3)
AMSS_DECL_TILES(2, 0)
AMSS_DECL_TILES(2, 1)
AMSS_DECL_TILES(2, 2)
AMSS_DECL_TILES(2, 3)
#undef AMSS_DECL_TILES
// Boundary fill and exchange of a sub-step over the work list W (nwork
// entries); check_u0 also tests the state's values at exchange points for
// NaN (the fill tests them at its own points).
void launch_ghost_boxes(const PatchDesc* P, const GhostBox* B, const GhostWork* W, int nwork, const Roles& m,
                        int stage, double tinny, bool check_u0, int* flag, hipStream_t s);
// Inter-patch ghost exchange of list slot `slot`.
void launch_seg_boxes(const PatchDesc* P, const SegRef* S, int nseg, int maxpts, int slot, hipStream_t s);
// Reverse of launch_seg_boxes (slot 0 of P): each ghost-box value's adjoint is
// added to the source point it was copied from (atomically: a source point
// can feed the ghosts of several patches) and the ghost value's adjoint zeroed.
void launch_seg_boxes_adjoint(const PatchDesc* P, const SegRef* S, int nseg, int maxpts, hipStream_t s);
// enforce_ga on every point of list slot `slot` of every patch.
void launch_enforce_patches(const PatchDesc* P, int npatch, long maxn, int slot, hipStream_t s);

}  // namespace amss
