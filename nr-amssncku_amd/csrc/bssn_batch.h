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
void launch_rk4_tiles(int gauge, const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m,
                      const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune, int* flag,
                      hipStream_t s);
#define AMSS_DECL_TILES(G, S)                                                                             \
  void launch_rk4_tiles_g##G##_s##S(const PatchDesc* P, const TileRef* T, int t0, int ntiles, const Roles& m, \
                                    const Coef& C, const Rk4Param& R, bool want_res, const BatchTune& tune,  \
                                    int* flag, hipStream_t s);
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
