// bssn_edge_tiles.h -- the boundary shells of an array and the LDS tiles
// over them.
//
// The points outside the interior box split into three shells by the order
// of their centred stencils; each shell is six slabs, one point deep along
// the slab's normal axis except where an axis is short.  The tiles of the
// order-2 and order-1 shells are 16 x 16 points of one layer of a slab (the
// two in-plane axes in index order, u fastest), one 256-thread workgroup per
// tile.  Every field the algebra opens is staged as the tile's window --
// the tile plus 3 points on every side, x widened to whole 16-byte chunks,
// with pitches fixed by the slab's normal axis -- so the boundary stencils
// (which stay within +-3 of the point, and whose mirrored reads below a
// symmetry plane land on indices 0..2) read LDS only, at offsets known at
// compile time.
//
// Host and device code share these functions; they hold no HIP types, so a
// host compiler can include this header on its own.
#pragma once

#if defined(__HIPCC__)
#define AMSS_HD __host__ __device__ inline
#else
#define AMSS_HD inline
#endif

namespace amss {

// The points of an outer box [OL, OH) outside an inner box [IL, IH) (clamped
// into it), enumerated as six disjoint slabs (z faces whole, then y faces
// over the inner z range, then x faces over the inner y and z ranges), so
// that every lane of a wave has a point: walking the whole array instead
// leaves an x-face row's points alone in a wave.  An inner box empty along
// any axis holds no point, and the slabs then cover the whole outer box.
struct Shell {
  int OL[3], OH[3], IL[3], IH[3];
  long cnt[6];
  AMSS_HD Shell(const int ol[3], const int oh[3], const int il[3], const int ih[3]) {
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
  AMSS_HD long size() const { return cnt[0] + cnt[1] + cnt[2] + cnt[3] + cnt[4] + cnt[5]; }
  // point t of the shell (t < size())
  AMSS_HD void point(long t, int& i, int& j, int& k) const {
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
// n: the array's shape, [lo, hi) its interior box.
AMSS_HD EdgeShells edge_shells_of(const int n[3], const int lo[3], const int hi[3], const int mlo[3]) {
  int z[3] = {0, 0, 0}, L1[3], H1[3], L2[3], H2[3];
  for (int a = 0; a < 3; ++a) {
    L1[a] = mlo[a] + 1 > 0 ? mlo[a] + 1 : 0;
    H1[a] = n[a] - 1;
    L2[a] = mlo[a] + 2 > 0 ? mlo[a] + 2 : 0;
    H2[a] = n[a] - 2;
    if (H2[a] < L2[a]) H2[a] = L2[a];
    if (H1[a] < L1[a]) H1[a] = L1[a];
  }
  int blo[3], bhi[3];
  bool empty = false;
  for (int a = 0; a < 3; ++a) {
    blo[a] = lo[a];
    bhi[a] = hi[a];
    empty = empty || hi[a] <= lo[a];
  }
  if (empty)  // no interior launch: the order-2 shell takes its box whole
    for (int a = 0; a < 3; ++a) bhi[a] = blo[a] = L2[a];
  return EdgeShells{{Shell(z, n, L1, H1), Shell(L1, H1, L2, H2), Shell(L2, H2, blo, bhi)}};
}

constexpr int kEdgeTile = 16;                        // points per in-plane tile axis
constexpr int kEdgeThreads = kEdgeTile * kEdgeTile;  // one point per thread

// The window of a tile whose slab has normal axis nrm: its extent along
// axis a (x: the 22 or 23 points from an even start, in whole chunks).
AMSS_HD constexpr int edge_window_extent(int nrm, int a) {
  return a == 0 ? (nrm == 0 ? 8 : kEdgeTile + 8) : (nrm == a ? 7 : kEdgeTile + 6);
}
// Its layout in LDS: rows of P0 doubles (P0 >= W0: an x-normal window's
// rows of 8 would put the neighbours along y, 16 dwords apart, on the same
// LDS banks; rows of 10 spread them), of which the DMA fills the first W0.
template <int NRM>
struct EdgeWindow {
  static constexpr int W0 = edge_window_extent(NRM, 0);
  static constexpr int W1 = edge_window_extent(NRM, 1);
  static constexpr int W2 = edge_window_extent(NRM, 2);
  static constexpr int P0 = NRM == 0 ? W0 + 2 : W0;
  static constexpr int P01 = P0 * W1;
  static constexpr int NE = P01 * W2;  // doubles, row padding included
  static constexpr int ROUNDS = (NE / 2 + kEdgeThreads - 1) / kEdgeThreads;  // chunks per thread
  static constexpr int PAD = 2 * ROUNDS * kEdgeThreads;                      // doubles per buffer
  static_assert(W0 % 2 == 0 && P0 % 2 == 0, "rows of whole 16-byte chunks");
};

struct EdgeTiles {
  int nslab = 0;
  int org[6][3];   // first point of each slab
  int end[6][3];   // one past its last point, per axis
  int nrm[6];      // normal axis (tiles are one layer deep along it)
  int nt[6][2];    // tiles along the two in-plane axes
  int first[7];    // first tile of each slab; first[nslab] = number of tiles
  AMSS_HD int tiles() const { return first[nslab]; }
};

// The in-plane axes of a slab with normal axis n, in index order.
AMSS_HD void plane_axes(int n, int& a0, int& a1) {
  a0 = n == 0 ? 1 : 0;
  a1 = n == 2 ? 1 : 2;
}

// The normal axis of a slab of extents e: the one it is thinnest along
// (ties: the highest axis, so a z-face keeps x fastest in its tiles).
AMSS_HD int slab_normal(const int e[3]) {
  int n = 2;
  for (int a = 1; a >= 0; --a)
    if (e[a] < e[n]) n = a;
  return n;
}

// Tiles of those slabs [lo, hi) (six at most; empty ones are skipped) whose
// normal axis is nrm.
AMSS_HD EdgeTiles make_edge_tiles(const int lo[][3], const int hi[][3], int nslab, int nrm) {
  EdgeTiles T;
  T.nslab = 0;
  T.first[0] = 0;
  for (int s = 0; s < nslab && s < 6; ++s) {
    int e[3];
    bool empty = false;
    for (int a = 0; a < 3; ++a) {
      e[a] = hi[s][a] - lo[s][a];
      empty = empty || e[a] <= 0;
    }
    if (empty || slab_normal(e) != nrm) continue;
    const int n = nrm;
    int a0, a1;
    plane_axes(n, a0, a1);
    const int q = T.nslab;
    for (int a = 0; a < 3; ++a) {
      T.org[q][a] = lo[s][a];
      T.end[q][a] = hi[s][a];
    }
    T.nrm[q] = n;
    T.nt[q][0] = (e[a0] + kEdgeTile - 1) / kEdgeTile;
    T.nt[q][1] = (e[a1] + kEdgeTile - 1) / kEdgeTile;
    T.first[q + 1] = T.first[q] + T.nt[q][0] * T.nt[q][1] * e[n];
    T.nslab = q + 1;
  }
  return T;
}

// The tiles of a shell's slabs (Shell's order) with normal axis nrm.
AMSS_HD EdgeTiles edge_tiles_of(const Shell& sh, int nrm) {
  const int lo[6][3] = {{sh.OL[0], sh.OL[1], sh.OL[2]}, {sh.OL[0], sh.OL[1], sh.IH[2]},
                        {sh.OL[0], sh.OL[1], sh.IL[2]}, {sh.OL[0], sh.IH[1], sh.IL[2]},
                        {sh.OL[0], sh.IL[1], sh.IL[2]}, {sh.IH[0], sh.IL[1], sh.IL[2]}};
  const int hi[6][3] = {{sh.OH[0], sh.OH[1], sh.IL[2]}, {sh.OH[0], sh.OH[1], sh.OH[2]},
                        {sh.OH[0], sh.IL[1], sh.IH[2]}, {sh.OH[0], sh.OH[1], sh.IH[2]},
                        {sh.IL[0], sh.IH[1], sh.IH[2]}, {sh.OH[0], sh.IH[1], sh.IH[2]}};
  return make_edge_tiles(lo, hi, 6, nrm);
}

// The tiles of the order-1 and order-2 shells in one launch: set
// s = 3 (order - 1) + normal axis holds blocks [first[s], first[s + 1]).
struct EdgeTileSet {
  EdgeTiles t[6];
  int first[7];
  AMSS_HD int tiles() const { return first[6]; }
  // the set of block b
  AMSS_HD int set_of(int b) const {
    int s = 0;
    while (s < 5 && b >= first[s + 1]) ++s;
    return s;
  }
};
AMSS_HD EdgeTileSet edge_tile_set(const EdgeShells& es) {
  EdgeTileSet S;
  S.first[0] = 0;
  for (int s = 0; s < 6; ++s) {
    S.t[s] = edge_tiles_of(es.s[1 + s / 3], s % 3);
    S.first[s + 1] = S.first[s] + S.t[s].tiles();
  }
  return S;
}

// One tile: its first point, its extent per axis (16, 16 in-plane, 1 along
// the normal), the slab's end, and the window [wlo, wlo + wl) staged in LDS
// (wl: EdgeWindow's extents for the slab's normal).
struct EdgeTile {
  int lo[3], ext[3], end[3];
  int a0, a1;            // in-plane axes of the threads' u and w
  int wlo[3], wl[3];
};

AMSS_HD EdgeTile edge_tile(const EdgeTiles& T, int b) {
  int s = 0;
  while (s + 1 < T.nslab && b >= T.first[s + 1]) ++s;
  int t = b - T.first[s];
  EdgeTile E;
  const int n = T.nrm[s];
  plane_axes(n, E.a0, E.a1);
  const int tu = t % T.nt[s][0];
  t /= T.nt[s][0];
  const int tw = t % T.nt[s][1];
  const int layer = t / T.nt[s][1];
  for (int a = 0; a < 3; ++a) {
    E.end[a] = T.end[s][a];
    E.lo[a] = T.org[s][a];
    E.ext[a] = 1;
  }
  E.lo[E.a0] += tu * kEdgeTile;
  E.lo[E.a1] += tw * kEdgeTile;
  E.lo[n] += layer;
  E.ext[E.a0] = E.ext[E.a1] = kEdgeTile;
  // x starts on an even index: the direct loads move 16 bytes from an even
  // index of an even-pitched array
  for (int a = 0; a < 3; ++a) {
    E.wlo[a] = E.lo[a] - 3;
    E.wl[a] = edge_window_extent(n, a);
  }
  E.wlo[0] -= E.wlo[0] & 1;
  return E;
}

// The point of thread tid (u = tid % 16 along a0, w = tid / 16 along a1);
// live is false past the slab's end, where the point is clamped into the
// slab so that the thread reads valid memory but must store nothing.
AMSS_HD void edge_tile_point(const EdgeTile& E, int tid, int q[3], bool& live) {
  for (int a = 0; a < 3; ++a) q[a] = E.lo[a];
  q[E.a0] += tid % kEdgeTile;
  q[E.a1] += tid / kEdgeTile;
  live = true;
  for (int a = 0; a < 3; ++a)
    if (q[a] >= E.end[a]) {
      live = false;
      q[a] = E.end[a] - 1;
    }
}

// Point t of the points of an nx x ny x nz array outside [3, n - 3) per axis
// (z slabs whole, then y slabs over the inner z range, then x slabs over the
// inner y and z ranges), in scalars only: t < ghost3_count(n).
AMSS_HD long ghost3_count(int nx, int ny, int nz) {
  return (long)nx * ny * nz - (long)(nx - 6) * (ny - 6) * (nz - 6);
}
AMSS_HD void ghost3_point(long t, int nx, int ny, int nz, int& i, int& j, int& k) {
  const long zs = 3L * nx * ny, ys = 3L * nx * (nz - 6), xs = 3L * (ny - 6) * (nz - 6);
  int x0 = 0, ex = nx, y0 = 0, ey = ny, z0 = 0;
  if (t < 2 * zs) {  // z slabs
    z0 = t < zs ? 0 : nz - 3;
    t -= t < zs ? 0 : zs;
  } else if ((t -= 2 * zs) < 2 * ys) {  // y slabs
    y0 = t < ys ? 0 : ny - 3;
    ey = 3;
    z0 = 3;
    t -= t < ys ? 0 : ys;
  } else {  // x slabs
    t -= 2 * ys;
    x0 = t < xs ? 0 : nx - 3;
    ex = 3;
    y0 = 3;
    ey = ny - 6;
    z0 = 3;
    t -= t < xs ? 0 : xs;
  }
  i = x0 + (int)(t % ex);
  j = y0 + (int)((t / ex) % ey);
  k = z0 + (int)(t / ((long)ex * ey));
}

}  // namespace amss
