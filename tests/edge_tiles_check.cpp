// Host check of the boundary shells and their LDS tiles
// (nr-amssncku_amd/csrc/bssn_edge_tiles.h), built by test_edge_tiles.py.
//
// For every array shape in the sweep, with and without symmetry planes:
//  * the three stencil-order shells and the interior box cover the array
//    once, and each shell point has the order its shell is for (the joint
//    room test of diff_new.f90:1084-1096 with the lowest readable index);
//  * the tiles of shells 1 and 2 give every shell point to exactly one live
//    thread, and every idle thread a point inside its slab;
//  * every index a boundary stencil may read at a live point -- within +-3
//    of it and inside [mlo, n - 1] per axis, mirrored to -m - 1 below 0 --
//    lies in the tile's window, and the window fits one LDS buffer;
//  * the windows start on even x, hold whole 16-byte chunks and fit one
//    LDS buffer, and the mirror image cells of a window below a symmetry
//    plane take their values from cells inside it;
//  * the frozen-ghost fill's enumeration (ghost3_point) gives every point
//    outside [3, n - 3) once and no other.
// Prints "ok <points>" or the first failure, and exits non-zero on failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bssn_edge_tiles.h"

using namespace amss;

static int fail(const char* what, const int n[3], const int mlo[3], int i, int j, int k) {
  std::printf("FAIL %s: n %d %d %d mlo %d %d %d at %d %d %d\n", what, n[0], n[1], n[2], mlo[0], mlo[1], mlo[2], i, j,
              k);
  return 1;
}

static int order_of(const int q[3], const int n[3], const int mlo[3]) {
  int r = 3;
  for (int a = 0; a < 3; ++a) {
    const int dn = q[a] - mlo[a], up = n[a] - 1 - q[a];
    r = r < dn ? r : dn;
    r = r < up ? r : up;
  }
  return r >= 3 ? 3 : r;
}

static int check_shape(const int n[3], const int mlo[3], long& points) {
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = 3;  // the interior box of a level patch: +-3 room in the array itself
    hi[a] = n[a] - 3;
  }
  const EdgeShells es = edge_shells_of(n, lo, hi, mlo);
  const long N = (long)n[0] * n[1] * n[2];
  std::vector<int> owner(N, -1);
  bool interior = true;
  for (int a = 0; a < 3; ++a) interior = interior && hi[a] > lo[a];
  if (interior)
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i) owner[i + n[0] * (j + (long)n[1] * k)] = 3;
  for (int o = 0; o < 3; ++o) {
    const Shell& sh = es.s[o];
    for (long t = 0; t < sh.size(); ++t) {
      int i, j, k;
      sh.point(t, i, j, k);
      const long p = i + n[0] * (j + (long)n[1] * k);
      if (owner[p] != -1) return fail("shell points overlap", n, mlo, i, j, k);
      owner[p] = o;
      const int q[3] = {i, j, k};
      const int want = order_of(q, n, mlo);
      // an order-2 shell point may have room 3 on a symmetry plane (the
      // interior box starts at 3 in the array itself)
      if (!(want == o || (o == 2 && want == 3))) return fail("shell order", n, mlo, i, j, k);
    }
  }
  for (long p = 0; p < N; ++p)
    if (owner[p] == -1) return fail("point in no shell", n, mlo, (int)(p % n[0]), (int)(p / n[0] % n[1]), (int)(p / n[0] / n[1]));

  for (int o = 1; o < 3; ++o) {
    const Shell& sh = es.s[o];
    std::vector<int> seen(N, 0);
    for (int nrm = 0; nrm < 3; ++nrm) {
      const EdgeTiles T = edge_tiles_of(sh, nrm);
      for (int b = 0; b < T.tiles(); ++b) {
        const EdgeTile t = edge_tile(T, b);
        if (t.wlo[0] & 1) return fail("window x not chunk-aligned", n, mlo, t.lo[0], t.lo[1], t.lo[2]);
        long ne = 1;
        for (int a = 0; a < 3; ++a) {
          if (t.wl[a] != edge_window_extent(nrm, a)) return fail("window extent", n, mlo, t.lo[0], t.lo[1], t.lo[2]);
          ne *= t.wl[a];
        }
        const int p0 = nrm == 0 ? EdgeWindow<0>::P0 : (nrm == 1 ? EdgeWindow<1>::P0 : EdgeWindow<2>::P0);
        const int pad = nrm == 0 ? EdgeWindow<0>::PAD : (nrm == 1 ? EdgeWindow<1>::PAD : EdgeWindow<2>::PAD);
        if ((t.wl[0] & 1) || p0 < t.wl[0] || ne / t.wl[0] * p0 > pad) return fail("window shape", n, mlo, t.lo[0], t.lo[1], t.lo[2]);
        // mirror image cells read their source inside the window
        for (int a = 0; a < 3; ++a)
          for (int m = t.wlo[a]; m < 0 && mlo[a] < 0; ++m)
            if (-m - 1 >= t.wlo[a] + t.wl[a]) return fail("mirror source outside the window", n, mlo, t.lo[0], t.lo[1], t.lo[2]);
        for (int tid = 0; tid < kEdgeThreads; ++tid) {
          int q[3];
          bool live;
          edge_tile_point(t, tid, q, live);
          for (int a = 0; a < 3; ++a)
            if (q[a] < 0 || q[a] >= n[a]) return fail("thread point outside the array", n, mlo, q[0], q[1], q[2]);
          const long p = q[0] + n[0] * (q[1] + (long)n[1] * q[2]);
          if (owner[p] != o) return fail(live ? "live point off its shell" : "idle point off its shell", n, mlo, q[0], q[1], q[2]);
          if (!live) continue;
          if (seen[p]++) return fail("shell point given twice", n, mlo, q[0], q[1], q[2]);
          ++points;
          // the window and the readable set are both products of per-axis
          // sets, so the check goes axis by axis; the stencils read indices
          // below 0 through the window's mirror cells
          for (int a = 0; a < 3; ++a)
            for (int d = -3; d <= 3; ++d) {
              const int r = q[a] + d;
              if (r < mlo[a] || r > n[a] - 1) continue;
              const int x = r - t.wlo[a];
              if (x < 0 || x >= t.wl[a]) return fail("stencil read outside the window", n, mlo, q[0], q[1], q[2]);
            }
        }
      }
    }
    for (long p = 0; p < N; ++p)
      if (owner[p] == o && seen[p] != 1) return fail("shell point in no tile", n, mlo, (int)(p % n[0]), (int)(p / n[0] % n[1]), (int)(p / n[0] / n[1]));
  }
  return 0;
}

static int check_ghost3(int nx, int ny, int nz) {
  const int n[3] = {nx, ny, nz}, z[3] = {0, 0, 0};
  std::vector<int> seen((long)nx * ny * nz, 0);
  const long cnt = ghost3_count(nx, ny, nz);
  for (long t = 0; t < cnt; ++t) {
    int i, j, k;
    ghost3_point(t, nx, ny, nz, i, j, k);
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return fail("ghost fill point outside the array", n, z, i, j, k);
    const bool inner = i >= 3 && i < nx - 3 && j >= 3 && j < ny - 3 && k >= 3 && k < nz - 3;
    if (inner) return fail("ghost fill point inside the interior box", n, z, i, j, k);
    if (seen[i + (long)nx * (j + (long)ny * k)]++) return fail("ghost fill point given twice", n, z, i, j, k);
  }
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const bool inner = i >= 3 && i < nx - 3 && j >= 3 && j < ny - 3 && k >= 3 && k < nz - 3;
        if (!inner && !seen[i + (long)nx * (j + (long)ny * k)]) return fail("ghost point not filled", n, z, i, j, k);
      }
  return 0;
}

int main() {
  for (int nx : {14, 15, 16, 22, 38, 134})
    for (int ny : {14, 17, 30})
      for (int nz : {14, 19, 22, 70})
        if (check_ghost3(nx, ny, nz)) return 1;
  long points = 0;
  int shapes = 0;
  std::vector<int> sizes;
  for (int s = 5; s <= 24; ++s) sizes.push_back(s);
  for (int s : {31, 32, 33, 40, 47, 70, 134}) sizes.push_back(s);
  for (int sym = 0; sym < 4; ++sym)
    for (int nx : sizes)
      for (int ny : {5, 6, 7, 9, 16, 19, 22, 23, 38})
        for (int nz : {5, 6, 8, 11, 20, 21, 37}) {
          // sym 0: no plane; 1: z plane (Symmetry 1); 2: x, y, z planes
          // (Symmetry 2); 3: x plane only
          const int mlo[3] = {sym >= 2 ? -3 : 0, sym == 2 ? -3 : 0, (sym == 1 || sym == 2) ? -3 : 0};
          const int n[3] = {nx, ny, nz};
          if (check_shape(n, mlo, points)) return 1;
          ++shapes;
        }
  const int big[][3] = {{134, 134, 134}, {70, 70, 70}, {102, 54, 86}};
  for (auto& b : big)
    for (int sym = 0; sym < 3; ++sym) {
      const int mlo[3] = {sym == 2 ? -3 : 0, sym == 2 ? -3 : 0, sym >= 1 ? -3 : 0};
      if (check_shape(b, mlo, points)) return 1;
      ++shapes;
    }
  std::printf("ok %d shapes, %ld live shell points\n", shapes, points);
  return 0;
}
