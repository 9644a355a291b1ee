// amss_xfer.h -- inter-level transfer between 2:1 cell-centred grids
// (prolongrestrict_cell.f90, Cell, ghost_width 3).
#pragma once
#include <hip/hip_runtime.h>

namespace amss {

// Index plan of one transfer, in the reference's 1-based cell numbering:
// the target loop range [lo, hi] (fine array for a prolongation, coarse for
// a restriction), the first cell of each array from the common base (lbf,
// lbc), and both array extents.
struct XferPlan {
  int lo[3], hi[3];
  int lbf[3], lbc[3];
  int ext_c[3], ext_f[3];
};

// Plan a transfer of the region [ll, uu] (prolong: a fine region; restrict:
// a coarse one).  Returns 0; 1 where the reference reports a bad region
// (and where a stencil would leave the source and its mirrored halo); 2 for
// grids that are not 2:1.
int xfer_plan(const double* llbc, const double* uubc, const int* extc, const double* llbf, const double* uubf,
              const int* extf, const double* ll, const double* uu, bool prolong, XferPlan& P);
// S: the field's parity per axis (for the mirrored halo below index 1).
void launch_prolong3(const double* func, double* funf, const XferPlan& P, const double S[3], hipStream_t s);
void launch_restrict3(const double* funf, double* func, const XferPlan& P, const double S[3], hipStream_t s);

}  // namespace amss
