// bssn_kernels.h -- launch interface of the gfx950 BSSN kernels.
#pragma once
#include "bssn_env.h"

namespace amss {

// Interior tile of the plain kernels: kTileX x kTileY points of one z-plane.
constexpr int kTileX = 32;
constexpr int kTileY = 8;
constexpr int kRhsThreads = kTileX * kTileY;

// An array of shape n (with ghosts), strides sy, sz, and a box [lo, hi).
// mlo: lowest index a boundary stencil may read along each axis, 0 or -3
// where the array's lower face lies on a symmetry plane (Symmetry 1: z,
// Symmetry 2: x, y, z; diff_new.f90:975-980, lopsidediff.f90:230-232,
// kodiss.f90:156-158); a negative index reads the mirror point with the
// field's parity on that axis.
struct Box {
  int n[3];
  int lo[3];
  int hi[3];
  long sy, sz;
  int mlo[3] = {0, 0, 0};
};

// Widths of the boundary layers filled after a sub-step, per face.
struct Faces {
  int lo[3], hi[3];
};

// One ghost-exchange box: source origin, destination origin, extent, strides.
struct SegDesc {
  int slo[3], dlo[3], ext[3];
  long ssy, ssz, dsy, dsz;
};

struct Rk4Param {
  int stage;     // 0 predictor, 1..3 correctors
  int variant;   // interior kernel: 0 direct (L1/L2 reads), 1 LDS-staged tiles
  double c;      // 0.5 dt, 0.5 dt, dt, dt/6
  double tinny;  // chitiny - 1 (bssn_class.C:179)
  int edge_order = -1;  // boundary launches: the one stencil-order shell to run (-1: all three)
  // interior launches: with fill_w = 3, the frozen-ghost fill of the points
  // outside [3, n - 3) per axis runs as extra workgroups of the LDS kernel's
  // launch where rk4_fuses_fill holds (the caller then launches no
  // launch_ghost_fill); 0: no fill
  int fill_w = 0;
  // right-hand sides of the LDS boundary kernel: edge_k_tiles tiles x NFIELD
  // x 256 doubles (null: the direct boundary kernel takes those shells)
  double* edge_k = nullptr;
  long edge_k_tiles = 0;
};

// RHS of compute_rhs_bssn on the interior box (and, with edge, on the rest
// of the array).  mat must point at 10 valid arrays (zeros for vacuum).
void launch_rhs(const FieldsIn& in, const FieldsIn& mat, const FieldsOut& out, const AuxOut& aux, const Box& box,
                const Coef& C, int gauge, int co, int* flag, bool edge, hipStream_t s);
void launch_rhs_g0(const FieldsIn& in, const FieldsIn& mat, const FieldsOut& out, const AuxOut& aux, const Box& box,
                   const Coef& C, int co, int* flag, bool edge, hipStream_t s);
void launch_rhs_g2(const FieldsIn& in, const FieldsIn& mat, const FieldsOut& out, const AuxOut& aux, const Box& box,
                   const Coef& C, int co, int* flag, bool edge, hipStream_t s);
// RHS fused with RK4 sub-step R.stage (residuals written for stage 0 when want_res).
// sg: Sommerfeld faces, applied by the boundary kernel (edge) on their layers.
// Whether the interior launch of launch_rk4 takes the LDS kernel (variant 1,
// an even row pitch, at least one whole tile) and, with R.fill_w = 3, runs
// the frozen-ghost fill in its launch: the fill's points are exactly those
// outside the interior box [3, n - 3), which no tile writes.  Sub-steps 0
// and 1 only: sub-step 3's kernel, which keeps fewer values in flight, ran
// longer by more than the separate fill launch takes.
inline bool rk4_fuses_fill(const Box& box, const Rk4Param& R) {
  if (R.fill_w != 3 || R.variant != 1 || R.stage > 1 || (box.sy % 2) != 0) return false;
  for (int a = 0; a < 3; ++a)
    if (box.lo[a] != 3 || box.hi[a] != box.n[a] - 3 || box.hi[a] - box.lo[a] < 8) return false;
  return true;
}

void launch_rk4(const FieldsIn& in, const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                const AuxOut& aux, const Box& box, const Coef& C, const Rk4Param& R, int gauge, bool want_res,
                int* flag, bool edge, hipStream_t s, bool interior = true, const SommerGeom& sg = SommerGeom{});
void launch_rk4_g0(const FieldsIn& in, const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                   const AuxOut& aux, const Box& box, const Coef& C, const Rk4Param& R, bool want_res, int* flag,
                   bool edge, hipStream_t s, bool interior, const SommerGeom& sg);
void launch_rk4_g2(const FieldsIn& in, const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                   const AuxOut& aux, const Box& box, const Coef& C, const Rk4Param& R, bool want_res, int* flag,
                   bool edge, hipStream_t s, bool interior, const SommerGeom& sg);
// sommerfeld_routbam on one field: f_rhs on the face layers of sg from f0
// (parity S per axis for the mirrored layer), velocity vel.
void launch_sommer_bam(const double* f0, double* f_rhs, const Shape& n, const SommerGeom& sg, double vel,
                       const double S[3], hipStream_t s);
// sommerfeld_rout(precor = 0) on one field: f on the outermost layer of the
// faces in the mask (bits as SommerGeom::faces) from f0, the base state's
// chi0 and lap0 and the coordinate arrays X, Y, Z (device); *bad is set when
// a point's interpolation block cannot be placed inside the array.
void launch_sommer_interp(const double* X, const double* Y, const double* Z, const double* chi0, const double* lap0,
                          const double* f0, double* f, const Shape& n, int faces, double dT, const double S[3],
                          int symmetry, int* bad, hipStream_t s);
void launch_ghost_fill(const FieldsIn& u0, const FieldsOut& out, const Box& box, const Faces& w, int stage,
                       double tinny, int* flag, hipStream_t s);
void launch_seg_copy(const FieldsIn& src, const FieldsOut& dst, const SegDesc& d, hipStream_t s);
void launch_enforce(const FieldsOut& f, long n, hipStream_t s);
void launch_lowerbound(double* chi, long n, double tinny, hipStream_t s);
void launch_rk4_update(const double* f0, double* f1, double* frhs, long n, double dT, int stage, hipStream_t s);
void launch_copy_box(const double* f0, double* f, const Box& box, hipStream_t s);
int launch_sumsq(const double* f, const Box& box, double* part, int maxblocks, hipStream_t s);

// out = c1 f1 + c2 f2 (+ c3 f3) over n points of all 24 fields (the time
// interpolation of prepare_inter_time_level).
void launch_time_interp(const FieldsIn& f1, const FieldsIn& f2, const FieldsIn& f3, const FieldsOut& out, long n,
                        double c1, double c2, double c3, bool three, hipStream_t s);

// One L2 reduction: dv times the sum of squares of f over the index box
// [lo, hi) (strides sy, sz), reduced through nparts workgroup partials at
// part[part0 ..] and then summed in a fixed order (deterministic results).
struct L2Job {
  const double* f;
  int lo[3], hi[3];
  long sy, sz;
  double dv;
  int part0, nparts;
};
// out[j] = job j's sum for every job (jobs: device table of njobs; maxparts:
// the largest nparts).  Stream-ordered, no synchronisation.
void launch_l2_jobs(const L2Job* jobs, int njobs, int maxparts, double* part, double* out, hipStream_t s);

}  // namespace amss
