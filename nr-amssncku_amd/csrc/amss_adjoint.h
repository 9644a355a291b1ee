// amss_adjoint.h -- discrete adjoint of one RK4 step of a refinement level
// (bssn_class::Step with frozen ghost zones of width 3, GAUGE 0, vacuum), on
// device arrays.  Per step, the stage states of every Block are re-evaluated
// from U^n (with the inter-patch exchange after each sub-step, as the level
// step does), then the step is run backwards: the RK4 combination, chi's
// lower bound, enforce_ga, the right-hand side (reverse algebra + transposed
// stencils) and the exchange (each exchanged ghost value's adjoint is added
// to the source point it was copied from).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bssn_batch.h"

namespace amss {

// Work arrays and stage operations of one Block.
class AdjointBlock {
 public:
  AdjointBlock(const int ex[3], const Coef& C, double tinny);
  ~AdjointBlock();
  bool ok() const { return err_.empty(); }
  const std::string& err() const { return err_; }
  long npts() const { return N_; }
  long padded_npts() const { return NP_; }
  const Box& box() const { return box_; }

  // Stage data (Y0..Y3, z0..z3: 192 grid functions) live in the work array
  // (slot -1) or in slot s of a tape reserved for a segment of steps.
  int reserve(int slots);
  void use_slot(int slot) { slot_ = slot; }
  // forward: Y0 = E(U); stage s: k = RHS(Y_s), z_s, Y_{s+1} (s = 3: into Unext
  // or a scratch array)
  void fwd_open(const double* const* U, hipStream_t s);
  void fwd_stage(int st, double c, double* const* Unext, hipStream_t s);
  // reverse: adj_final from lam; per stage: RHS adjoint into lY (or lY0 for st = 0)
  void rev_final(const double* const* lam, double c, hipStream_t s);
  void rev_rhs(int st, hipStream_t s);
  void rev_stage(int st, double c, hipStream_t s);  // st = 1..3: consumes lY of Y_st
  void rev_open(const double* const* U, double* const* lam, hipStream_t s);
  void rhs_vjp(const double* const* U, const double* const* lam, double* const* out, hipStream_t s);
  // a 24-field run of the work slab as a PatchDesc "slot 0" (for the exchange kernels)
  PatchDesc desc(int which, int slot) const;
  double* stage_output(int st) const;  // Y_{st+1} (st = 0..2) or the scratch U^{n+1} (st = 3)
  double* lY() const;

 private:
  FieldsIn fin(int first) const;
  FieldsOut fout(int first) const;
  double* padded(int first) const;  // element 0 of padded array `first`
  FieldsIn pin(int first) const;
  FieldsOut pout(int first) const;
  Box box_;
  Coef C_;
  double tinny_;
  long N_ = 0;
  double* buf_ = nullptr;
  double* padbuf_ = nullptr;  // padded arrays: NP_ = N_ + 2 pad_ doubles each, pad_ = 3 planes
  long pad_ = 0, NP_ = 0;
  double* tape_ = nullptr;
  int slots_ = 0;
  int slot_ = -1;
  double* base_of(int first, int slot) const;
  std::string err_;
};

class LevelAdjoint {
 public:
  // shapes: 3 ints per patch; bboxes: 6 doubles per patch; all patches share
  // one grid spacing, (bbox max - min) / shape unless h (3 doubles) is given
  LevelAdjoint(int npatch, const int* shapes, const double* bboxes, double eps, double chitiny,
               const double* h = nullptr);
  ~LevelAdjoint();
  bool ok() const { return err_.empty(); }
  const std::string& err() const { return err_; }
  int npatch() const { return (int)blocks_.size(); }
  // U, Unext, lam: npatch x 24 device pointers (patch-major).  slot >= 0:
  // keep / use the stage data of this step in that tape slot (backward then
  // skips re-evaluating the step); -1: the work arrays (backward re-evaluates).
  int reserve(int slots);
  int forward(double dT, const double* const* U, double* const* Unext, hipStream_t s, int slot = -1);
  int backward(double dT, const double* const* U, double* const* lam, hipStream_t s, int slot = -1);
  int rhs_vjp(const double* const* U, const double* const* lam, double* const* out, hipStream_t s);
  size_t bytes() const;

 private:
  // table: 0..2 = Y1..Y3, 3 = the U^{n+1} scratch / dJ/dY run
  void exchange(int table, int slot, hipStream_t s);  // copy source interiors into ghost boxes
  void exchange_adjoint(hipStream_t s);                // add the dJ/dY run's ghost-box adjoints to the sources
  PatchDesc* table(int which, int slot);
  std::vector<AdjointBlock*> blocks_;
  std::vector<SegRef> hseg_;
  SegRef* dseg_ = nullptr;
  long seg_max_ = 0;
  std::vector<PatchDesc*> dtab_;  // device PatchDesc tables: Y1, Y2, Y3 of the work arrays, the dJ/dY run,
                                  // then Y1..Y3 of each tape slot
  std::string err_;
};

}  // namespace amss
