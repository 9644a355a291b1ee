// amss_amr.h -- a hierarchy of refinement levels on one GPU, advanced with
// Berger-Oliger sub-cycling in the order of bssn_class::RecursiveStep
// (bssn_class.C:2323-2403) and RestrictProlong (:5724-5904): each level is a
// BssnLevel (its own Step), the levels exchange data through the 6-point
// restriction / prolongation kernels and the time interpolation of
// prepare_inter_time_level (Parallel.C:4329-4441).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "amss_host.h"
#include "amss_xfer.h"

namespace amss {

class AmrHierarchy {
 public:
  // nlev levels; npatch[l] patches on level l; shapes / bboxes of all
  // patches, level after level (as amss_level_create).  Level 0 uses
  // boundary mode bc0 and ghost width 3; a finer level halves the spacing,
  // its patch arrays carry the reference's buffer zone (buffer_width = 6
  // cells, macrodef.h:88) inside their bbox on every face but a lower face
  // on a symmetry plane, and its Step copies f0 onto the outermost layer
  // (sommerfeld_rout with cor, bssn_class.C:3124-3135).
  AmrHierarchy(int nlev, const int* npatch, const int* shapes, const double* bboxes, int gauge, double eps,
               double chitiny, int bc0, int symmetry);
  ~AmrHierarchy();
  bool ok() const { return err_.empty(); }
  const std::string& err() const { return err_; }
  int nlev() const { return (int)lev_.size(); }
  BssnLevel* level(int l) { return lev_[l]; }
  // nsteps coarse steps of size dT (RecursiveStep(0) each); mode and
  // want_res as BssnLevel::step, for every level step
  int step(double dT, int nsteps, int mode, int want_res);
  double time() const { return phys_time_; }
  double level_time(int l) const { return lt_[l]; }
  // after a restart (checkpoint::readcheck_cgh): PhysTime and each level's
  // time; the start time keeps its value (bssn_class.C:699-731)
  void restore_time(double t, const double* lt) {
    phys_time_ = t;
    for (size_t l = 0; l < lt_.size(); ++l) lt_[l] = lt[l];
  }

  static constexpr int kBufferWidth = 6;

 private:
  struct Xfer {
    int c, f;     // coarse / fine patch
    XferPlan P;
  };
  int recursive_step(int l, double dT, int mode, int want_res);
  int restrict_prolong(int l, bool same_time, bool three_levels);
  int plan_transfers(int l);

  std::vector<BssnLevel*> lev_;
  std::vector<std::vector<std::array<int, 6>>> buf_;  // per level and patch: lli[3], uui[3]
  std::vector<std::vector<Xfer>> restrict_, prolong_;  // [l]: between levels l - 1 and l
  std::vector<double> lt_;                             // each level's time (GH->Lt)
  int symmetry_;
  double phys_time_ = 0.0, start_time_ = 0.0;
  hipStream_t stream_ = nullptr;
  std::string err_;
};

}  // namespace amss
