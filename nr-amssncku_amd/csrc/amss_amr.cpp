// amss_amr.cpp -- Berger-Oliger hierarchy of BssnLevels on one GPU:
// RecursiveStep / RestrictProlong of bssn_class.C (:2323-2403, :5724-5904)
// with the regions of Parallel::Restrict / OutBdLow2Hi (Parallel.C:4487-4640)
// planned once on the host (one level's Blocks are its patches here).
#include "amss_amr.h"

#include <algorithm>
#include <cmath>

#include "../../include/amss_bssn.h"

namespace amss {

namespace {

struct PBox {  // a box of physical coordinates (cell faces)
  double lo[3], hi[3];
};

// The overlap of two boxes, empty by the Cell rule of build_gstl
// (Parallel.C:3205-3245): less than half the smaller spacing wide.
bool overlap(const PBox& a, const PBox& b, const double hmin[3], PBox& out) {
  for (int d = 0; d < 3; ++d) {
    out.lo[d] = std::max(a.lo[d], b.lo[d]);
    out.hi[d] = std::min(a.hi[d], b.hi[d]);
    if (out.hi[d] - out.lo[d] < hmin[d] / 2) return false;
  }
  return true;
}

long cells(const PBox& b, const double h[3]) {
  long n = 1;
  for (int d = 0; d < 3; ++d) n *= std::lround((b.hi[d] - b.lo[d]) / h[d]);
  return n;
}

bool on_plane(int axis, int symmetry, double x, double h) {
  return (axis == 2 ? symmetry > 0 : symmetry > 1) && std::fabs(x) < h / 2;
}

}  // namespace

AmrHierarchy::AmrHierarchy(int nlev, const int* npatch, const int* shapes, const double* bboxes, int gauge,
                           double eps, double chitiny, int bc0, int symmetry)
    : symmetry_(symmetry) {
  if (nlev < 1) { err_ = "at least one level"; return; }
  int off = 0;
  for (int l = 0; l < nlev; ++l) {
    if (npatch[l] < 1) { err_ = "every level needs a patch"; return; }
    const int bc = l == 0 ? bc0 : AMSS_BC_COPY_LAYER;
    auto* L = new BssnLevel(npatch[l], shapes + 3 * off, bboxes + 6 * off, gauge, eps, chitiny, bc, 3, symmetry);
    lev_.push_back(L);
    if (!L->ok()) { err_ = "level " + std::to_string(l) + ": " + L->err(); return; }
    buf_.emplace_back();
    for (int p = 0; p < npatch[l]; ++p) {
      std::array<int, 6> w{};
      const Block& b = L->block(p);
      for (int a = 0; a < 3 && l > 0; ++a) {
        w[a] = on_plane(a, symmetry, b.bbox[a], b.getdX(a)) ? 0 : kBufferWidth;
        w[3 + a] = kBufferWidth;
      }
      buf_[l].push_back(w);
    }
    off += npatch[l];
  }
  lt_.assign(nlev, 0.0);
  restrict_.resize(nlev);
  prolong_.resize(nlev);
  for (int l = 1; l < nlev; ++l)
    if (plan_transfers(l)) return;
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) err_ = "hipStreamCreate failed";
}

AmrHierarchy::~AmrHierarchy() {
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
  for (BssnLevel* L : lev_) delete L;
}

// The transfers between level l - 1 and level l:
//  restriction: every fine patch's cells less its buffer zone
//    (build_owned_gsl2) onto the whole of every coarse patch
//    (build_complete_gsl), Parallel::Restrict;
//  prolongation: every fine patch's buffer zone (build_buffer_gsl) from the
//    coarse patches less their buffer zone and three ghost cells, except on
//    a symmetry plane (build_owned_gsl4), Parallel::OutBdLow2Hi.
int AmrHierarchy::plan_transfers(int l) {
  BssnLevel& C = *lev_[l - 1];
  BssnLevel& F = *lev_[l];
  const double hc = C.block(0).getdX(0), hf = F.block(0).getdX(0);
  for (int a = 0; a < 3; ++a)
    if (std::fabs(C.block(0).getdX(a) - 2 * F.block(0).getdX(a)) > 1e-10 * C.block(0).getdX(a)) {
      err_ = "level " + std::to_string(l) + " must halve the spacing of level " + std::to_string(l - 1);
      return AMSS_ERR_ARG;
    }
  for (int f = 0; f < F.npatch(); ++f) {
    const Block& fb = F.block(f);
    const auto& wf = buf_[l][f];
    double hF[3], hmin[3];
    for (int a = 0; a < 3; ++a) {
      hF[a] = fb.getdX(a);
      hmin[a] = hF[a];
      // the fine patch must sit on the coarse cells (2:1 cell-centred nesting)
      const double q = (fb.bbox[a] - C.block(0).bbox[a]) / C.block(0).getdX(a);
      if (std::fabs(q - std::round(q)) > 1e-6) {
        err_ = "a fine patch of level " + std::to_string(l) + " is not aligned with the coarse cells";
        return AMSS_ERR_ARG;
      }
    }
    PBox owned2, inner;
    for (int a = 0; a < 3; ++a) {
      owned2.lo[a] = inner.lo[a] = fb.bbox[a] + wf[a] * hF[a];
      owned2.hi[a] = inner.hi[a] = fb.bbox[3 + a] - wf[3 + a] * hF[a];
    }
    // the buffer zone as disjoint slabs: z faces whole, then y, then x
    std::vector<PBox> bufs;
    PBox rest;
    for (int a = 0; a < 3; ++a) {
      rest.lo[a] = fb.bbox[a];
      rest.hi[a] = fb.bbox[3 + a];
    }
    for (int a = 2; a >= 0; --a) {
      if (wf[a] > 0) {
        PBox s = rest;
        s.hi[a] = inner.lo[a];
        bufs.push_back(s);
      }
      if (wf[3 + a] > 0) {
        PBox s = rest;
        s.lo[a] = inner.hi[a];
        bufs.push_back(s);
      }
      rest.lo[a] = inner.lo[a];
      rest.hi[a] = inner.hi[a];
    }
    long want = 0, got = 0;
    for (const PBox& b : bufs) want += cells(b, hF);
    for (int c = 0; c < C.npatch(); ++c) {
      const Block& cb = C.block(c);
      const auto& wc = buf_[l - 1][c];
      PBox complete, owned4;
      for (int a = 0; a < 3; ++a) {
        const double h = cb.getdX(a);
        complete.lo[a] = cb.bbox[a];
        complete.hi[a] = cb.bbox[3 + a];
        owned4.lo[a] = on_plane(a, symmetry_, cb.bbox[a], h) ? cb.bbox[a] : cb.bbox[a] + (wc[a] + 3) * h;
        owned4.hi[a] = cb.bbox[3 + a] - (wc[3 + a] + 3) * h;
      }
      const int extc[3] = {cb.shape[0], cb.shape[1], cb.shape[2]};
      const int extf[3] = {fb.shape[0], fb.shape[1], fb.shape[2]};
      PBox r;
      if (overlap(owned2, complete, hmin, r)) {
        Xfer x{c, f, {}};
        if (xfer_plan(cb.bbox, cb.bbox + 3, extc, fb.bbox, fb.bbox + 3, extf, r.lo, r.hi, false, x.P)) {
          err_ = "a restriction region of level " + std::to_string(l) + " does not fit its arrays";
          return AMSS_ERR_ARG;
        }
        restrict_[l].push_back(x);
      }
      for (const PBox& b : bufs)
        if (overlap(b, owned4, hmin, r)) {
          Xfer x{c, f, {}};
          if (xfer_plan(cb.bbox, cb.bbox + 3, extc, fb.bbox, fb.bbox + 3, extf, r.lo, r.hi, true, x.P)) {
            err_ = "a prolongation region of level " + std::to_string(l) + " does not fit its arrays";
            return AMSS_ERR_ARG;
          }
          prolong_[l].push_back(x);
          got += cells(r, hF);
        }
    }
    if (got != want) {
      err_ = "the buffer zone of a patch of level " + std::to_string(l) +
             " is not inside the coarse level's owned cells (" + std::to_string(got) + " of " + std::to_string(want) +
             " cells)";
      return AMSS_ERR_ARG;
    }
  }
  (void)hc;
  (void)hf;
  return 0;
}

int AmrHierarchy::step(double dT, int nsteps, int mode, int want_res) {
  if (!ok()) return AMSS_ERR_ARG;
  for (int n = 0; n < nsteps; ++n) {
    const int r = recursive_step(0, dT, mode, want_res);
    if (r) return r;
  }
  return 0;
}

// RecursiveStep (bssn_class.C:2323-2403), every finer level refined in time
// (trfls = 0): level l takes 2^l steps of dT / 2^l per coarse step, and after
// each of them the transfers with level l - 1.
int AmrHierarchy::recursive_step(int l, double dT, int mode, int want_res) {
  const double dTl = dT * std::pow(0.5, l);
  const int nit = l == 0 ? 1 : 2;
  for (int i = 0; i < nit; ++i) {
    const bool same_time = i == nit - 1;  // YN: level l catches up with level l - 1
    const int r = lev_[l]->step(dTl, 1, mode, want_res);
    if (r) {
      err_ = "level " + std::to_string(l) + ": " + lev_[l]->err();
      return r;
    }
    lt_[l] += dTl;
    if (l + 1 < nlev()) {
      if (int q = recursive_step(l + 1, dT, mode, want_res)) return q;
    } else {
      phys_time_ += dT * std::pow(0.5, l);
    }
    // three time levels of the coarse level exist after its first step
    // (fgt(PhysTime - dT_lev, StartTime, dT_lev / 2))
    const bool three = phys_time_ - dTl > start_time_ + dTl / 2;
    if (int q = restrict_prolong(l, same_time, three)) return q;
  }
  return 0;
}

#define ACHK(call)                                                          \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      err_ = std::string(#call ": ") + hipGetErrorString(e_);               \
      return AMSS_ERR_HIP;                                                  \
    }                                                                       \
  } while (0)

// RestrictProlong(lev, YN, BB, StateList, OldStateList, SynchList_cor)
// (bssn_class.C:5724-5904, RPB 0, MIXOUTB 0).  Between the fine level's two
// steps the coarse level is one fine step ahead: its state is interpolated
// in time into its SynchList_pre (average, or average2 over three levels),
// the fine state restricted into that copy, and the fine buffer zone
// prolonged from it; when both levels reach the same time the fine state is
// restricted into the coarse StateList and the buffer prolonged from it.
int AmrHierarchy::restrict_prolong(int l, bool same_time, bool three) {
  if (l == 0) return 0;
  BssnLevel& C = *lev_[l - 1];
  BssnLevel& F = *lev_[l];
  const int target = same_time ? AMSS_LIST_STATE : AMSS_LIST_PRE;
  if (!same_time) {
    for (int c = 0; c < C.npatch(); ++c) {
      FieldsIn f1, f2, f3;
      FieldsOut out;
      for (int v = 0; v < NFIELD; ++v) {
        f1.p[v] = C.field(c, AMSS_LIST_STATE, v);  // t + dt
        f2.p[v] = C.field(c, AMSS_LIST_OLD, v);    // t
        f3.p[v] = C.field(c, AMSS_LIST_COR, v);    // t - dt (after the rotation)
        out.p[v] = C.field(c, AMSS_LIST_PRE, v);
      }
      if (three)  // average2 (fmisc.f90:1569-1592): 3/8, 3/4, -1/8 at t + dt/2
        launch_time_interp(f1, f2, f3, out, C.npts(c), 3.0 / 8.0, 3.0 / 4.0, -1.0 / 8.0, true, stream_);
      else  // average (fmisc.f90:1532-1544)
        launch_time_interp(f1, f2, f3, out, C.npts(c), 0.5, 0.5, 0.0, false, stream_);
    }
  }
  for (const Xfer& x : restrict_[l])
    for (int v = 0; v < NFIELD; ++v) {
      const double S[3] = {(double)parity(v, 0), (double)parity(v, 1), (double)parity(v, 2)};
      launch_restrict3(F.field(x.f, AMSS_LIST_STATE, v), C.field(x.c, target, v), x.P, S, stream_);
    }
  ACHK(hipGetLastError());
  ACHK(hipStreamSynchronize(stream_));
  if (int r = C.sync_list(target)) return r;  // Parallel::Sync on the coarse level
  ACHK(hipStreamSynchronize(C.stream()));
  for (const Xfer& x : prolong_[l])
    for (int v = 0; v < NFIELD; ++v) {
      const double S[3] = {(double)parity(v, 0), (double)parity(v, 1), (double)parity(v, 2)};
      launch_prolong3(C.field(x.c, target, v), F.field(x.f, AMSS_LIST_STATE, v), x.P, S, stream_);
    }
  ACHK(hipGetLastError());
  ACHK(hipStreamSynchronize(stream_));
  if (int r = F.sync_list(AMSS_LIST_STATE)) return r;  // Parallel::Sync on the fine level
  ACHK(hipStreamSynchronize(F.stream()));
  return 0;
}

}  // namespace amss
