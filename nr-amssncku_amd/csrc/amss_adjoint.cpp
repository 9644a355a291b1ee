// amss_adjoint.cpp -- work arrays, stage sequencing and the inter-patch
// exchange of the level step adjoint.
#include "amss_adjoint.h"

#include <algorithm>
#include <array>
#include <cmath>

#include "../../include/amss_bssn.h"
#include "bssn_adjoint.h"
#include "bssn_adjoint_launch.h"

namespace amss {

namespace {
// work arrays of a Block (in grid functions of N doubles)
constexpr int kY0 = 0;                   // enforced U^n (24)
constexpr int kY = kY0 + NFIELD;         // Y1..Y3, the inputs of sub-steps 1-3 (3 x 24)
constexpr int kZ = kY + 3 * NFIELD;      // z0..z3, sub-step outputs before bound/enforce (4 x 24)
constexpr int kStage = kZ + 4 * NFIELD;  // grid functions of one step's stage data
constexpr int kK = kZ + 4 * NFIELD;      // k of the current sub-step (24)
constexpr int kAcc = kK + NFIELD;        // RK4 accumulator (24)
constexpr int kLY = kAcc + NFIELD;       // U^{n+1} scratch (forward); dJ/dY of the sub-step being reversed
constexpr int kLY0 = kLY + NFIELD;       // dJ/dY0 (24)
constexpr int kFeat = kLY0 + NFIELD;     // stencil values of the reverse algebra (kAdjInputs)
constexpr int kFields = kFeat + kAdjInputs;
// padded arrays (3 zero planes before and after each, never written outside
// the interior, so the gather reads them at any stencil offset unchecked)
constexpr int kBars = 0;                   // feature adjoints (kAdjFeatures)
constexpr int kLK = kBars + kAdjFeatures;  // dJ/dk0..k3 (4 x 24)
constexpr int kPadFields = kLK + 4 * NFIELD;
}  // namespace

AdjointBlock::AdjointBlock(const int ex[3], const Coef& C, double tinny) : C_(C), tinny_(tinny) {
  for (int a = 0; a < 3; ++a) {
    if (ex[a] < 7) {
      err_ = "every axis needs at least 7 points";
      return;
    }
    box_.n[a] = ex[a];
    box_.lo[a] = 3;
    box_.hi[a] = ex[a] - 3;
  }
  box_.sy = ex[0];
  box_.sz = (long)ex[0] * ex[1];
  N_ = (long)ex[0] * ex[1] * ex[2];
  hipError_t e = hipMalloc(&buf_, (size_t)kFields * N_ * sizeof(double));
  if (e != hipSuccess) {
    buf_ = nullptr;
    err_ = std::string("hipMalloc of the adjoint work arrays: ") + hipGetErrorString(e);
    return;
  }
  pad_ = 3 * box_.sz;
  NP_ = N_ + 2 * pad_;
  e = hipMalloc(&padbuf_, (size_t)kPadFields * NP_ * sizeof(double));
  if (e == hipSuccess) e = hipMemset(padbuf_, 0, (size_t)kPadFields * NP_ * sizeof(double));
  if (e != hipSuccess) {
    if (padbuf_) (void)hipFree(padbuf_);
    padbuf_ = nullptr;
    err_ = std::string("hipMalloc of the adjoint work arrays: ") + hipGetErrorString(e);
  }
}

AdjointBlock::~AdjointBlock() {
  if (buf_) (void)hipFree(buf_);
  if (padbuf_) (void)hipFree(padbuf_);
  if (tape_) (void)hipFree(tape_);
}

int AdjointBlock::reserve(int slots) {
  if (slots <= slots_) return 0;
  if (tape_) (void)hipFree(tape_);
  tape_ = nullptr;
  slots_ = 0;
  const hipError_t e = hipMalloc(&tape_, (size_t)slots * kStage * N_ * sizeof(double));
  if (e != hipSuccess) {
    tape_ = nullptr;
    err_ = std::string("hipMalloc of the adjoint tape: ") + hipGetErrorString(e);
    return AMSS_ERR_HIP;
  }
  slots_ = slots;
  return 0;
}

// grid function `first` (stage data from the current slot)
double* AdjointBlock::base_of(int first, int slot) const {
  if (first < kStage && slot >= 0) return tape_ + ((long)slot * kStage + first) * N_;
  return buf_ + (long)first * N_;
}
double* AdjointBlock::padded(int first) const { return padbuf_ + (long)first * NP_ + pad_; }
FieldsIn AdjointBlock::pin(int first) const {
  FieldsIn f;
  for (int v = 0; v < NFIELD; ++v) f.p[v] = padded(first + v);
  return f;
}
FieldsOut AdjointBlock::pout(int first) const {
  FieldsOut f;
  for (int v = 0; v < NFIELD; ++v) f.p[v] = padded(first + v);
  return f;
}
FieldsIn AdjointBlock::fin(int first) const {
  FieldsIn f;
  const double* b = base_of(first, slot_);
  for (int v = 0; v < NFIELD; ++v) f.p[v] = b + (long)v * N_;
  return f;
}
FieldsOut AdjointBlock::fout(int first) const {
  FieldsOut f;
  double* b = base_of(first, slot_);
  for (int v = 0; v < NFIELD; ++v) f.p[v] = b + (long)v * N_;
  return f;
}
static FieldsIn as_in(const double* const* p) {
  FieldsIn f;
  for (int v = 0; v < NFIELD; ++v) f.p[v] = p[v];
  return f;
}
static FieldsOut as_out(double* const* p) {
  FieldsOut f;
  for (int v = 0; v < NFIELD; ++v) f.p[v] = p[v];
  return f;
}

PatchDesc AdjointBlock::desc(int which, int slot) const {
  PatchDesc d;
  d.slab = base_of(which, slot);
  d.n = N_;
  d.sy = box_.n[0];
  d.sz = box_.n[0] * box_.n[1];
  for (int a = 0; a < 3; ++a) d.shape[a] = box_.n[a];
  return d;
}
double* AdjointBlock::stage_output(int st) const { return base_of(st < 3 ? kY + st * NFIELD : kLY, slot_); }
double* AdjointBlock::lY() const { return buf_ + (long)kLY * N_; }

void AdjointBlock::fwd_open(const double* const* U, hipStream_t s) {
  adj_launch_copy_enforce(as_in(U), fout(kY0), N_, s);
}

void AdjointBlock::fwd_stage(int st, double c, double* const* Unext, hipStream_t s) {
  const FieldsIn in = st == 0 ? fin(kY0) : fin(kY + (st - 1) * NFIELD);
  adj_launch_rhs_k(in, fout(kK), box_, C_, s);
  const FieldsOut yn = st < 3 ? fout(kY + st * NFIELD) : (Unext ? as_out(Unext) : fout(kLY));
  adj_launch_fwd_stage(fin(kY0), fin(kK), fout(kAcc), fout(kZ + st * NFIELD), yn, box_, st, c, tinny_, s);
}

void AdjointBlock::rev_final(const double* const* lam, double c, hipStream_t s) {
  const FieldsOut lk[4] = {pout(kLK), pout(kLK + NFIELD), pout(kLK + 2 * NFIELD), pout(kLK + 3 * NFIELD)};
  adj_launch_final(as_in(lam), fin(kZ + 3 * NFIELD), fin(kY0), fout(kLY0), lk, box_, c, tinny_, s);
}

void AdjointBlock::rev_rhs(int st, hipStream_t s) {
  const FieldsIn Ys = st == 0 ? fin(kY0) : fin(kY + (st - 1) * NFIELD);
  const FieldsIn lks = pin(kLK + st * NFIELD);
  adj_launch_point(Ys, lks, padded(kBars), NP_, box_, C_, s, buf_ + (long)kFeat * N_);
  if (st == 0)
    adj_launch_gather(Ys, lks, true, padded(kBars), NP_, fout(kLY0), box_, C_, true, s);
  else
    adj_launch_gather(Ys, lks, true, padded(kBars), NP_, fout(kLY), box_, C_, false, s);
}

void AdjointBlock::rev_stage(int st, double c, hipStream_t s) {
  adj_launch_stage(fin(kLY), fin(kZ + (st - 1) * NFIELD), fin(kY0), fout(kLY0), pout(kLK + (st - 1) * NFIELD),
                   box_, c, tinny_, s);
}

void AdjointBlock::rev_open(const double* const* U, double* const* lam, hipStream_t s) {
  adj_launch_open(as_in(U), fin(kLY0), as_out(lam), N_, s);
}

void AdjointBlock::rhs_vjp(const double* const* U, const double* const* lam, double* const* out, hipStream_t s) {
  adj_launch_point(as_in(U), as_in(lam), padded(kBars), NP_, box_, C_, s, buf_ + (long)kFeat * N_);
  adj_launch_gather(as_in(U), as_in(lam), false, padded(kBars), NP_, as_out(out), box_, C_, false, s);
}

// ---------------------------------------------------------------------------
static Coef coef_of(const double h[3], double eps) {
  Coef C;
  for (int a = 0; a < 3; ++a) {
    C.h[a] = h[a];
    C.d1[a] = 1.0 / 12.0 / h[a];
    C.d1lo[a] = 1.0 / 2.0 / h[a];
    C.d2[a] = (1.0 / 12.0) / (h[a] * h[a]);
    C.d2lo[a] = 1.0 / (h[a] * h[a]);
    C.ko[a] = (eps > 0 ? eps : 0.0) / 64.0 / h[a];
  }
  const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
  for (int q = 0; q < 3; ++q) {
    C.dm[q] = (1.0 / 144.0) / (h[pa[q]] * h[pb[q]]);
    C.dmlo[q] = 0.25 / (h[pa[q]] * h[pb[q]]);
  }
  return C;
}

LevelAdjoint::LevelAdjoint(int npatch, const int* shapes, const double* bboxes, double eps, double chitiny,
                           const double* hgiven) {
  const int gw = 3;
  std::vector<std::array<double, 3>> h(npatch);
  for (int p = 0; p < npatch; ++p)
    for (int a = 0; a < 3; ++a) {
      h[p][a] = hgiven ? hgiven[a] : (bboxes[6 * p + 3 + a] - bboxes[6 * p + a]) / shapes[3 * p + a];
      if (!(h[p][a] > 0.0)) {
        err_ = "bad patch box";
        return;
      }
      if (std::fabs(h[p][a] - h[0][a]) > 1e-12 * h[0][a]) {
        err_ = "all patches of a level must share one grid spacing";
        return;
      }
    }
  for (int p = 0; p < npatch; ++p) {
    auto* b = new AdjointBlock(shapes + 3 * p, coef_of(h[p].data(), eps), chitiny - 1.0);
    blocks_.push_back(b);
    if (!b->ok()) {
      err_ = b->err();
      return;
    }
  }
  // exchange boxes: the ghost points of patch A inside the interior of patch B
  double gmin[3] = {1e300, 1e300, 1e300};
  for (int p = 0; p < npatch; ++p)
    for (int a = 0; a < 3; ++a) gmin[a] = std::min(gmin[a], bboxes[6 * p + a]);
  std::vector<std::array<int, 3>> org(npatch);
  for (int p = 0; p < npatch; ++p)
    for (int a = 0; a < 3; ++a) org[p][a] = (int)std::lround((bboxes[6 * p + a] - gmin[a]) / h[0][a]);
  for (int A = 0; A < npatch; ++A)
    for (int B = 0; B < npatch; ++B) {
      if (A == B) continue;
      SegRef r;
      bool empty = false;
      for (int a = 0; a < 3; ++a) {
        const int alo = org[A][a], ahi = org[A][a] + shapes[3 * A + a];
        const int blo = org[B][a] + gw, bhi = org[B][a] + shapes[3 * B + a] - gw;
        const int lo = std::max(alo, blo), hi = std::min(ahi, bhi);
        if (hi <= lo) {
          empty = true;
          break;
        }
        r.dlo[a] = lo - org[A][a];
        r.slo[a] = lo - org[B][a];
        r.ext[a] = hi - lo;
      }
      if (empty) continue;
      r.src = B;
      r.dst = A;
      hseg_.push_back(r);
      seg_max_ = std::max(seg_max_, (long)r.ext[0] * r.ext[1] * r.ext[2]);
    }
  if (hseg_.size() > 65535) {
    err_ = "too many exchange boxes";
    return;
  }
  auto upload = [&](const void* src, size_t bytes, void** dst) -> bool {
    if (hipMalloc(dst, bytes) != hipSuccess || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
      err_ = "device tables of the level adjoint";
      return false;
    }
    return true;
  };
  if (!hseg_.empty() && !upload(hseg_.data(), hseg_.size() * sizeof(SegRef), (void**)&dseg_)) return;
  // PatchDesc tables of the exchanged runs: Y1, Y2, Y3 and the scratch / dJ/dY run
  for (int which : {kY, kY + NFIELD, kY + 2 * NFIELD, kLY}) {
    std::vector<PatchDesc> t;
    for (AdjointBlock* b : blocks_) t.push_back(b->desc(which, -1));
    PatchDesc* d = nullptr;
    if (!upload(t.data(), t.size() * sizeof(PatchDesc), (void**)&d)) return;
    dtab_.push_back(d);
  }
}

int LevelAdjoint::reserve(int slots) {
  const int have = ((int)dtab_.size() - 4) / 3;
  if (slots <= have) return 0;
  for (AdjointBlock* b : blocks_)
    if (int r = b->reserve(slots)) {
      err_ = b->err();
      return r;
    }
  for (size_t i = 4; i < dtab_.size(); ++i) (void)hipFree(dtab_[i]);
  dtab_.resize(4);
  for (int sl = 0; sl < slots; ++sl)
    for (int which : {kY, kY + NFIELD, kY + 2 * NFIELD}) {
      std::vector<PatchDesc> t;
      for (AdjointBlock* b : blocks_) t.push_back(b->desc(which, sl));
      PatchDesc* d = nullptr;
      if (hipMalloc(&d, t.size() * sizeof(PatchDesc)) != hipSuccess ||
          hipMemcpy(d, t.data(), t.size() * sizeof(PatchDesc), hipMemcpyHostToDevice) != hipSuccess) {
        err_ = "device tables of the adjoint tape";
        return AMSS_ERR_HIP;
      }
      dtab_.push_back(d);
    }
  return 0;
}

PatchDesc* LevelAdjoint::table(int which, int slot) {
  if (which == 3 || slot < 0) return dtab_[which];
  return dtab_[4 + 3 * slot + which];
}

LevelAdjoint::~LevelAdjoint() {
  (void)hipDeviceSynchronize();
  for (AdjointBlock* b : blocks_) delete b;
  if (dseg_) (void)hipFree(dseg_);
  for (PatchDesc* d : dtab_) (void)hipFree(d);
}

size_t LevelAdjoint::bytes() const {
  size_t n = 0;
  const int slots = ((int)dtab_.size() - 4) / 3;
  for (AdjointBlock* b : blocks_)
    n += ((size_t)(kFields + (size_t)slots * kStage) * b->npts() + (size_t)kPadFields * b->padded_npts()) *
         sizeof(double);
  return n;
}

void LevelAdjoint::exchange(int tab, int slot, hipStream_t s) {
  if (!hseg_.empty()) launch_seg_boxes(table(tab, slot), dseg_, (int)hseg_.size(), (int)seg_max_, 0, s);
}
void LevelAdjoint::exchange_adjoint(hipStream_t s) {
  if (!hseg_.empty()) launch_seg_boxes_adjoint(dtab_[3], dseg_, (int)hseg_.size(), (int)seg_max_, s);
}

int LevelAdjoint::forward(double dT, const double* const* U, double* const* Unext, hipStream_t s, int slot) {
  if (slot >= ((int)dtab_.size() - 4) / 3) {
    err_ = "tape slot not reserved";
    return AMSS_ERR_ARG;
  }
  const double c[4] = {0.5 * dT, 0.5 * dT, dT, dT / 6.0};
  const int np = npatch();
  for (AdjointBlock* b : blocks_) b->use_slot(slot);
  for (int p = 0; p < np; ++p) blocks_[p]->fwd_open(U + p * NFIELD, s);
  for (int st = 0; st < 4; ++st) {
    for (int p = 0; p < np; ++p) blocks_[p]->fwd_stage(st, c[st], nullptr, s);
    exchange(st, slot, s);  // tables 0..2 = Y1..Y3, 3 = the U^{n+1} scratch
  }
  if (Unext)
    for (int p = 0; p < np; ++p) {
      const long N = blocks_[p]->npts();
      FieldsIn src;
      for (int v = 0; v < NFIELD; ++v) src.p[v] = blocks_[p]->stage_output(3) + v * N;
      adj_launch_copy_fields(src, as_out(Unext + p * NFIELD), N, s);
    }
  if (hipGetLastError() != hipSuccess) {
    err_ = "level adjoint forward: kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

int LevelAdjoint::backward(double dT, const double* const* U, double* const* lam, hipStream_t s, int slot) {
  if (slot < 0) {
    if (int r = forward(dT, U, nullptr, s)) return r;
  } else if (slot >= ((int)dtab_.size() - 4) / 3) {
    err_ = "tape slot not reserved";
    return AMSS_ERR_ARG;
  }
  const double c[4] = {0.5 * dT, 0.5 * dT, dT, dT / 6.0};
  const int np = npatch();
  for (AdjointBlock* b : blocks_) b->use_slot(slot);
  // dJ/dU^{n+1} into each Block's dJ/dY run, then the final exchange reversed
  for (int p = 0; p < np; ++p) {
    const long N = blocks_[p]->npts();
    FieldsOut dst;
    for (int v = 0; v < NFIELD; ++v) dst.p[v] = blocks_[p]->lY() + v * N;
    adj_launch_copy_fields(as_in(lam + p * NFIELD), dst, N, s);
  }
  exchange_adjoint(s);
  for (int p = 0; p < np; ++p) {
    double* ly = blocks_[p]->lY();
    const long N = blocks_[p]->npts();
    const double* ptr[NFIELD];
    for (int v = 0; v < NFIELD; ++v) ptr[v] = ly + v * N;
    blocks_[p]->rev_final(ptr, c[3], s);
  }
  for (int st = 3; st >= 1; --st) {
    for (int p = 0; p < np; ++p) blocks_[p]->rev_rhs(st, s);  // dJ/dY_st into the dJ/dY run
    exchange_adjoint(s);  // Y_st was exchanged after sub-step st-1
    for (int p = 0; p < np; ++p) blocks_[p]->rev_stage(st, c[st - 1], s);
  }
  for (int p = 0; p < np; ++p) blocks_[p]->rev_rhs(0, s);
  for (int p = 0; p < np; ++p) blocks_[p]->rev_open(U + p * NFIELD, lam + p * NFIELD, s);
  if (hipGetLastError() != hipSuccess) {
    err_ = "level adjoint backward: kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

int LevelAdjoint::rhs_vjp(const double* const* U, const double* const* lam, double* const* out, hipStream_t s) {
  for (AdjointBlock* b : blocks_) b->use_slot(-1);
  for (int p = 0; p < npatch(); ++p)
    blocks_[p]->rhs_vjp(U + p * NFIELD, lam + p * NFIELD, out + p * NFIELD, s);
  if (hipGetLastError() != hipSuccess) {
    err_ = "adjoint rhs: kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

}  // namespace amss
