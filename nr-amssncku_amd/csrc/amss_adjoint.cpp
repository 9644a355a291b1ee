// amss_adjoint.cpp -- the step adjoint's workspace and stage sequencing.
#include "amss_adjoint.h"

#include "../../include/amss_bssn.h"
#include "bssn_adjoint.h"
#include "bssn_adjoint_launch.h"

namespace amss {

namespace {
// work arrays (in fields of N doubles)
constexpr int kY0 = 0;                            // enforced U^n (24)
constexpr int kY = kY0 + NFIELD;                  // Y1..Y3, the inputs of stages 1-3 (3 x 24)
constexpr int kZ = kY + 3 * NFIELD;               // z0..z3, stage outputs before bound/enforce (4 x 24)
constexpr int kK = kZ + 4 * NFIELD;               // k of the current stage (24)
constexpr int kAcc = kK + NFIELD;                 // RK4 accumulator (24)
constexpr int kBars = kAcc + NFIELD;              // feature bars
constexpr int kLY = kBars + kAdjFeatures;         // dJ/dY of the stage being reversed (24)
constexpr int kLY0 = kLY + NFIELD;                // dJ/dY0 (24)
constexpr int kLK = kLY0 + NFIELD;                // dJ/dk0..k3 (4 x 24)
constexpr int kFields = kLK + 4 * NFIELD;
}  // namespace

AdjointStep::AdjointStep(const int ex[3], const Coef& C, double chitiny) : C_(C), tinny_(chitiny - 1.0) {
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
  const hipError_t e = hipMalloc(&buf_, (size_t)kFields * N_ * sizeof(double));
  if (e != hipSuccess) {
    buf_ = nullptr;
    err_ = std::string("hipMalloc of the adjoint workspace: ") + hipGetErrorString(e);
  }
}

AdjointStep::~AdjointStep() {
  if (buf_) (void)hipFree(buf_);
}

FieldsIn AdjointStep::fin(int first) const {
  FieldsIn f;
  for (int v = 0; v < NFIELD; ++v) f.p[v] = buf_ + (long)(first + v) * N_;
  return f;
}
FieldsOut AdjointStep::fout(int first) const {
  FieldsOut f;
  for (int v = 0; v < NFIELD; ++v) f.p[v] = buf_ + (long)(first + v) * N_;
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

int AdjointStep::forward(double dT, const double* const* U, double* const* Unext, hipStream_t s) {
  const double c[4] = {0.5 * dT, 0.5 * dT, dT, dT / 6.0};
  adj_launch_copy_enforce(as_in(U), fout(kY0), N_, s);
  for (int st = 0; st < 4; ++st) {
    const FieldsIn in = st == 0 ? fin(kY0) : fin(kY + (st - 1) * NFIELD);
    adj_launch_rhs_k(in, fout(kK), box_, C_, s);
    const FieldsOut yn = st < 3 ? fout(kY + st * NFIELD) : (Unext ? as_out(Unext) : fout(kLY));
    adj_launch_fwd_stage(fin(kY0), fin(kK), fout(kAcc), fout(kZ + st * NFIELD), yn, box_, st, c[st], tinny_, s);
  }
  if (hipGetLastError() != hipSuccess) {
    err_ = "adjoint forward: kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

int AdjointStep::backward(double dT, const double* const* U, double* const* lam, hipStream_t s) {
  if (int r = forward(dT, U, nullptr, s)) return r;
  const double c[4] = {0.5 * dT, 0.5 * dT, dT, dT / 6.0};
  const FieldsOut lk[4] = {fout(kLK), fout(kLK + NFIELD), fout(kLK + 2 * NFIELD), fout(kLK + 3 * NFIELD)};
  adj_launch_final(as_in(lam), fin(kZ + 3 * NFIELD), fin(kY0), fout(kLY0), lk, box_, c[3], tinny_, s);
  double* bars = buf_ + (long)kBars * N_;
  for (int st = 3; st >= 1; --st) {
    const FieldsIn Ys = fin(kY + (st - 1) * NFIELD);
    const FieldsIn lks = fin(kLK + st * NFIELD);
    adj_launch_point(Ys, lks, bars, box_, C_, s);
    adj_launch_gather(Ys, lks, bars, fout(kLY), box_, C_, false, s);
    adj_launch_stage(fin(kLY), fin(kZ + (st - 1) * NFIELD), fin(kY0), fout(kLY0), lk[st - 1], box_, c[st - 1],
                     tinny_, s);
  }
  adj_launch_point(fin(kY0), fin(kLK), bars, box_, C_, s);
  adj_launch_gather(fin(kY0), fin(kLK), bars, fout(kLY0), box_, C_, true, s);
  adj_launch_open(as_in(U), fin(kLY0), as_out(lam), N_, s);
  if (hipGetLastError() != hipSuccess) {
    err_ = "adjoint backward: kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

int AdjointStep::rhs_vjp(const double* const* U, const double* const* lam, double* const* out, hipStream_t s) {
  double* bars = buf_ + (long)kBars * N_;
  adj_launch_point(as_in(U), as_in(lam), bars, box_, C_, s);
  adj_launch_gather(as_in(U), as_in(lam), bars, as_out(out), box_, C_, false, s);
  if (hipGetLastError() != hipSuccess) {
    err_ = "adjoint rhs: kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

}  // namespace amss
