// amss_level.cpp -- Blocks, lists and the RK4 Step of one refinement level
// on one GPU, in the order of bssn_class::Step (bssn_class.C:2981-3901).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/amss_bssn.h"
#include "amss_host.h"
#include "bssn_edge_tiles.h"

namespace amss {

static const char* kNames[NFIELD] = {"phi",  "trK",  "gxx", "gxy",  "gxz", "gyy", "gyz", "gzz",
                                     "Axx",  "Axy",  "Axz", "Ayy",  "Ayz", "Azz", "Gmx", "Gmy",
                                     "Gmz",  "Lap",  "Sfx", "Sfy",  "Sfz", "dtSfx", "dtSfy", "dtSfz"};
static const char* kListSfx[5] = {"0", "o", "", "1", "_rhs"};
static const char* kConsNames[7] = {"Cons_Ham", "Cons_Px", "Cons_Py", "Cons_Pz", "Cons_Gx", "Cons_Gy", "Cons_Gz"};

namespace {

// Half-open box of array indices.
struct IBox {
  int lo[3], hi[3];
  bool intersects(const IBox& o) const {
    for (int a = 0; a < 3; ++a)
      if (std::max(lo[a], o.lo[a]) >= std::min(hi[a], o.hi[a])) return false;
    return true;
  }
};

// The points of the disjoint boxes `keep` that no box of `cut` contains, as
// disjoint boxes: the cell grid of all box faces, marked, then merged greedily
// along x, y and z.
void subtract_boxes(const std::vector<IBox>& keep, const std::vector<IBox>& cut, std::vector<IBox>& out) {
  out.clear();
  if (keep.empty()) return;
  std::vector<int> c[3];
  for (int a = 0; a < 3; ++a) {
    for (const IBox& b : keep) {
      c[a].push_back(b.lo[a]);
      c[a].push_back(b.hi[a]);
    }
    for (const IBox& b : cut) {
      c[a].push_back(b.lo[a]);
      c[a].push_back(b.hi[a]);
    }
    std::sort(c[a].begin(), c[a].end());
    c[a].erase(std::unique(c[a].begin(), c[a].end()), c[a].end());
  }
  const int n0 = (int)c[0].size() - 1, n1 = (int)c[1].size() - 1, n2 = (int)c[2].size() - 1;
  auto index = [&](const std::vector<int>& v, int x) {
    return (int)(std::lower_bound(v.begin(), v.end(), x) - v.begin());
  };
  std::vector<char> on((size_t)n0 * n1 * n2, 0);
  auto at = [&](int i, int j, int k) -> char& { return on[i + (size_t)n0 * (j + (size_t)n1 * k)]; };
  auto mark = [&](const IBox& b, char val) {
    const int i0 = index(c[0], b.lo[0]), i1 = index(c[0], b.hi[0]);
    const int j0 = index(c[1], b.lo[1]), j1 = index(c[1], b.hi[1]);
    const int k0 = index(c[2], b.lo[2]), k1 = index(c[2], b.hi[2]);
    for (int k = k0; k < k1; ++k)
      for (int j = j0; j < j1; ++j)
        for (int i = i0; i < i1; ++i) at(i, j, k) = val;
  };
  for (const IBox& b : keep) mark(b, 1);
  for (const IBox& b : cut) mark(b, 0);
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n0; ++i) {
        if (!at(i, j, k)) continue;
        int i1 = i + 1;
        while (i1 < n0 && at(i1, j, k)) ++i1;
        auto row = [&](int jj, int kk) {
          for (int x = i; x < i1; ++x)
            if (!at(x, jj, kk)) return false;
          return true;
        };
        int j1 = j + 1;
        while (j1 < n1 && row(j1, k)) ++j1;
        int k1 = k + 1;
        while (k1 < n2) {
          bool ok = true;
          for (int y = j; y < j1 && ok; ++y) ok = row(y, k1);
          if (!ok) break;
          ++k1;
        }
        for (int z = k; z < k1; ++z)
          for (int y = j; y < j1; ++y)
            for (int x = i; x < i1; ++x) at(x, y, z) = 0;
        out.push_back(IBox{{c[0][i], c[1][j], c[2][k]}, {c[0][i1], c[1][j1], c[2][k1]}});
      }
}

}  // namespace

void Block::swapList(MyList<var>* a, MyList<var>* b) {
  while (a && b) {
    std::swap(fgfs[a->data->sgfn], fgfs[b->data->sgfn]);
    a = a->next;
    b = b->next;
  }
}

#define LCHK(call)                                                                   \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      err_ = std::string(#call ": ") + hipGetErrorString(e_);                        \
      return;                                                                        \
    }                                                                                \
  } while (0)
#define RCHK(call)                                                                   \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      err_ = std::string(#call ": ") + hipGetErrorString(e_);                        \
      return AMSS_ERR_HIP;                                                           \
    }                                                                                \
  } while (0)

BssnLevel::BssnLevel(int npatch, const int* shapes, const double* bboxes, int gauge, double eps, double chitiny,
                     int bc, int gw, int symmetry)
    : gauge_(gauge), bc_(bc), gw_(gw), symmetry_(symmetry), eps_(eps), chitiny_(chitiny) {
  if (gauge != 0 && gauge != 2) { err_ = "gauge must be 0 or 2"; return; }
  if (bc < 0 || bc > 3) { err_ = "unknown boundary mode"; return; }
  if (gw < 0 || gw > 8) { err_ = "ghost width out of range"; return; }
  if (symmetry < 0 || symmetry > 2) { err_ = "Symmetry must be 0, 1 or 2"; return; }
  if (symmetry != 0 && bc == AMSS_BC_FROZEN_GHOSTS) {
    err_ = "a symmetry plane needs a boundary mode of the reference (copy layer, none or Sommerfeld)";
    return;
  }
  // grid-function registry: 5 time levels x 24 fields, then 7 residuals
  // (bssn_class.C:183-361), parity vectors as in the reference.
  int sg = 0;
  for (int l = 0; l < 5; ++l)
    for (int v = 0; v < NFIELD; ++v) {
      var* x = new var;
      std::snprintf(x->name, sizeof x->name, "%s%s", kNames[v], kListSfx[l]);
      x->sgfn = sg++;
      for (int a = 0; a < 3; ++a) x->SoA[a] = parity(v, a);
      x->propspeed = (v == CHI || v == TRK || v == LAP) ? std::sqrt(2.0) : 1.0;
      vars_.push_back(x);
    }
  for (int a = 0; a < 7; ++a) {
    var* x = new var;
    std::snprintf(x->name, sizeof x->name, "%s", kConsNames[a]);
    x->sgfn = sg++;
    x->SoA[0] = x->SoA[1] = x->SoA[2] = 1.0;
    x->propspeed = 1.0;
    vars_.push_back(x);
  }
  MyList<var>** lists[5] = {&StateList, &OldStateList, &SynchList_pre, &SynchList_cor, &RHSList};
  // registry order of time levels: 0 = state, 1 = old, 2 = pre, 3 = cor, 4 = rhs
  for (int l = 0; l < 5; ++l) {
    *lists[l] = new MyList<var>(vars_[l * NFIELD]);
    for (int v = 1; v < NFIELD; ++v) (*lists[l])->insert(vars_[l * NFIELD + v]);
  }
  ConsList = new MyList<var>(vars_[5 * NFIELD]);
  for (int a = 1; a < 7; ++a) ConsList->insert(vars_[5 * NFIELD + a]);

  double h0 = 0.0;
  for (int pi = 0; pi < npatch; ++pi) {
    Block* b = new Block;
    const int* sh = shapes + 3 * pi;
    const double* bb = bboxes + 6 * pi;
    for (int a = 0; a < 3; ++a) {
      b->shape[a] = sh[a];
      b->bbox[a] = bb[a];
      b->bbox[3 + a] = bb[3 + a];
      if (sh[a] < 7) { err_ = "every axis needs at least 7 points"; delete b; return; }
      // cell-centred coordinates (Block.C:61-64)
      const double h = (bb[3 + a] - bb[a]) / sh[a];
      b->X[a].resize(sh[a]);
      for (int j = 0; j < sh[a]; ++j) b->X[a][j] = bb[a] + (j + 0.5) * h;
      if (pi == 0 && a == 0) h0 = h;
      if (std::fabs(h - h0) > 1e-12 * h0) { err_ = "all patches of a level must share one grid spacing"; delete b; return; }
    }
    if (b->npts() >= (1L << 29)) { err_ = "more than 2^29 points in one block"; delete b; return; }
    b->fngfs = (int)vars_.size();
    const size_t n = b->npts();
    LCHK(hipMalloc(&b->slab, n * b->fngfs * sizeof(double)));
    LCHK(hipMemset(b->slab, 0, n * b->fngfs * sizeof(double)));
    b->fgfs.resize(b->fngfs);
    for (int i = 0; i < b->fngfs; ++i) b->fgfs[i] = b->slab + i * n;
    blocks_.push_back(b);
    Patch P;
    P.lev = 0;
    for (int a = 0; a < 3; ++a) {
      P.shape[a] = sh[a];
      P.bbox[a] = bb[a];
      P.bbox[3 + a] = bb[3 + a];
    }
    P.blb = P.ble = new MyList<Block>(b);
    patches_.push_back(P);
    const double hh[3] = {b->getdX(0), b->getdX(1), b->getdX(2)};
    coef_.push_back(Coef{});
    Coef& C = coef_.back();
    for (int a = 0; a < 3; ++a) {
      C.h[a] = hh[a];
      C.d1[a] = 1.0 / 12.0 / hh[a];
      C.d1lo[a] = 1.0 / 2.0 / hh[a];
      C.d2[a] = (1.0 / 12.0) / (hh[a] * hh[a]);
      C.d2lo[a] = 1.0 / (hh[a] * hh[a]);
      C.ko[a] = (eps > 0 ? eps : 0.0) / 64.0 / hh[a];
    }
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    for (int q = 0; q < 3; ++q) {
      C.dm[q] = (1.0 / 144.0) / (hh[pa[q]] * hh[pb[q]]);
      C.dmlo[q] = 0.25 / (hh[pa[q]] * hh[pb[q]]);
    }
    // Symmetry planes (diff_new.f90:975-980): z for Symmetry 1 and 2, x and y
    // for 2, where the first coordinate lies within one spacing of 0.  The
    // boundary kernel's stencils read mirror points there.
    Box eb;
    for (int a = 0; a < 3; ++a) {
      eb.n[a] = sh[a];
      eb.lo[a] = 3;
      eb.hi[a] = sh[a] - 3;
      const bool plane = (a == 2 ? symmetry > 0 : symmetry > 1) && std::fabs(b->X[a][0]) < hh[a];
      eb.mlo[a] = plane ? -3 : 0;
    }
    eb.sy = sh[0];
    eb.sz = (long)sh[0] * sh[1];
    edge_box_.push_back(eb);
    // Faces of the one-layer boundaries: those of the patch bbox the array
    // reaches, less the lower faces on a symmetry plane
    // (sommerfeld_rout.f90:61-119, 272-325)
    int fc = 0;
    for (int a = 0; a < 3; ++a) {
      const bool sym_face = (a == 2 ? symmetry > 0 : symmetry == 2) && std::fabs(bb[a]) < hh[a] / 2;
      if (std::fabs(b->X[a][sh[a] - 1] - bb[3 + a]) < hh[a]) fc |= 1 << a;
      if (std::fabs(b->X[a][0] - bb[a]) < hh[a] && !sym_face) fc |= 1 << (3 + a);
    }
    faces_.push_back(fc);
    SommerGeom g;
    if (bc == AMSS_BC_COPY_LAYER) g.copy = fc;
    if (bc == AMSS_BC_SOMMERFELD) {
      g.faces = fc;
      for (int a = 0; a < 3; ++a) {
        g.x0[a] = b->X[a][0];
        g.h[a] = hh[a];
        g.lo[a] = eb.mlo[a] < 0 ? -1 : 0;  // one mirrored layer (symmetry_bd(1, ...))
      }
    }
    sommer_.push_back(g);
  }
  LCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  if (bc != AMSS_BC_FROZEN_GHOSTS || gw < 3) {
    for (const Box& eb : edge_box_)
      edge_k_tiles_ = std::max<long>(edge_k_tiles_, edge_tile_set(edge_shells_of(eb.n, eb.lo, eb.hi, eb.mlo)).tiles());
    if (edge_k_tiles_ > 0) LCHK(hipMalloc(&edge_k_, edge_k_tiles_ * NFIELD * kEdgeThreads * sizeof(double)));
    LCHK(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    for (int o = 0; o < 3; ++o) {
      LCHK(hipStreamCreateWithFlags(&side_[o], hipStreamNonBlocking));
      LCHK(hipEventCreateWithFlags(&join_ev_[o], hipEventDisableTiming));
    }
  }
  LCHK(hipMalloc(&dflag_, sizeof(int)));
  build_ghost_segments();
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    ncu_ = prop.multiProcessorCount;
  build_batch_tables();
  build_l2_jobs();
}

BssnLevel::~BssnLevel() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (Block* b : blocks_) {
    if (b->slab) (void)hipFree(b->slab);
    delete b;
  }
  for (Patch& P : patches_) delete P.blb;
  for (MyList<var>* l : {StateList, OldStateList, SynchList_pre, SynchList_cor, RHSList, ConsList})
    if (l) {
      l->clearList();
      delete l;
    }
  for (var* v : vars_) delete v;
  for (Ev& e : ev_) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  clear_graphs();
  if (dflag_) (void)hipFree(dflag_);
  if (edge_k_) (void)hipFree(edge_k_);
  if (dpd_) (void)hipFree(dpd_);
  for (TileRef* t : dtiles_)
    if (t) (void)hipFree(t);
  if (dgbox_) (void)hipFree(dgbox_);
  if (dgwork_) (void)hipFree(dgwork_);
  for (void* p : {(void*)dl2_, (void*)dl2part_, (void*)dl2last_, (void*)dseries_})
    if (p) (void)hipFree(p);
  for (int o = 0; o < 3; ++o) {
    if (side_[o]) {
      (void)hipStreamSynchronize(side_[o]);
      (void)hipStreamDestroy(side_[o]);
    }
    if (join_ev_[o]) (void)hipEventDestroy(join_ev_[o]);
  }
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// The boundary kernels on the shell of patch pi (the reference's edge
// stencils), one side stream per stencil-order shell (where the LDS kernels
// take orders 1 and 2, both run on the order-1 stream and the order-2 call
// queues nothing), each after what stream_ has queued so far: they read the
// stage input and write only shell points, so they overlap the interior
// kernel.  join_edges makes stream_ wait for them.
int BssnLevel::launch_edges(const Block& b, int pi, MyList<var>* in, MyList<var>* out, const AuxOut& aux,
                            const Rk4Param& R, bool res) {
  RCHK(hipEventRecord(fork_ev_, stream_));
  for (int o = 0; o < 3; ++o) {
    RCHK(hipStreamWaitEvent(side_[o], fork_ev_, 0));
    Rk4Param Ro = R;
    Ro.edge_order = o;
    launch_rk4(list_in(b, in), list_in(b, StateList), list_out(b, RHSList), list_out(b, out), aux, edge_box_[pi],
               coef_[pi], Ro, gauge_, res, dflag_, true, side_[o], /*interior=*/false, sommer_[pi]);
    RCHK(hipEventRecord(join_ev_[o], side_[o]));
  }
  return 0;
}

int BssnLevel::join_edges() {
  for (int o = 0; o < 3; ++o) RCHK(hipStreamWaitEvent(stream_, join_ev_[o], 0));
  return 0;
}

double* BssnLevel::field(int patch, int list, int v) {
  if (patch < 0 || patch >= (int)blocks_.size()) return nullptr;
  MyList<var>* l = nullptr;
  switch (list) {
    case AMSS_LIST_STATE: l = StateList; break;
    case AMSS_LIST_OLD: l = OldStateList; break;
    case AMSS_LIST_PRE: l = SynchList_pre; break;
    case AMSS_LIST_COR: l = SynchList_cor; break;
    case AMSS_LIST_RHS: l = RHSList; break;
    case AMSS_LIST_RES: l = ConsList; break;
    default: return nullptr;
  }
  for (int i = 0; i < v && l; ++i) l = l->next;
  if (!l || v < 0) return nullptr;
  return blocks_[patch]->fgfs[l->data->sgfn];
}

long BssnLevel::npts(int patch) const {
  if (patch < 0 || patch >= (int)blocks_.size()) return 0;
  return blocks_[patch]->npts();
}

FieldsIn BssnLevel::list_in(const Block& b, MyList<var>* l) const {
  FieldsIn f;
  for (int v = 0; v < NFIELD; ++v, l = l->next) f.p[v] = b.fgfs[l->data->sgfn];
  return f;
}
FieldsOut BssnLevel::list_out(const Block& b, MyList<var>* l) const {
  FieldsOut f;
  for (int v = 0; v < NFIELD; ++v, l = l->next) f.p[v] = b.fgfs[l->data->sgfn];
  return f;
}

// Ghost points of patch A that lie in the interior of patch B receive B's
// values after every sub-step (the intra-level part of Parallel::Sync).
void BssnLevel::build_ghost_segments() {
  segs_.clear();
  if (blocks_.size() < 2) return;
  double gmin[3] = {1e300, 1e300, 1e300};
  for (Block* b : blocks_)
    for (int a = 0; a < 3; ++a) gmin[a] = std::min(gmin[a], b->bbox[a]);
  const double h = blocks_[0]->getdX(0);
  std::vector<std::array<int, 3>> org(blocks_.size());
  for (size_t i = 0; i < blocks_.size(); ++i)
    for (int a = 0; a < 3; ++a) org[i][a] = (int)std::lround((blocks_[i]->bbox[a] - gmin[a]) / h);
  const int gw = gw_;
  for (size_t A = 0; A < blocks_.size(); ++A)
    for (size_t B = 0; B < blocks_.size(); ++B) {
      if (A == B) continue;
      GhostSeg s;
      bool empty = false;
      for (int a = 0; a < 3; ++a) {
        const int alo = org[A][a], ahi = org[A][a] + blocks_[A]->shape[a];            // A's full array
        const int blo = org[B][a] + gw, bhi = org[B][a] + blocks_[B]->shape[a] - gw;  // B's interior
        const int lo = std::max(alo, blo), hi = std::min(ahi, bhi);
        if (hi <= lo) { empty = true; break; }
        s.dlo[a] = lo - org[A][a];
        s.slo[a] = lo - org[B][a];
        s.ext[a] = hi - lo;
      }
      if (empty) continue;
      s.src = (int)B;
      s.dst = (int)A;
      segs_.push_back(s);
    }
}

int BssnLevel::sync_list(int list) {
  MyList<var>* l = nullptr;
  switch (list) {
    case AMSS_LIST_STATE: l = StateList; break;
    case AMSS_LIST_OLD: l = OldStateList; break;
    case AMSS_LIST_PRE: l = SynchList_pre; break;
    case AMSS_LIST_COR: l = SynchList_cor; break;
    case AMSS_LIST_RHS: l = RHSList; break;
    default: err_ = "sync_list: not an evolved-field list"; return AMSS_ERR_ARG;
  }
  sync(l, 0);
  return 0;
}

void BssnLevel::sync(MyList<var>* list, int /*stage*/) {
  for (const GhostSeg& g : segs_) {
    SegDesc d;
    for (int a = 0; a < 3; ++a) {
      d.slo[a] = g.slo[a];
      d.dlo[a] = g.dlo[a];
      d.ext[a] = g.ext[a];
    }
    const Block& S = *blocks_[g.src];
    const Block& D = *blocks_[g.dst];
    d.ssy = S.shape[0];
    d.ssz = (long)S.shape[0] * S.shape[1];
    d.dsy = D.shape[0];
    d.dsz = (long)D.shape[0] * D.shape[1];
    launch_seg_copy(list_in(S, list), list_out(D, list), d, stream_);
  }
}

// One RK4 step, one launch per patch and sub-step.
int BssnLevel::Step(double dT, bool want_res) {
  const double tinny = chitiny_ - 1.0;  // bssn_class.C:179
  const double cst[4] = {0.5 * dT, 0.5 * dT, dT, (1.0 / 6.0) * dT};
  // predictor's enforce_ga on the state, in place (bssn_class.C:3051-3057)
  for (Block* b : blocks_) launch_enforce(list_out(*b, StateList), b->npts(), stream_);
  for (int it = 0; it < 4; ++it) {
    MyList<var>* in = it == 0 ? StateList : SynchList_pre;
    MyList<var>* out = it == 0 ? SynchList_pre : SynchList_cor;
    for (size_t pi = 0; pi < blocks_.size(); ++pi) {
      Block& b = *blocks_[pi];
      Box box;
      for (int a = 0; a < 3; ++a) {
        box.n[a] = b.shape[a];
        box.lo[a] = 3;
        box.hi[a] = b.shape[a] - 3;
      }
      box.sy = b.shape[0];
      box.sz = (long)b.shape[0] * b.shape[1];
      AuxOut aux;
      for (int a = 0; a < 18; ++a) aux.gam[a] = nullptr;
      for (int a = 0; a < 6; ++a) aux.ric[a] = nullptr;
      MyList<var>* c = ConsList;
      for (int a = 0; a < 7; ++a, c = c->next) aux.res[a] = b.fgfs[c->data->sgfn];
      Rk4Param R{it, variant_, cst[it], tinny};
      R.edge_k = edge_k_;
      R.edge_k_tiles = edge_k_tiles_;
      // Sub-step 2 writes the list sub-step 0 wrote (the pre/cor swap), whose
      // frozen points already hold the same values unless the zone is deeper
      // than the interior kernel's start (gw > 3); a 3-deep fill can run in
      // the interior kernel's launch.
      const bool frozen_fill = bc_ == AMSS_BC_FROZEN_GHOSTS && !(it == 2 && gw_ == 3);
      if (frozen_fill && gw_ == 3) R.fill_w = 3;
      const bool fused_fill = frozen_fill && rk4_fuses_fill(box, R);
      const bool edge = (bc_ != AMSS_BC_FROZEN_GHOSTS) || gw_ < 3;
      if (edge)
        if (int r = launch_edges(b, (int)pi, in, out, aux, R, want_res && it == 0)) return r;
      Ev* ev = nullptr;
      if (int r = time_event_begin(it, &ev)) return r;
      launch_rk4(list_in(b, in), list_in(b, StateList), list_out(b, RHSList), list_out(b, out), aux, box,
                 coef_[pi], R, gauge_, want_res && it == 0, dflag_, false, stream_);
      if (ev) RCHK(hipEventRecord(ev->b, stream_));
      if (edge)
        if (int r = join_edges()) return r;
      if (bc_ == AMSS_BC_FROZEN_GHOSTS) {
        Faces w{{gw_, gw_, gw_}, {gw_, gw_, gw_}};
        if (frozen_fill && !fused_fill)
          launch_ghost_fill(list_in(b, StateList), list_out(b, out), box, w, it, tinny, dflag_, stream_);
      } else if (bc_ == AMSS_BC_COPY_LAYER) {
        // sommerfeld_rout(precor = 1) faces (sommerfeld_rout.f90:61-137)
        Faces w{{0, 0, 0}, {0, 0, 0}};
        for (int a = 0; a < 3; ++a) {
          w.hi[a] = (faces_[pi] >> a) & 1;
          w.lo[a] = (faces_[pi] >> (3 + a)) & 1;
        }
        launch_ghost_fill(list_in(b, StateList), list_out(b, out), box, w, it, tinny, dflag_, stream_);
      }
    }
    sync(out, it);  // Parallel::Sync (bssn_class.C:3340, 3710)
    if (it > 0 && it < 3)
      for (Block* b : blocks_) b->swapList(SynchList_pre, SynchList_cor);  // :3765-3781
  }
  rotate_lists();
  if (hipGetLastError() != hipSuccess) {
    err_ = "kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

// rotate: the corrector output becomes the state, the state the old state
// (bssn_class.C:3845-3859)
void BssnLevel::rotate_lists() {
  for (Block* b : blocks_) {
    b->swapList(StateList, SynchList_cor);
    b->swapList(OldStateList, SynchList_cor);
  }
}

int BssnLevel::time_event_begin(int stage, Ev** ev) {
  *ev = nullptr;
  if (!timing_ || capturing_) return 0;
  if (ev_used_ == ev_.size()) {
    Ev e{};
    RCHK(hipEventCreate(&e.a));
    RCHK(hipEventCreate(&e.b));
    ev_.push_back(e);
  }
  *ev = &ev_[ev_used_++];
  (*ev)->stage = stage;
  RCHK(hipEventRecord((*ev)->a, stream_));
  return 0;
}

// A list's slot in every Block's slab: lists move as whole 24-field runs.
int BssnLevel::slot_of(MyList<var>* l) const {
  const Block& b = *blocks_[0];
  return (int)((b.fgfs[l->data->sgfn] - b.slab) / ((long)NFIELD * b.npts()));
}

// Device tables of the batched step: patch views, the interior tiles of every
// patch for each tile shape (whole tiles only; what is left of an interior
// goes to the direct kernel), the boundary boxes of every patch and the
// inter-patch exchange boxes.
void BssnLevel::build_batch_tables() {
  const int np = (int)blocks_.size();
  hpd_.clear();
  maxn_ = 0;
  for (Block* b : blocks_) {
    PatchDesc d;
    d.slab = b->slab;
    d.n = b->npts();
    d.sy = b->shape[0];
    d.sz = b->shape[0] * b->shape[1];
    for (int a = 0; a < 3; ++a) d.shape[a] = b->shape[a];
    hpd_.push_back(d);
    maxn_ = std::max(maxn_, d.n);
  }
  LCHK(hipMalloc(&dpd_, np * sizeof(PatchDesc)));
  LCHK(hipMemcpy(dpd_, hpd_.data(), np * sizeof(PatchDesc), hipMemcpyHostToDevice));
  for (int sh = 0; sh < kNumTileShapes; ++sh) {
    const int T[3] = {tile_dim(sh, 0), tile_dim(sh, 1), tile_dim(sh, 2)};
    htiles_[sh].clear();
    rem_[sh].clear();
    for (int pi = 0; pi < np; ++pi) {
      const Block& b = *blocks_[pi];
      Box box;
      for (int a = 0; a < 3; ++a) {
        box.n[a] = b.shape[a];
        box.lo[a] = 3;
        box.hi[a] = b.shape[a] - 3;
      }
      box.sy = b.shape[0];
      box.sz = (long)b.shape[0] * b.shape[1];
      int core[3];
      for (int a = 0; a < 3; ++a) core[a] = 3 + (box.hi[a] - 3) / T[a] * T[a];
      // tiles need an even row pitch (16-byte aligned direct loads); a patch
      // with an odd one is all remainder (direct kernel)
      if (b.shape[0] % 2 != 0) core[0] = core[1] = core[2] = 3;
      if (core[0] > 3 && core[1] > 3 && core[2] > 3) {
        if (b.shape[0] > 65535 || b.shape[1] > 65535 || b.shape[2] > 65535 || pi > 65535) {
          err_ = "patch too large for the tile list";
          return;
        }
        for (int k = 3; k < core[2]; k += T[2])
          for (int j = 3; j < core[1]; j += T[1])
            for (int i = 3; i < core[0]; i += T[0])
              htiles_[sh].push_back(TileRef{(unsigned short)pi, (unsigned short)i, (unsigned short)j,
                                            (unsigned short)k});
      } else {
        for (int a = 0; a < 3; ++a) core[a] = 3;
      }
      for (int a = 0; a < 3; ++a) {  // remainder slab [core[a], hi[a]) along a
        Box r = box;
        for (int c = 0; c < a; ++c) r.hi[c] = core[c];
        r.lo[a] = core[a];
        if (r.hi[0] > r.lo[0] && r.hi[1] > r.lo[1] && r.hi[2] > r.lo[2]) rem_[sh].push_back({pi, r});
      }
    }
    if ((long)htiles_[sh].size() > (1L << 31) - 1) { err_ = "too many tiles"; return; }
    if (!htiles_[sh].empty()) {
      LCHK(hipMalloc(&dtiles_[sh], htiles_[sh].size() * sizeof(TileRef)));
      LCHK(hipMemcpy(dtiles_[sh], htiles_[sh].data(), htiles_[sh].size() * sizeof(TileRef),
                     hipMemcpyHostToDevice));
    }
  }
  // boundary boxes: the points outside [w.lo, n - w.hi) of each patch, as at
  // most 6 disjoint slabs (z faces whole, then y faces, then x faces), less
  // the points an exchange box overwrites
  std::vector<std::vector<IBox>> fill(np), cover(np);
  if (bc_ == AMSS_BC_FROZEN_GHOSTS || bc_ == AMSS_BC_COPY_LAYER) {
    for (int pi = 0; pi < np; ++pi) {
      const Block& b = *blocks_[pi];
      int wl[3], wh[3];
      for (int a = 0; a < 3; ++a) {
        if (bc_ == AMSS_BC_FROZEN_GHOSTS) {
          wl[a] = wh[a] = gw_;
        } else {
          wh[a] = (faces_[pi] >> a) & 1;
          wl[a] = (faces_[pi] >> (3 + a)) & 1;
        }
      }
      int lo[3] = {0, 0, 0}, hi[3] = {b.shape[0], b.shape[1], b.shape[2]};
      for (int a = 2; a >= 0; --a) {
        for (int side = 0; side < 2; ++side) {
          IBox f;
          for (int c = 0; c < 3; ++c) {
            f.lo[c] = lo[c];
            f.hi[c] = hi[c];
          }
          if (side == 0) f.hi[a] = wl[a];
          else f.lo[a] = b.shape[a] - wh[a];
          if (f.hi[0] > f.lo[0] && f.hi[1] > f.lo[1] && f.hi[2] > f.lo[2]) fill[pi].push_back(f);
        }
        lo[a] = wl[a];
        hi[a] = b.shape[a] - wh[a];
      }
    }
  }
  hgbox_.clear();
  std::vector<GhostBox> segbox;
  ghost_fused_ = true;
  for (const GhostSeg& g : segs_) {
    GhostBox r;
    r.kind = 1;
    r.src = g.src;
    r.dst = g.dst;
    IBox d, src;
    for (int a = 0; a < 3; ++a) {
      r.slo[a] = g.slo[a];
      r.dlo[a] = g.dlo[a];
      r.ext[a] = g.ext[a];
      d.lo[a] = g.dlo[a];
      d.hi[a] = g.dlo[a] + g.ext[a];
      src.lo[a] = g.slo[a];
      src.hi[a] = g.slo[a] + g.ext[a];
    }
    cover[g.dst].push_back(d);
    for (const IBox& f : fill[g.src])
      if (f.intersects(src)) ghost_fused_ = false;  // the copy must read the filled values
    segbox.push_back(r);
  }
  for (int pi = 0; pi < np; ++pi) {
    std::vector<IBox> left;
    subtract_boxes(fill[pi], cover[pi], left);
    for (const IBox& f : left) {
      GhostBox r{};
      r.kind = 0;
      r.src = r.dst = pi;
      for (int a = 0; a < 3; ++a) {
        r.dlo[a] = f.lo[a];
        r.ext[a] = f.hi[a] - f.lo[a];
      }
      hgbox_.push_back(r);
    }
  }
  const int nfillbox = (int)hgbox_.size();
  hgbox_.insert(hgbox_.end(), segbox.begin(), segbox.end());
  std::vector<GhostWork> work;
  for (int bi = 0; bi < (int)hgbox_.size(); ++bi) {
    const GhostBox& r = hgbox_[bi];
    const long npts = (long)r.ext[0] * r.ext[1] * r.ext[2];  // < 2^29: inside one Block
    for (long t = 0; t < npts; t += 256) work.push_back(GhostWork{bi, (int)t});
    if (bi == nfillbox - 1) nfill_work_ = (int)work.size();
  }
  if (nfillbox == 0) nfill_work_ = 0;
  nseg_work_ = (int)work.size() - nfill_work_;
  if (work.size() > (size_t)INT32_MAX) {
    err_ = "too many ghost points";
    return;
  }
  if (!hgbox_.empty()) {
    LCHK(hipMalloc(&dgbox_, hgbox_.size() * sizeof(GhostBox)));
    LCHK(hipMemcpy(dgbox_, hgbox_.data(), hgbox_.size() * sizeof(GhostBox), hipMemcpyHostToDevice));
  }
  if (!work.empty()) {
    LCHK(hipMalloc(&dgwork_, work.size() * sizeof(GhostWork)));
    LCHK(hipMemcpy(dgwork_, work.data(), work.size() * sizeof(GhostWork), hipMemcpyHostToDevice));
  }
}

// One RK4 step of the whole level, one launch per kind of work and sub-step
// (same order of operations as Step).
int BssnLevel::StepBatched(double dT, bool want_res) {
  const double tinny = chitiny_ - 1.0;
  const double cst[4] = {0.5 * dT, 0.5 * dT, dT, (1.0 / 6.0) * dT};
  const int np = (int)blocks_.size();
  const int sh = tune_.shape;
  launch_enforce_patches(dpd_, np, maxn_, slot_of(StateList), stream_);
  for (int it = 0; it < 4; ++it) {
    MyList<var>* in = it == 0 ? StateList : SynchList_pre;
    MyList<var>* out = it == 0 ? SynchList_pre : SynchList_cor;
    const Roles m{slot_of(in), slot_of(StateList), slot_of(RHSList), slot_of(out)};
    Rk4Param R{it, 1, cst[it], tinny};
    R.edge_k = edge_k_;
    R.edge_k_tiles = edge_k_tiles_;
    const bool res = want_res && it == 0;
    // Sub-step 2 writes the list sub-step 0 wrote (the pre/cor swap after
    // sub-step 1), whose fill points already hold the same values when
    // nothing else writes ghost points (frozen ghosts 3 deep, no edge kernel;
    // wider frozen zones overlap the tiles, which start at index 3, so their
    // fill must follow the tiles in every sub-step).
    const bool frozen = bc_ == AMSS_BC_FROZEN_GHOSTS && gw_ == 3;
    const bool fill = !(it == 2 && frozen);
    // With frozen ghosts the fill reads only the state and writes only
    // points no tile writes: it runs in extra workgroups of the tile launch.
    const int ntiles = (int)htiles_[sh].size();
    const bool fill_in_tiles = frozen && ghost_fused_ && launch_rk4_tiles_fuses_fill(tune_, ntiles);
    FillWork fw;
    if (fill_in_tiles && fill) fw = FillWork{dgbox_, dgwork_, nfill_work_};
    Ev* ev = nullptr;
    if (int r = time_event_begin(it, &ev)) return r;
    launch_rk4_tiles(gauge_, dpd_, dtiles_[sh], 0, ntiles, m, coef_[0], R, res, tune_, dflag_, stream_, fw);
    if (ev) RCHK(hipEventRecord(ev->b, stream_));
    AuxOut aux;
    for (int a = 0; a < 18; ++a) aux.gam[a] = nullptr;
    for (int a = 0; a < 6; ++a) aux.ric[a] = nullptr;
    Rk4Param Rd = R;
    Rd.variant = 0;
    for (const auto& pr : rem_[sh]) {
      const Block& b = *blocks_[pr.first];
      for (int a = 0; a < 7; ++a) aux.res[a] = b.fgfs[kResidualBase + a];
      launch_rk4(list_in(b, in), list_in(b, StateList), list_out(b, RHSList), list_out(b, out), aux, pr.second,
                 coef_[pr.first], Rd, gauge_, res, dflag_, false, stream_, true);
    }
    if (bc_ != AMSS_BC_FROZEN_GHOSTS || gw_ < 3) {
      for (int pi = 0; pi < np; ++pi) {
        const Block& b = *blocks_[pi];
        for (int a = 0; a < 7; ++a) aux.res[a] = b.fgfs[kResidualBase + a];
        launch_rk4(list_in(b, in), list_in(b, StateList), list_out(b, RHSList), list_out(b, out), aux,
                   edge_box_[pi], coef_[pi], R, gauge_, res, dflag_, true, stream_, /*interior=*/false,
                   sommer_[pi]);
      }
    }
    // The state is the same in all sub-steps, so sub-step 0 tests it for NaN
    // at the exchange points as well and the later sub-steps need not.
    if (fill_in_tiles) {
      launch_ghost_boxes(dpd_, dgbox_, dgwork_ + nfill_work_, nseg_work_, m, it, tinny, it == 0, dflag_, stream_);
    } else if (ghost_fused_ && fill) {
      launch_ghost_boxes(dpd_, dgbox_, dgwork_, nfill_work_ + nseg_work_, m, it, tinny, it == 0, dflag_, stream_);
    } else {
      if (fill) launch_ghost_boxes(dpd_, dgbox_, dgwork_, nfill_work_, m, it, tinny, false, dflag_, stream_);
      launch_ghost_boxes(dpd_, dgbox_, dgwork_ + nfill_work_, nseg_work_, m, it, tinny, it == 0, dflag_, stream_);
    }
    if (it > 0 && it < 3)
      for (Block* b : blocks_) b->swapList(SynchList_pre, SynchList_cor);
  }
  rotate_lists();
  if (hipGetLastError() != hipSuccess) {
    err_ = "kernel launch failed";
    return AMSS_ERR_HIP;
  }
  return 0;
}

// One batched step replayed from a hipGraph.  The list slots return to the
// same assignment every 3 steps, so at most 3 graphs (keyed by the slot
// assignment at the start of the step, and dT) cover a run.
int BssnLevel::StepGraph(double dT, bool want_res) {
  const int key = (((slot_of(StateList) * 5 + slot_of(OldStateList)) * 5 + slot_of(SynchList_pre)) * 5 +
                   slot_of(SynchList_cor)) * 5 + slot_of(RHSList);
  for (const Graph& g : graphs_)
    if (g.key == key && g.dT == dT && g.res == want_res) {
      RCHK(hipGraphLaunch(g.exec, stream_));
      // the pre/cor swaps of sub-steps 1 and 2 cancel; only the rotation is left
      rotate_lists();
      return 0;
    }
  hipGraph_t graph = nullptr;
  RCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  capturing_ = true;
  const int r = StepBatched(dT, want_res);
  capturing_ = false;
  const hipError_t e = hipStreamEndCapture(stream_, &graph);
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  if (e != hipSuccess) {
    err_ = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
    return AMSS_ERR_HIP;
  }
  Graph g{key, dT, want_res, nullptr};
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    err_ = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
    return AMSS_ERR_HIP;
  }
  graphs_.push_back(g);
  RCHK(hipGraphLaunch(g.exec, stream_));
  return 0;
}

void BssnLevel::clear_graphs() {
  for (Graph& g : graphs_) (void)hipGraphExecDestroy(g.exec);
  graphs_.clear();
}

int BssnLevel::set_tune(const BatchTune& t) {
  if (t.shape < 0 || t.shape >= kNumTileShapes || t.grid < 0 || (t.remap != 0 && t.remap != 1)) {
    err_ = "bad tile shape, grid or order";
    return AMSS_ERR_ARG;
  }
  tune_ = t;
  tuned_ = true;
  clear_graphs();
  return 0;
}

// Candidates: the tile shapes x {one workgroup per tile, persistent with 1x
// and 2x the resident workgroups (2 waves per SIMD)} x {list order,
// XCD-contiguous order}.  Each is timed (best of 2 after a warm-up) on
// sub-step 0 of the current state, which writes only the RHS and stage lists
// that the next step overwrites anyway.  ms (may be NULL) receives the times
// in candidate order.
int BssnLevel::autotune(double* ms, int nms) {
  const double tinny = chitiny_ - 1.0;
  const Roles m{slot_of(StateList), slot_of(StateList), slot_of(RHSList), slot_of(SynchList_pre)};
  const Rk4Param R{0, 1, 0.0, tinny};
  hipEvent_t e0, e1;
  RCHK(hipEventCreate(&e0));
  RCHK(hipEventCreate(&e1));
  double best = 1e300;
  BatchTune pick = tune_;
  int c = 0;
  AuxOut aux;
  for (int a = 0; a < 18; ++a) aux.gam[a] = nullptr;
  for (int a = 0; a < 6; ++a) aux.ric[a] = nullptr;
  Rk4Param Rd = R;
  Rd.variant = 0;
  for (int sh = 0; sh < kNumTileShapes; ++sh)
    for (int gm = 0; gm < 3; ++gm)
      for (int rm = 0; rm < 2; ++rm, ++c) {
        // resident workgroups per CU at 2 waves per SIMD: 512 threads / tile size
        const int per_cu = 512 / (tile_dim(sh, 0) * tile_dim(sh, 1) * tile_dim(sh, 2));
        const BatchTune t{sh, gm * per_cu * ncu_, rm};
        float tbest = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
          RCHK(hipEventRecord(e0, stream_));
          launch_rk4_tiles(gauge_, dpd_, dtiles_[sh], 0, (int)htiles_[sh].size(), m, coef_[0], R, false, t, dflag_,
                           stream_);
          for (const auto& pr : rem_[sh]) {
            const Block& b = *blocks_[pr.first];
            for (int a = 0; a < 7; ++a) aux.res[a] = b.fgfs[kResidualBase + a];
            launch_rk4(list_in(b, StateList), list_in(b, StateList), list_out(b, RHSList),
                       list_out(b, SynchList_pre), aux, pr.second, coef_[pr.first], Rd, gauge_, false, dflag_, false,
                       stream_, true);
          }
          RCHK(hipEventRecord(e1, stream_));
          RCHK(hipEventSynchronize(e1));
          float t_ms = 0.f;
          RCHK(hipEventElapsedTime(&t_ms, e0, e1));
          if (rep > 0) tbest = std::min(tbest, t_ms);
        }
        if (ms && c < nms) ms[c] = tbest;
        if (tbest < best) {
          best = tbest;
          pick = t;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (hipGetLastError() != hipSuccess) {
    err_ = "kernel launch failed";
    return AMSS_ERR_HIP;
  }
  tune_ = pick;
  tuned_ = true;
  clear_graphs();
  return 0;
}

int BssnLevel::step(double dT, int nsteps, int mode, int want_res) {
  if (mode < 0 || mode > 2) {
    err_ = "mode must be 0 (per patch), 1 (batched) or 2 (batched, hipGraph)";
    return AMSS_ERR_ARG;
  }
  if (nsteps < 0) {
    err_ = "negative step count";
    return AMSS_ERR_ARG;
  }
  if (want_res < 0 || want_res > 2) {
    err_ = "want_ham must be 0, 1 (last step) or 2 (every step)";
    return AMSS_ERR_ARG;
  }
  if (mode > 0 && !tuned_)
    if (int r = autotune(nullptr, 0)) return r;
  const int njobs = (int)hl2_.size();
  if (want_res == 2 && nsteps > series_cap_) {
    RCHK(hipStreamSynchronize(stream_));
    if (dseries_) RCHK(hipFree(dseries_));
    dseries_ = nullptr;
    series_cap_ = 0;
    RCHK(hipMalloc(&dseries_, (size_t)nsteps * njobs * sizeof(double)));
    series_cap_ = nsteps;
  }
  RCHK(hipMemsetAsync(dflag_, 0, sizeof(int), stream_));
  ev_used_ = 0;
  for (int n = 0; n < nsteps; ++n) {
    const bool res = want_res == 2 || (want_res == 1 && n == nsteps - 1);
    int r = 0;
    if (mode == 0)
      r = Step(dT, res);
    else if (mode == 1)
      r = StepBatched(dT, res);
    else
      r = StepGraph(dT, res);
    if (r) return r;
    // the predictor's residuals of this step -> its norms (Constraint_Out)
    if (want_res == 2) launch_l2_jobs(dl2_, njobs, l2maxparts_, dl2part_, dseries_ + (size_t)n * njobs, stream_);
  }
  int flag = 0;
  RCHK(hipMemcpyAsync(&flag, dflag_, sizeof(int), hipMemcpyDeviceToHost, stream_));
  if (want_res == 2) {
    series_.assign((size_t)nsteps * njobs, 0.0);
    if (nsteps > 0)
      RCHK(hipMemcpyAsync(series_.data(), dseries_, series_.size() * sizeof(double), hipMemcpyDeviceToHost,
                          stream_));
    series_steps_ = nsteps;
  }
  RCHK(hipStreamSynchronize(stream_));
  for (size_t i = 0; i < ev_used_; ++i) {
    float ms = 0.f;
    RCHK(hipEventElapsedTime(&ms, ev_[i].a, ev_[i].b));
    ms_[ev_[i].stage] += ms;
    nl_[ev_[i].stage] += 1;
  }
  ev_used_ = 0;
  if (want_res && nsteps > 0) ham_valid_ = true;
  if (hipGetLastError() != hipSuccess) {
    err_ = "kernel launch failed";
    return AMSS_ERR_HIP;
  }
  if (flag) {
    err_ = "NaN found in a right-hand-side input";
    return AMSS_ERR_NONFINITE;
  }
  return 0;
}

int BssnLevel::set_variant(int v) {
  if (v != 0 && v != 1) {
    err_ = "kernel variant must be 0 (direct) or 1 (LDS tiles)";
    return AMSS_ERR_ARG;
  }
  variant_ = v;
  return 0;
}

int BssnLevel::set_timing(bool on) {
  timing_ = on;
  for (int a = 0; a < 4; ++a) {
    ms_[a] = 0.0;
    nl_[a] = 0;
  }
  return 0;
}

int BssnLevel::kernel_ms(double* ms4, long* n4) {
  for (int a = 0; a < 4; ++a) {
    ms4[a] = ms_[a];
    n4[a] = nl_[a];
  }
  return 0;
}

// One L2 job per (patch, residual) in Constraint_Out's order (Ham, Px, Py,
// Pz, Gx, Gy, Gz; bssn_class.C:7475-7481) over the points l2normhelper sums
// (fmisc.f90:1283-1301): the ghost layers of a Block count where the Block
// reaches its patch's bbox, which a level's Block always does, so under the
// reference's boundary modes every point counts.  With frozen ghost zones
// (whose residuals are never evaluated) only [gw, n - gw) counts.
void BssnLevel::build_l2_jobs() {
  hl2_.clear();
  int part = 0;
  l2maxparts_ = 1;
  for (Block* b : blocks_) {
    const int w = bc_ == AMSS_BC_FROZEN_GHOSTS ? gw_ : 0;
    L2Job J;
    long N = 1;
    for (int a = 0; a < 3; ++a) {
      J.lo[a] = w;
      J.hi[a] = std::max(w, b->shape[a] - w);
      N *= J.hi[a] - J.lo[a];
    }
    J.sy = b->shape[0];
    J.sz = (long)b->shape[0] * b->shape[1];
    J.dv = b->getdX(0) * b->getdX(1) * b->getdX(2);
    J.nparts = (int)std::min<long>(512, std::max<long>(1, (N + 4095) / 4096));
    l2maxparts_ = std::max(l2maxparts_, J.nparts);
    MyList<var>* c = ConsList;
    for (int r = 0; r < 7; ++r, c = c->next) {
      J.f = b->fgfs[c->data->sgfn];
      J.part0 = part;
      part += J.nparts;
      hl2_.push_back(J);
    }
  }
  LCHK(hipMalloc(&dl2_, hl2_.size() * sizeof(L2Job)));
  LCHK(hipMemcpy(dl2_, hl2_.data(), hl2_.size() * sizeof(L2Job), hipMemcpyHostToDevice));
  LCHK(hipMalloc(&dl2part_, (size_t)std::max(part, 1) * sizeof(double)));
  LCHK(hipMalloc(&dl2last_, hl2_.size() * sizeof(double)));
}

// sqrt of the per-patch sums (one patch, or all patches added: -1)
int BssnLevel::norms_of(const double* sums, int patch, double* out7) const {
  const int np = (int)blocks_.size();
  if (patch < -1 || patch >= np) return AMSS_ERR_ARG;
  for (int r = 0; r < 7; ++r) {
    double t = 0.0;
    for (int p = (patch < 0 ? 0 : patch); p < (patch < 0 ? np : patch + 1); ++p) t += sums[p * 7 + r];
    out7[r] = std::sqrt(t);
  }
  return 0;
}

int BssnLevel::constraint_norms(int patch, double* out7) {
  if (!ham_valid_) {
    err_ = "no constraint residuals computed yet (step with want_ham)";
    return AMSS_ERR_ARG;
  }
  if (patch < -1 || patch >= (int)blocks_.size()) {
    err_ = "patch index out of range";
    return AMSS_ERR_ARG;
  }
  std::vector<double> sums(hl2_.size());
  launch_l2_jobs(dl2_, (int)hl2_.size(), l2maxparts_, dl2part_, dl2last_, stream_);
  RCHK(hipMemcpyAsync(sums.data(), dl2last_, sums.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  RCHK(hipStreamSynchronize(stream_));
  return norms_of(sums.data(), patch, out7);
}

int BssnLevel::constraint_series(int patch, double* out, int maxsteps, int* nsteps) {
  if (patch < -1 || patch >= (int)blocks_.size()) {
    err_ = "patch index out of range";
    return AMSS_ERR_ARG;
  }
  const int n = std::min(maxsteps, series_steps_);
  const size_t nj = hl2_.size();
  for (int s = 0; s < n; ++s) norms_of(series_.data() + s * nj, patch, out + 7 * s);
  *nsteps = series_steps_;
  return 0;
}

int BssnLevel::ham_l2(double* out) {
  double v[7];
  if (int r = constraint_norms(-1, v)) return r;
  *out = v[0];
  return 0;
}

}  // namespace amss
