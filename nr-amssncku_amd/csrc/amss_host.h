// amss_host.h -- host-side storage surface of the BSSN path, shaped like the
// reference's (Block.h:9-32, var.h:5-25, MyList.h:8-108, MPatch.h:11-48), with
// grid functions resident in HBM: a Block owns one device slab and fgfs[i]
// are device pointers into it.  BssnLevel::Step follows bssn_class::Step
// (bssn_class.C:2981-3901) for one refinement level on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "bssn_batch.h"

namespace amss {

// A grid-function descriptor (var.h:8-12): name, slot number, parities.
struct var {
  char name[20];
  int sgfn;
  double SoA[3];
  double propspeed;
};

// Singly linked list walked in lockstep by Step (MyList.h:8-108).
template <class T>
struct MyList {
  T* data;
  MyList* next;
  explicit MyList(T* d) : data(d), next(nullptr) {}
  void insert(T* d) {
    MyList* p = this;
    while (p->next) p = p->next;
    p->next = new MyList(d);
  }
  void clearList() {
    MyList* p = next;
    while (p) {
      MyList* q = p->next;
      delete p;
      p = q;
    }
    next = nullptr;
  }
};

// Block (Block.h:13-20): shape, bbox, cell-centred coordinates, rank and
// level, and fgfs[] -- here device pointers into one HBM slab per Block.
struct Block {
  int shape[3];
  double bbox[6];
  std::vector<double> X[3];  // host copies of the coordinates
  int rank = 0, lev = 0, cgpu = 1;
  int fngfs = 0;
  double* slab = nullptr;        // fngfs * N doubles, one allocation
  std::vector<double*> fgfs;     // fgfs[i] = slab + i * N
  long npts() const { return (long)shape[0] * shape[1] * shape[2]; }
  double getdX(int dir) const { return X[dir][1] - X[dir][0]; }
  // Block::swapList (Block.C:182-199): swap the slot pointers of two lists.
  void swapList(MyList<var>* a, MyList<var>* b);
};

// Patch (MPatch.h:13-19): one Block per patch on a single GPU.
struct Patch {
  int lev = 0;
  int shape[3];
  double bbox[6];
  MyList<Block>* blb = nullptr;
  MyList<Block>* ble = nullptr;
};

// A contiguous copy of a box of points from one patch to another
// (the single-rank form of Parallel::Sync's transfer segments).
struct GhostSeg {
  int src, dst;       // patch indices
  int slo[3], dlo[3]; // box origins in source / destination index space
  int ext[3];         // box extent
};

class BssnLevel {
 public:
  struct Ev { hipEvent_t a, b; int stage; };
  BssnLevel(int npatch, const int* shapes, const double* bboxes, int gauge, double eps, double chitiny, int bc,
            int gw, int symmetry);
  ~BssnLevel();
  bool ok() const { return err_.empty(); }
  const std::string& err() const { return err_; }

  double* field(int patch, int list, int v);
  // want_res: 0 none; 1 residuals in the last step's predictor; 2 in every
  // predictor (the reference's co = 0), with the seven constraint norms of
  // every patch recorded per step
  int step(double dT, int nsteps, int mode, int want_res);
  int ham_l2(double* out);
  // the seven constraint L2 norms (Ham, Px, Py, Pz, Gx, Gy, Gz) of one patch,
  // or of all patches together (patch -1), from the last residuals
  int constraint_norms(int patch, double* out7);
  // the per-step norms recorded by the last step() with want_res 2
  int constraint_series(int patch, double* out, int maxsteps, int* nsteps);
  hipStream_t stream() const { return stream_; }
  // HIP-event timing of the fused RHS+RK4 interior launches (per stage)
  int set_timing(bool on);
  // interior kernel variant: 0 direct L1/L2 reads, 1 LDS-staged tiles
  int set_variant(int v);
  int variant() const { return variant_; }
  int kernel_ms(double* ms4, long* n4);
  // batched-step tile shape / grid / order: time every candidate on a
  // sub-step of the current state and keep the fastest; or set one
  int autotune(double* ms, int nms);
  int set_tune(const BatchTune& t);
  const BatchTune& tune() const { return tune_; }
  int npatch() const { return (int)patches_.size(); }
  const Patch& patch(int i) const { return patches_[i]; }
  const Block& block(int i) const { return *blocks_[i]; }
  long npts(int patch) const;
  // Parallel::Sync of one list (AMSS_LIST_*) on the level's stream
  int sync_list(int list);

  // the reference's lists (bssn_class.C:412-535)
  MyList<var>* OldStateList = nullptr;
  MyList<var>* StateList = nullptr;
  MyList<var>* SynchList_pre = nullptr;
  MyList<var>* SynchList_cor = nullptr;
  MyList<var>* RHSList = nullptr;
  MyList<var>* ConsList = nullptr;

 private:
  int Step(double dT, bool want_res);  // one RK4 step, one launch per patch and sub-step
  int StepBatched(double dT, bool want_res);
  int StepGraph(double dT, bool want_res);
  void build_l2_jobs();
  int norms_of(const double* sums, int patch, double* out7) const;
  void rotate_lists();  // the end-of-step list rotation (bssn_class.C:3845-3859)
  int slot_of(MyList<var>* l) const;
  int time_event_begin(int stage, Ev** ev);
  void build_ghost_segments();
  void build_batch_tables();
  void clear_graphs();
  void sync(MyList<var>* list, int stage);
  FieldsIn list_in(const Block& b, MyList<var>* l) const;
  FieldsOut list_out(const Block& b, MyList<var>* l) const;

  std::vector<Patch> patches_;
  std::vector<Block*> blocks_;
  std::vector<var*> vars_;
  std::vector<GhostSeg> segs_;
  std::vector<Coef> coef_;
  std::vector<SommerGeom> sommer_;  // per patch; no faces unless bc_ == AMSS_BC_SOMMERFELD
  std::vector<Box> edge_box_;       // per patch: the boundary kernel's box, mirror bounds at symmetry planes
  std::vector<int> faces_;          // per patch: one-layer boundary faces (bit a upper, bit 3 + a lower)
  int gauge_, bc_, gw_, symmetry_;
  double eps_, chitiny_;
  hipStream_t stream_ = nullptr;
  // the boundary kernels' three stencil-order shells run on side streams,
  // concurrently with the interior kernel (they write disjoint points)
  hipStream_t side_[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork_ev_ = nullptr, join_ev_[3] = {nullptr, nullptr, nullptr};
  int launch_edges(const Block& b, int pi, MyList<var>* in, MyList<var>* out, const AuxOut& aux, const Rk4Param& R,
                   bool res);
  int join_edges();
  int* dflag_ = nullptr;
  // right-hand sides of the LDS boundary kernel (Rk4Param::edge_k), sized
  // for the largest patch's order-1 and order-2 shells
  double* edge_k_ = nullptr;
  long edge_k_tiles_ = 0;
  bool ham_valid_ = false;
  // constraint norms: one L2 job per (patch, residual), their partials, the
  // last evaluation and the per-step series (device, then host copy)
  std::vector<L2Job> hl2_;
  L2Job* dl2_ = nullptr;
  double* dl2part_ = nullptr;
  int l2maxparts_ = 1;
  double* dl2last_ = nullptr;
  double* dseries_ = nullptr;
  int series_cap_ = 0;
  std::vector<double> series_;
  int series_steps_ = 0;
  bool timing_ = false;
  int variant_ = 1;
  std::vector<Ev> ev_;       // event pairs of the current step() call
  size_t ev_used_ = 0;
  // batched step: device tables of patches, tiles (per tile shape), boundary
  // boxes and exchange boxes; interior remainders that no whole tile covers
  std::vector<PatchDesc> hpd_;
  PatchDesc* dpd_ = nullptr;
  std::vector<TileRef> htiles_[kNumTileShapes];
  TileRef* dtiles_[kNumTileShapes] = {};
  std::vector<std::pair<int, Box>> rem_[kNumTileShapes];
  // ghost boxes (fill boxes, then exchange boxes) and their work list: fill
  // entries [0, nfill_work_), exchange entries after them
  std::vector<GhostBox> hgbox_;
  GhostBox* dgbox_ = nullptr;
  GhostWork* dgwork_ = nullptr;
  int nfill_work_ = 0, nseg_work_ = 0;
  bool ghost_fused_ = true;  // no exchange source overlaps a fill box: one launch
  long maxn_ = 0;
  int ncu_ = 256;
  BatchTune tune_{0, 0, 1};
  bool tuned_ = false;
  bool capturing_ = false;
  struct Graph { int key; double dT; bool res; hipGraphExec_t exec; };
  std::vector<Graph> graphs_;
  double ms_[4] = {0, 0, 0, 0};
  long nl_[4] = {0, 0, 0, 0};
  std::string err_;
};

}  // namespace amss
