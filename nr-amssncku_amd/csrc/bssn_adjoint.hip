// bssn_adjoint.hip -- kernels of the discrete adjoint of one RK4 step
// (GAUGE 0, vacuum, Symmetry 0, frozen ghost zones of width 3):
//   * adj_feature_lds_kernel / adj_algebra_kernel: the stencil values of the
//     reverse algebra, then the reverse of the pointwise algebra -> feature bars,
//   * adj_gather_kernel: transposed stencils (first/second derivatives,
//     upwind advection, KO) -> dJ/d(input state) at every point,
//   * the forward re-evaluation of the step's stages, and the pointwise
//     reverse of the RK4 combination, chi's lower bound and enforce_ga.
#include "bssn_adjoint.h"
#include "bssn_adjoint_launch.h"
#include "bssn_kernels_impl.h"

namespace amss {

namespace {

__device__ __forceinline__ bool inside(const Box& b, int i, int j, int k) {
  return i >= b.lo[0] && i < b.hi[0] && j >= b.lo[1] && j < b.hi[1] && k >= b.lo[2] && k < b.hi[2];
}

// k only (no Christoffel / Ricci / residual outputs)
struct KSink {
  const FieldsOut& O;
  long p;
  __device__ __forceinline__ void put(int v, double k) { O.p[v][p] = k; }
  __device__ __forceinline__ void gam(int, double) {}
  __device__ __forceinline__ void ric(int, double) {}
  __device__ __forceinline__ void res(int, double) {}
};

__global__ void __launch_bounds__(kRhsThreads) rhs_k_kernel(FieldsIn in, FieldsOut out, Box box, Coef C) {
  int i, j, k;
  if (!interior_index(box, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  InteriorEnv E{in, in, C, p, box.sy, box.sz};
  KSink S{out, p};
  bssn_rhs_point<0, false, false, false>(E, S);
}

// rhs_k_kernel on 8x8x4 tiles with the forward kernel's LDS environment.
__global__ void __launch_bounds__(kRhsThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
rhs_k_lds_kernel(FieldsIn in, FieldsOut out, Box box, Coef C) {
  using Tl = LdsTile<kLdsTX, kLdsTY, kLdsTZ>;
  __shared__ __attribute__((aligned(16))) double tiles[kLdsBuffers * Tl::NPAD];
  const int tx = threadIdx.x % kLdsTX, ty = tile_row((threadIdx.x / kLdsTX) % kLdsTY, kLdsTY),
            tz = threadIdx.x / (kLdsTX * kLdsTY);
  const int gx = (box.hi[0] - box.lo[0]) / kLdsTX, gy = (box.hi[1] - box.lo[1]) / kLdsTY;
  const int b = xcd_spread(blockIdx.x, gridDim.x);
  const int i0 = box.lo[0] + (b % gx) * kLdsTX, j0 = box.lo[1] + ((b / gx) % gy) * kLdsTY;
  const int k0 = box.lo[2] + (b / (gx * gy)) * kLdsTZ;
  const long p = (i0 + tx) + (long)(j0 + ty) * box.sy + (long)(k0 + tz) * box.sz;
  lds_double* t0 = (lds_double*)tiles;
  LdsEnv<kLdsTX, kLdsTY, kLdsTZ> E{in, in, C, t0, t0 + (kLdsBuffers / 2) * Tl::NPAD, p,
                                  (tx + 3) + Tl::LX * ((ty + 3) + Tl::LY * (tz + 3)),
                                  (int)((i0 - 3) + (j0 - 3) * box.sy + (k0 - 3) * box.sz), (int)box.sy, (int)box.sz,
                                  box.n[0] * box.n[1] * box.n[2], (int)threadIdx.x};
  E.prologue();
  KSink S{out, p};
  bssn_rhs_point<0, false, false, false>(E, S);
}

// Run `lds` (tile grid size, tile-aligned core box) on the largest sub-box of
// whole 8x8x4 tiles and `direct` (box) on the remainder slabs; all direct
// when the row pitch is odd (the 16-byte halo loads need an even one).
template <class L, class D>
void split_tiles(const Box& box, L lds, D direct) {
  Box core = box;
  const int T[3] = {kLdsTX, kLdsTY, kLdsTZ};
  for (int a = 0; a < 3; ++a)
    core.hi[a] = box.sy % 2 == 0 ? box.lo[a] + (box.hi[a] - box.lo[a]) / T[a] * T[a] : box.lo[a];
  if (!box_empty(core)) {
    lds((unsigned)((core.hi[0] - core.lo[0]) / kLdsTX) * ((core.hi[1] - core.lo[1]) / kLdsTY) *
            ((core.hi[2] - core.lo[2]) / kLdsTZ),
        core);
  } else {
    for (int a = 0; a < 3; ++a) core.hi[a] = core.lo[a];
  }
  for (int a = 0; a < 3; ++a) {  // remainder slab [core.hi[a], box.hi[a]) along axis a
    Box rem = box;
    for (int c = 0; c < a; ++c) rem.hi[c] = core.hi[c];
    rem.lo[a] = core.hi[a];
    if (!box_empty(rem)) direct(rem);
  }
}

// Every stencil value the reverse algebra reads, at the interior points: a
// light kernel (high occupancy hides the stencil loads' latency), so the
// algebra kernel, which runs at one wave per SIMD, reads plain arrays.
__global__ void __launch_bounds__(256) adj_feature_kernel(FieldsIn in, FieldsIn lam, double* feat, long fs, Box box,
                                                          Coef C) {
  int i, j, k;
  if (!interior_index(box, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  InteriorEnv E{in, in, C, p, box.sy, box.sz};
  for (int v = 0; v < NFIELD; ++v) {
    feat[v * fs + p] = E.val(v);
    const int gs = adj_grad_slot(v), hs = adj_hess_slot(v);
    if (gs >= 0) {
      double d[3];
      E.grad(v, d);
#pragma unroll
      for (int a = 0; a < 3; ++a) feat[adj_grad_index(gs, a) * fs + p] = d[a];
    }
    if (hs >= 0) {
      double h[6];
      E.hess(v, h);
#pragma unroll
      for (int t = 0; t < 6; ++t) feat[adj_hess_index(hs, t) * fs + p] = h[t];
    }
  }
  double l[NFIELD], ub[3];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) l[v] = lam.p[v][p];
  upwind_beta_sums(E, l, ub);
#pragma unroll
  for (int a = 0; a < 3; ++a) feat[(kAdjFeatures + a) * fs + p] = ub[a];
}

// The same on 8x8x4 tiles with every field's halo tile staged in LDS (the
// forward kernel's environment: direct global->LDS loads, two field groups
// per barrier); tiles in XCD-contiguous order.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
adj_feature_lds_kernel(FieldsIn in, FieldsIn lam, double* feat, long fs, Box box, Coef C) {
  using Tl = LdsTile<kLdsTX, kLdsTY, kLdsTZ>;
  __shared__ __attribute__((aligned(16))) double tiles[kLdsBuffers * Tl::NPAD];
  const int tx = threadIdx.x % kLdsTX, ty = tile_row((threadIdx.x / kLdsTX) % kLdsTY, kLdsTY),
            tz = threadIdx.x / (kLdsTX * kLdsTY);
  const int gx = (box.hi[0] - box.lo[0]) / kLdsTX, gy = (box.hi[1] - box.lo[1]) / kLdsTY;
  const int b = xcd_spread(blockIdx.x, gridDim.x);
  const int i0 = box.lo[0] + (b % gx) * kLdsTX, j0 = box.lo[1] + ((b / gx) % gy) * kLdsTY;
  const int k0 = box.lo[2] + (b / (gx * gy)) * kLdsTZ;
  const long p = (i0 + tx) + (long)(j0 + ty) * box.sy + (long)(k0 + tz) * box.sz;
  const int N = box.n[0] * box.n[1] * box.n[2];
  lds_double* t0 = (lds_double*)tiles;
  // lam and beta at the point, loaded before the first halo loads (so no wait
  // for them stalls on those): every field's upwind term needs all three beta
  double l[NFIELD], bp[3];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) l[v] = lam.p[v][p];
#pragma unroll
  for (int a = 0; a < 3; ++a) bp[a] = in.p[BETAX + a][p];
  __builtin_amdgcn_sched_barrier(0);
  LdsEnv<kLdsTX, kLdsTY, kLdsTZ> E{in, in, C, t0, t0 + (kLdsBuffers / 2) * Tl::NPAD, p,
                                  (tx + 3) + Tl::LX * ((ty + 3) + Tl::LY * (tz + 3)),
                                  (int)((i0 - 3) + (j0 - 3) * box.sy + (k0 - 3) * box.sz), (int)box.sy, (int)box.sz, N,
                                  (int)threadIdx.x};
  E.prologue();
  double ub[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int g = 0; g < NFIELD; ++g) {
    const int v = group_field(g);
    E.begin(v);
    feat[v * fs + p] = E.val(v);
    const int gs = adj_grad_slot(v), hs = adj_hess_slot(v);
    // (fences between the stencil kinds: issued all at once, a group's LDS
    // reads would need more registers than there are)
    if (gs >= 0) {
      double d[3];
      E.grad(v, d);
#pragma unroll
      for (int a = 0; a < 3; ++a) feat[adj_grad_index(gs, a) * fs + p] = d[a];
      amss_fence<decltype(E)>();
    }
    if (hs >= 0) {
      double h[6];
      E.hess(v, h);
#pragma unroll
      for (int t = 0; t < 6; ++t) feat[adj_hess_index(hs, t) * fs + p] = h[t];
      amss_fence<decltype(E)>();
    }
    double up[3], dn[3];
    E.updn(v, up, dn);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ub[a] += l[v] * upwind_dbeta(bp[a], up[a], dn[a]);
      asm volatile("" : "+v"(ub[a]));  // summed here, not deferred (that keeps 144 stencil values live)
    }
    E.end();
    amss_fence<decltype(E)>();
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) feat[(kAdjFeatures + a) * fs + p] = ub[a];
}

static_assert(kAdjStride == kRhsThreads, "the algebra kernel's stash is strided by its workgroup size");
__global__ void __launch_bounds__(kAdjStride) adj_algebra_kernel(const double* feat, long fs, FieldsIn lam,
                                                          double* bars, long stride, Box box, Coef C) {
  int i, j, k;
  if (!interior_index(box, i, j, k)) return;
  const long p = i + j * box.sy + k * box.sz;
  const FeatureEnv E{feat, fs, p, C};
  double l[NFIELD];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) l[v] = lam.p[v][p];
  const AdjSink O{bars, stride, p};
  __shared__ double stash[kAdjStash * kAdjStride];
  bssn_rhs_adjoint_point(E, l, O, (lds_double*)stash + threadIdx.x);
}

// Loads of one array at q + o for uniform offsets o in [-3 sz, 3 sz]: a
// buffer resource based 3 planes before element 0, the thread's q in the
// vector offset and o in the scalar one (one instruction per load).
struct Taps {
  __amdgpu_buffer_rsrc_t r;
  unsigned vo;
  int bias;
  __device__ __forceinline__ Taps(const double* f, long q, int sz3)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)(f - sz3), (short)0, (int)0xFFFFFFFF, 0x00020000)),
        vo((unsigned)(q * 8)),
        bias(sz3) {}
  __device__ __forceinline__ double operator()(int o) const {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, (unsigned)((o + bias) * 8), 0));
  }
};

// dJ/dU(q) = value bar + sum over the output points p whose stencils read q
// of (stencil weight) x (feature bar at p).  The bars are zero outside the
// interior and readable 3 planes beyond the array, so they are read at every
// offset unchecked; lam likewise when LAMPAD (else lam, and beta with it, is
// read only at interior points).  A workgroup is 256 consecutive points of
// one group of 4 fields (blockIdx.y): the upwind weights of beta at the 21
// output points a field's advection term reads are formed once per point and
// used for all 4 fields; the groups mix fields with and without Hessians so
// that they cost about the same.
constexpr int kGatherGroups = 6, kGatherPer = 4;
__constant__ signed char kGatherField[kGatherGroups][kGatherPer] = {
    {CHI, DXX, TRK, AXX}, {GXY, GXZ, GAMX, AXY}, {DYY, GYZ, GAMY, AXZ},
    {DZZ, LAP, GAMZ, AYY}, {BETAX, BETAX + 1, AYZ, AZZ}, {BETAX + 2, DTSFX, DTSFX + 1, DTSFX + 2}};
template <bool LAMPAD>
__global__ void __launch_bounds__(256) adj_gather_kernel(FieldsIn in, FieldsIn lam, const double* __restrict__ bars,
                                                         long bs, FieldsOut out, Box box, Coef C, int accumulate) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= N) return;
  const int i = (int)(q % box.n[0]), j = (int)((q / box.n[0]) % box.n[1]), k = (int)(q / ((long)box.n[0] * box.n[1]));
  const int sy = (int)box.sy, sz = (int)box.sz, sz3 = 3 * sz;
  const int st[3] = {1, sy, sz};
  // advection + KO weights at p = q - o e_a: d1 (up(o) max(beta,0) - up(-o) min(beta,0)) + ko kw(o)
  constexpr double up[7] = {0.0, 0.0, -3.0, -10.0, 18.0, -6.0, 1.0};  // offsets -3..3 from p
  constexpr double kw[7] = {1.0, -6.0, 15.0, -20.0, 15.0, -6.0, 1.0};
  double wt[3][7];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const Taps B(in.p[BETAX + a], q, sz3);
#pragma unroll
    for (int o = -3; o <= 3; ++o) {
      bool use = true;
      if (!LAMPAD) {
        const int pi = i - (a == 0 ? o : 0), pj = j - (a == 1 ? o : 0), pk = k - (a == 2 ? o : 0);
        use = inside(box, pi, pj, pk);
      }
      const double b = use ? B(-o * st[a]) : 0.0;
      wt[a][o + 3] = use ? C.d1[a] * (up[o + 3] * fmax(b, 0.0) - up[3 - o] * fmin(b, 0.0)) + C.ko[a] * kw[o + 3] : 0.0;
    }
  }
  for (int f = 0; f < kGatherPer; ++f) {
    const int v = kGatherField[blockIdx.y][f];
    double acc = bars[v * bs + q];
    const int gs = adj_grad_slot(v);
    if (gs >= 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const Taps b(bars + adj_grad_index(gs, a) * bs, q, sz3);
        const int s = st[a];
        acc += C.d1[a] * ((b(2 * s) - b(-2 * s)) + 8.0 * (b(-s) - b(s)));
      }
    }
    const int hs = adj_hess_slot(v);
    if (hs >= 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {  // diagonal second derivatives
        const Taps b(bars + adj_hess_index(hs, a == 0 ? 0 : (a == 1 ? 3 : 5)) * bs, q, sz3);
        const int s = st[a];
        acc += C.d2[a] * (16.0 * (b(-s) + b(s)) - (b(-2 * s) + b(2 * s)) - 30.0 * b(0));
      }
      const int ta[3] = {0, 0, 1}, tb[3] = {1, 2, 2}, tt[3] = {1, 2, 4};
      constexpr int os[4] = {-2, -1, 1, 2};
      constexpr double w[4] = {1.0, -8.0, 8.0, -1.0};  // d1 weights at offsets -2, -1, 1, 2
#pragma unroll
      for (int m = 0; m < 3; ++m) {  // mixed: the 4 x 4 cross stencil is its own transpose
        const Taps b(bars + adj_hess_index(hs, tt[m]) * bs, q, sz3);
        const int sa = st[ta[m]], sb = st[tb[m]];
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          double r = 0.0;
#pragma unroll
          for (int y = 0; y < 4; ++y) r += w[y] * b(os[x] * sa + os[y] * sb);
          s += w[x] * r;
        }
        acc += C.dm[m] * s;
      }
    }
    // advection and KO read lam at the output points p = q - o e_a (the KO
    // stencil is symmetric, so its transpose uses the same points)
    const Taps L(lam.p[v], q, sz3);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int o = -3; o <= 3; ++o) {
        if (!LAMPAD) {
          const int pi = i - (a == 0 ? o : 0), pj = j - (a == 1 ? o : 0), pk = k - (a == 2 ? o : 0);
          if (!inside(box, pi, pj, pk)) continue;
        }
        acc += wt[a][o + 3] * L(-o * st[a]);
      }
    }
    if (accumulate)
      out.p[v][q] += acc;
    else
      out.p[v][q] = acc;
  }
}

// ---- enforce_ga and its vector-Jacobian product ------------------------------
__device__ void enforce_vjp(const double din[6], const double ain[6], double gb[6], double ab[6]) {
  constexpr double W[6] = {1.0, 2.0, 2.0, 1.0, 2.0, 1.0};
  double g[6];
  for (int s = 0; s < 6; ++s) g[s] = din[s] + (sdiag(s) ? 1.0 : 0.0);
  const double det = g[0] * g[3] * g[5] + g[1] * g[4] * g[2] + g[2] * g[1] * g[4] - g[2] * g[3] * g[2] -
                     g[1] * g[1] * g[5] - g[0] * g[4] * g[4];
  const double sc = 1.0 / pow(det, 1.0 / 3.0);
  double p[6];
  for (int s = 0; s < 6; ++s) p[s] = g[s] * sc;
  const double c[6] = {p[3] * p[5] - p[4] * p[4], p[2] * p[4] - p[1] * p[5], p[1] * p[4] - p[3] * p[2],
                       p[0] * p[5] - p[2] * p[2], p[1] * p[2] - p[0] * p[4], p[0] * p[3] - p[1] * p[1]};
  double trA = 0.0;
  for (int t = 0; t < 6; ++t) trA += W[t] * c[t] * ain[t];
  // outputs: d'_s = p_s - diag, a'_s = a_s - p_s trA / 3
  double pb[6], abo[6], trA_b = 0.0;
  for (int s = 0; s < 6; ++s) {
    pb[s] = gb[s] - (1.0 / 3.0) * trA * ab[s];
    trA_b -= (1.0 / 3.0) * p[s] * ab[s];
    abo[s] = ab[s];
  }
  double cb[6];
  for (int t = 0; t < 6; ++t) {
    abo[t] += trA_b * W[t] * c[t];
    cb[t] = trA_b * W[t] * ain[t];
  }
  pb[3] += cb[0] * p[5]; pb[5] += cb[0] * p[3]; pb[4] -= 2.0 * cb[0] * p[4];
  pb[2] += cb[1] * p[4]; pb[4] += cb[1] * p[2]; pb[1] -= cb[1] * p[5]; pb[5] -= cb[1] * p[1];
  pb[1] += cb[2] * p[4]; pb[4] += cb[2] * p[1]; pb[3] -= cb[2] * p[2]; pb[2] -= cb[2] * p[3];
  pb[0] += cb[3] * p[5]; pb[5] += cb[3] * p[0]; pb[2] -= 2.0 * cb[3] * p[2];
  pb[1] += cb[4] * p[2]; pb[2] += cb[4] * p[1]; pb[0] -= cb[4] * p[4]; pb[4] -= cb[4] * p[0];
  pb[0] += cb[5] * p[3]; pb[3] += cb[5] * p[0]; pb[1] -= 2.0 * cb[5] * p[1];
  // p = g det(g)^(-1/3): d det / d g_t = W_t cof_t(g)
  const double C0[6] = {g[3] * g[5] - g[4] * g[4], g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[3] * g[2],
                        g[0] * g[5] - g[2] * g[2], g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] * g[1]};
  double dot = 0.0;
  for (int s = 0; s < 6; ++s) dot += pb[s] * g[s];
  for (int t = 0; t < 6; ++t) {
    gb[t] = sc * pb[t] - dot * (1.0 / 3.0) * (sc / det) * W[t] * C0[t];
    ab[t] = abo[t];
  }
}

// Forward stage s of the step at every point:
//   interior: acc (s = 0: k; 1, 2: += 2k), z = Y0 + c k (s = 3: Y0 + c (k + acc)),
//             Ynext = LB(z), enforce_ga for s < 3;
//   boundary: Ynext = the frozen-ghost fill of Y0 (LB on chi, enforce_ga for s < 3).
__global__ void __launch_bounds__(256) fwd_stage_kernel(FieldsIn Y0, FieldsIn kk, FieldsOut acc, FieldsOut z,
                                                        FieldsOut Yn, Box box, int stage, double c, double tinny) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= N) return;
  const int i = (int)(q % box.n[0]), j = (int)((q / box.n[0]) % box.n[1]), k = (int)(q / ((long)box.n[0] * box.n[1]));
  double u[NFIELD];
  if (inside(box, i, j, k)) {
    for (int v = 0; v < NFIELD; ++v) {
      const double kv = kk.p[v][q];
      double w;
      if (stage == 0) {
        acc.p[v][q] = kv;
        w = Y0.p[v][q] + c * kv;
      } else if (stage < 3) {
        acc.p[v][q] = acc.p[v][q] + 2.0 * kv;
        w = Y0.p[v][q] + c * kv;
      } else {
        w = Y0.p[v][q] + c * (kv + acc.p[v][q]);
      }
      z.p[v][q] = w;
      u[v] = w;
    }
  } else {
    for (int v = 0; v < NFIELD; ++v) u[v] = Y0.p[v][q];
  }
  u[CHI] = u[CHI] < tinny ? tinny : u[CHI];
  if (stage < 3) enforce_point(u + DXX, u + AXX);
  for (int v = 0; v < NFIELD; ++v) Yn.p[v][q] = u[v];
}

// Reverse of the last combination: lam = dJ/dU^{n+1} ->
//   interior: lz = lam (chi masked by the lower bound), lY0 = lz,
//             lk3 = c lz, lk0 = c lz, lk1 = lk2 = 2 c lz;
//   boundary: lY0 = lam through the stage-3 fill (chi masked), lk = 0.
__global__ void __launch_bounds__(256) adj_final_kernel(FieldsIn lam, FieldsIn z3, FieldsIn Y0, FieldsOut lY0,
                                                        FieldsOut lk0, FieldsOut lk1, FieldsOut lk2, FieldsOut lk3,
                                                        Box box, double c, double tinny) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= N) return;
  const int i = (int)(q % box.n[0]), j = (int)((q / box.n[0]) % box.n[1]), k = (int)(q / ((long)box.n[0] * box.n[1]));
  const bool in = inside(box, i, j, k);
  for (int v = 0; v < NFIELD; ++v) {
    double l = lam.p[v][q];
    if (v == CHI && (in ? z3.p[CHI][q] : Y0.p[CHI][q]) < tinny) l = 0.0;
    lY0.p[v][q] = l;
    lk0.p[v][q] = in ? c * l : 0.0;
    lk1.p[v][q] = in ? 2.0 * c * l : 0.0;
    lk2.p[v][q] = in ? 2.0 * c * l : 0.0;
    lk3.p[v][q] = in ? c * l : 0.0;
  }
}

// Reverse of stage s < 3 given lY = dJ/dY_{s+1} (RHS adjoint of stage s+1):
//   interior: Y = E(LB(z)): lz = LB' E'^T lY; lY0 += lz; lk_s += c lz;
//   boundary: Y = fill(Y0): lY0 += LB' E'(Y0)^T lY.
__global__ void __launch_bounds__(256) adj_stage_kernel(FieldsIn lY, FieldsIn z, FieldsIn Y0, FieldsOut lY0,
                                                        FieldsOut lk, Box box, double c, double tinny) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= N) return;
  const int i = (int)(q % box.n[0]), j = (int)((q / box.n[0]) % box.n[1]), k = (int)(q / ((long)box.n[0] * box.n[1]));
  const bool in = inside(box, i, j, k);
  const FieldsIn& W = in ? z : Y0;
  double l[NFIELD], w[NFIELD];
  for (int v = 0; v < NFIELD; ++v) {
    l[v] = lY.p[v][q];
    w[v] = W.p[v][q];
  }
  enforce_vjp(w + DXX, w + AXX, l + DXX, l + AXX);
  if (w[CHI] < tinny) l[CHI] = 0.0;
  for (int v = 0; v < NFIELD; ++v) {
    lY0.p[v][q] += l[v];
    if (in) lk.p[v][q] += c * l[v];
  }
}

// dJ/dU^n = E'(U^n)^T lY0 (the step's opening enforce_ga on every point).
__global__ void __launch_bounds__(256) adj_open_kernel(FieldsIn U, FieldsIn lY0, FieldsOut lam, long N) {
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= N) return;
  double l[NFIELD], d[6], a[6];
  for (int v = 0; v < NFIELD; ++v) l[v] = lY0.p[v][q];
  for (int s = 0; s < 6; ++s) {
    d[s] = U.p[DXX + s][q];
    a[s] = U.p[AXX + s][q];
  }
  enforce_vjp(d, a, l + DXX, l + AXX);
  for (int v = 0; v < NFIELD; ++v) lam.p[v][q] = l[v];
}

__global__ void __launch_bounds__(256) copy_enforce_kernel(FieldsIn U, FieldsOut Y, long N) {
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= N) return;
  double u[NFIELD];
  for (int v = 0; v < NFIELD; ++v) u[v] = U.p[v][q];
  enforce_point(u + DXX, u + AXX);
  for (int v = 0; v < NFIELD; ++v) Y.p[v][q] = u[v];
}

// 24 fields src -> dst (blockIdx.y = field): one launch for a whole state
__global__ void __launch_bounds__(256) copy_fields_kernel(FieldsIn src, FieldsOut dst, long N) {
  const double* __restrict__ a = src.p[blockIdx.y];
  double* __restrict__ b = dst.p[blockIdx.y];
  for (long q = blockIdx.x * 256L + threadIdx.x; q < N; q += gridDim.x * 256L) b[q] = a[q];
}

unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void adj_launch_rhs_k(const FieldsIn& in, const FieldsOut& out, const Box& box, const Coef& C, hipStream_t s) {
  split_tiles(
      box, [&](unsigned gr, const Box& core) { rhs_k_lds_kernel<<<gr, kRhsThreads, 0, s>>>(in, out, core, C); },
      [&](const Box& rem) { rhs_k_kernel<<<interior_grid(rem), kRhsThreads, 0, s>>>(in, out, rem, C); });
}
void adj_launch_point(const FieldsIn& in, const FieldsIn& lam, double* bars, long stride, const Box& box,
                      const Coef& C, hipStream_t s, double* feat) {
  const long fs = (long)box.n[0] * box.n[1] * box.n[2];
  split_tiles(
      box, [&](unsigned gr, const Box& core) { adj_feature_lds_kernel<<<gr, 256, 0, s>>>(in, lam, feat, fs, core, C); },
      [&](const Box& rem) { adj_feature_kernel<<<interior_grid(rem), 256, 0, s>>>(in, lam, feat, fs, rem, C); });
  adj_algebra_kernel<<<interior_grid(box), kAdjStride, 0, s>>>(feat, fs, lam, bars, stride, box, C);
}

void adj_launch_gather(const FieldsIn& in, const FieldsIn& lam, bool lam_padded, const double* bars, long stride,
                       const FieldsOut& out, const Box& box, const Coef& C, bool accumulate, hipStream_t s) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  const dim3 g(nblk(N), kGatherGroups);
  if (lam_padded)
    adj_gather_kernel<true><<<g, 256, 0, s>>>(in, lam, bars, stride, out, box, C, accumulate ? 1 : 0);
  else
    adj_gather_kernel<false><<<g, 256, 0, s>>>(in, lam, bars, stride, out, box, C, accumulate ? 1 : 0);
}
void adj_launch_fwd_stage(const FieldsIn& Y0, const FieldsIn& k, const FieldsOut& acc, const FieldsOut& z,
                          const FieldsOut& Yn, const Box& box, int stage, double c, double tinny, hipStream_t s) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  fwd_stage_kernel<<<nblk(N), 256, 0, s>>>(Y0, k, acc, z, Yn, box, stage, c, tinny);
}
void adj_launch_final(const FieldsIn& lam, const FieldsIn& z3, const FieldsIn& Y0, const FieldsOut& lY0,
                      const FieldsOut* lk, const Box& box, double c, double tinny, hipStream_t s) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  adj_final_kernel<<<nblk(N), 256, 0, s>>>(lam, z3, Y0, lY0, lk[0], lk[1], lk[2], lk[3], box, c, tinny);
}
void adj_launch_stage(const FieldsIn& lY, const FieldsIn& z, const FieldsIn& Y0, const FieldsOut& lY0,
                      const FieldsOut& lk, const Box& box, double c, double tinny, hipStream_t s) {
  const long N = (long)box.n[0] * box.n[1] * box.n[2];
  adj_stage_kernel<<<nblk(N), 256, 0, s>>>(lY, z, Y0, lY0, lk, box, c, tinny);
}
void adj_launch_open(const FieldsIn& U, const FieldsIn& lY0, const FieldsOut& lam, long N, hipStream_t s) {
  adj_open_kernel<<<nblk(N), 256, 0, s>>>(U, lY0, lam, N);
}
void adj_launch_copy_fields(const FieldsIn& src, const FieldsOut& dst, long N, hipStream_t s) {
  const unsigned bx = nblk(N) < 1024u ? nblk(N) : 1024u;
  copy_fields_kernel<<<dim3(bx, NFIELD), 256, 0, s>>>(src, dst, N);
}
void adj_launch_copy_enforce(const FieldsIn& U, const FieldsOut& Y, long N, hipStream_t s) {
  copy_enforce_kernel<<<nblk(N), 256, 0, s>>>(U, Y, N);
}

}  // namespace amss
