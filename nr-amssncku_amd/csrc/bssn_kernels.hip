// bssn_kernels.hip -- gfx950 kernels of the BSSN RHS + RK4 path.
//
//  The point kernels (RHS, RHS fused with RK4) live in bssn_kernels_impl.h
//  and are instantiated per gauge in bssn_gauge0.hip / bssn_gauge2.hip.
//  This file holds the small pointwise kernels and the dispatch:
//  * ghost_fill : the boundary fill of a stage output; seg_copy: ghost exchange.
//  * enforce, lowerbound, rk4_update, copy_box, l2 partial sums.
#include <algorithm>

#include "amss_interp.h"
#include "bssn_kernels_impl.h"

namespace amss {

// Boundary fill after an RK4 sub-step: every point within w.lo[a] of the low
// face or w.hi[a] of the high face of axis a takes the sub-step's base state
// U0, then the lower bound on chi and, for sub-steps 0-2, the enforce_ga the
// next RHS applies to its input -- exactly the sequence rk4 -> copy ->
// lowerboundset -> enforce_ga the reference runs on such points
// (bssn_class.C:3119-3156, 3415-3525).
// The shell outside [w.lo, n - w.hi) is cut into at most 6 disjoint slabs
// (both z faces whole, then the y faces of what is left, then the x faces);
// blockIdx.y picks the slab.  Every field is loaded before any is stored, so a
// wave has all its loads in flight at once.
__device__ __forceinline__ bool face_slab(const Box& box, const Faces& w, int f, int lo[3], int ext[3]) {
  int l[3] = {0, 0, 0}, h[3] = {box.n[0], box.n[1], box.n[2]};
  for (int a = 2; a >= 0; --a) {
    for (int side = 0; side < 2; ++side, --f) {
      if (f != 0) continue;
      for (int c = 0; c < 3; ++c) {
        lo[c] = l[c];
        ext[c] = h[c] - l[c];
      }
      lo[a] = side == 0 ? 0 : box.n[a] - w.hi[a];
      ext[a] = side == 0 ? w.lo[a] : w.hi[a];
      return ext[0] > 0 && ext[1] > 0 && ext[2] > 0;
    }
    l[a] = w.lo[a];
    h[a] = box.n[a] - w.hi[a];
  }
  return false;
}

__global__ void __launch_bounds__(256) ghost_fill_kernel(FieldsIn u0, FieldsOut out, Box box, Faces w, int stage,
                                                         double tinny, int* flag) {
  int lo[3], ext[3];
  if (!face_slab(box, w, blockIdx.y, lo, ext)) return;
  const unsigned npts = (unsigned)ext[0] * ext[1] * ext[2];
  for (unsigned t = blockIdx.x * 256u + threadIdx.x; t < npts; t += gridDim.x * 256u) {
    const int i = lo[0] + (int)(t % ext[0]);
    const int j = lo[1] + (int)((t / ext[0]) % ext[1]);
    const int k = lo[2] + (int)(t / ((unsigned)ext[0] * ext[1]));
    const long p = i + j * box.sy + k * box.sz;
    double u[NFIELD];
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) u[v] = u0.p[v][p];
    bool bad = false;
#pragma unroll
    for (int v = 0; v < 21; ++v) bad |= !(u[v] == u[v]);  // bssn_rhs.f90:110-113
    if (bad) *flag = 1;
    if (stage < 3) enforce_point(u + DXX, u + AXX);
    u[CHI] = u[CHI] < tinny ? tinny : u[CHI];
#pragma unroll
    for (int v = 0; v < NFIELD; ++v) out.p[v][p] = u[v];
  }
}

// Ghost exchange between patches of one level: copy all 24 fields of a box
// from the source patch into the destination patch (Parallel::Sync for one
// rank: transfer + data_packer, Parallel.C:3336-3423, 3508-3591, 3677-3775).
__global__ void seg_copy_kernel(FieldsIn src, FieldsOut dst, SegDesc d) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long N = (long)d.ext[0] * d.ext[1] * d.ext[2];
  if (t >= N) return;
  const int a = t % d.ext[0], b = (t / d.ext[0]) % d.ext[1], c = t / ((long)d.ext[0] * d.ext[1]);
  const long ps = (d.slo[0] + a) + (d.slo[1] + b) * d.ssy + (d.slo[2] + c) * d.ssz;
  const long pd = (d.dlo[0] + a) + (d.dlo[1] + b) * d.dsy + (d.dlo[2] + c) * d.dsz;
  double u[NFIELD];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) u[v] = src.p[v][ps];
#pragma unroll
  for (int v = 0; v < NFIELD; ++v) dst.p[v][pd] = u[v];
}

__global__ void enforce_kernel(FieldsOut f, long n) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  double ga[12];
#pragma unroll
  for (int s = 0; s < 12; ++s) ga[s] = f.p[DXX + s][p];
  enforce_point(ga, ga + 6);
#pragma unroll
  for (int s = 0; s < 12; ++s) f.p[DXX + s][p] = ga[s];
}

__global__ void lowerbound_kernel(double* chi, long n, double tinny) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double c = chi[p];
  if (c < tinny) chi[p] = tinny;
}

// rungekutta4_rout on one field (rungekutta4_rout.f90:94-139).
__global__ void rk4_update_kernel(const double* __restrict__ f0, double* __restrict__ f1, double* __restrict__ frhs,
                                  long n, double dT, int stage) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (stage == 0) {
    f1[p] = f0[p] + (0.5 * dT) * frhs[p];
  } else if (stage == 1) {
    const double k = f1[p];
    frhs[p] = frhs[p] + 2.0 * k;
    f1[p] = f0[p] + (0.5 * dT) * k;
  } else if (stage == 2) {
    const double k = f1[p];
    frhs[p] = frhs[p] + 2.0 * k;
    f1[p] = f0[p] + dT * k;
  } else {
    f1[p] = f0[p] + ((1.0 / 6.0) * dT) * (f1[p] + frhs[p]);
  }
}

// f = f0 on the box [lo, hi] (inclusive) -- one face layer of sommerfeld_rout.
__global__ void copy_box_kernel(const double* __restrict__ f0, double* __restrict__ f, Box box) {
  const int nx = box.hi[0] - box.lo[0] + 1, ny = box.hi[1] - box.lo[1] + 1, nz = box.hi[2] - box.lo[2] + 1;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nx * ny * nz) return;
  const int i = box.lo[0] + t % nx, j = box.lo[1] + (t / nx) % ny, k = box.lo[2] + t / ((long)nx * ny);
  const long p = i + j * box.sy + k * box.sz;
  f[p] = f0[p];
}

// Partial sums of f^2 over the box [lo, hi) per block (fmisc.f90:1263-1309).
__global__ void __launch_bounds__(256) sumsq_kernel(const double* __restrict__ f, Box box, double* __restrict__ part) {
  __shared__ double red[256];
  const int nx = box.hi[0] - box.lo[0], ny = box.hi[1] - box.lo[1], nz = box.hi[2] - box.lo[2];
  const long N = (long)nx * ny * nz;
  double s = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N; t += (long)gridDim.x * blockDim.x) {
    const int i = box.lo[0] + t % nx, j = box.lo[1] + (t / nx) % ny, k = box.lo[2] + t / ((long)nx * ny);
    const double v = f[i + j * box.sy + k * box.sz];
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// prepare_inter_time_level (Parallel.C:4329-4441; average / average2 of
// fmisc.f90:1532-1592) on the 24 fields of one patch:
// out = c1 f1 + c2 f2 (+ c3 f3 when three).
__global__ void __launch_bounds__(256) time_interp_kernel(FieldsIn f1, FieldsIn f2, FieldsIn f3, FieldsOut out,
                                                          long n, double c1, double c2, double c3, int three) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int v = blockIdx.y;
  double r = c1 * f1.p[v][p] + c2 * f2.p[v][p];
  if (three) r = r + c3 * f3.p[v][p];
  out.p[v][p] = r;
}

// Partial sums of squares of the L2 jobs: workgroup x of job y covers every
// (nparts * 256)-th point of the job's box, starting at x * 256 + lane.
__global__ void __launch_bounds__(256) l2_parts_kernel(const L2Job* __restrict__ jobs, double* __restrict__ part) {
  const L2Job& J = jobs[blockIdx.y];
  if ((int)blockIdx.x >= J.nparts) return;  // the whole workgroup
  __shared__ double red[256];
  const int nx = J.hi[0] - J.lo[0], ny = J.hi[1] - J.lo[1], nz = J.hi[2] - J.lo[2];
  const long N = (long)nx * ny * nz;
  double s = 0.0;
  const long stride = (long)J.nparts * 256;
  if (N + stride <= 0xFFFFFFFFL) {
    // the same points in the same order, with 32-bit index arithmetic
    const unsigned nxu = (unsigned)nx, nyu = (unsigned)ny, nxy = nxu * nyu;
    for (unsigned t = blockIdx.x * 256u + threadIdx.x; t < (unsigned)N; t += (unsigned)stride) {
      const unsigned r = t / nxu;
      const int i = J.lo[0] + (int)(t - r * nxu), j = J.lo[1] + (int)(r % nyu), k = J.lo[2] + (int)(t / nxy);
      const double v = J.f[i + j * J.sy + k * J.sz];
      s += v * v;
    }
  } else {
    for (long t = blockIdx.x * 256L + threadIdx.x; t < N; t += stride) {
      const int i = J.lo[0] + (int)(t % nx), j = J.lo[1] + (int)((t / nx) % ny), k = J.lo[2] + (int)(t / ((long)nx * ny));
      const double v = J.f[i + j * J.sy + k * J.sz];
      s += v * v;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[J.part0 + blockIdx.x] = red[0];
}

// One wave per job: lane t sums partials t, t + 64, ... in order, then a
// fixed pairwise tree (the same sum on every run).
__global__ void __launch_bounds__(64) l2_final_kernel(const L2Job* __restrict__ jobs, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  const L2Job& J = jobs[blockIdx.x];
  __shared__ double red[64];
  double s = 0.0;
  for (int b = threadIdx.x; b < J.nparts; b += 64) s += part[J.part0 + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] * J.dv;
}

// --------------------------------------------------------------------------
// Host-side launchers (stream-ordered, no synchronisation).

void launch_rhs(const FieldsIn& in, const FieldsIn& mat, const FieldsOut& out, const AuxOut& aux, const Box& box,
                const Coef& C, int gauge, int co, int* flag, bool edge, hipStream_t s) {
  if (gauge == 2) launch_rhs_g2(in, mat, out, aux, box, C, co, flag, edge, s);
  else launch_rhs_g0(in, mat, out, aux, box, C, co, flag, edge, s);
}

void launch_rk4(const FieldsIn& in, const FieldsIn& u0, const FieldsOut& acc, const FieldsOut& out,
                const AuxOut& aux, const Box& box, const Coef& C, const Rk4Param& R, int gauge, bool want_res,
                int* flag, bool edge, hipStream_t s, bool interior, const SommerGeom& sg) {
  if (gauge == 2) launch_rk4_g2(in, u0, acc, out, aux, box, C, R, want_res, flag, edge, s, interior, sg);
  else launch_rk4_g0(in, u0, acc, out, aux, box, C, R, want_res, flag, edge, s, interior, sg);
}

// sommerfeld_routbam (sommerfeld_rout.f90:223-449) on one field: a thread per
// array point; points off the face layers return at once.
__global__ void __launch_bounds__(256) sommer_bam_kernel(const double* __restrict__ f0, double* __restrict__ frhs,
                                                         Shape n, SommerGeom sg, double vel, double S0, double S1,
                                                         double S2) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long N = (long)n.n[0] * n.n[1] * n.n[2];
  if (t >= N) return;
  const int p[3] = {(int)(t % n.n[0]), (int)((t / n.n[0]) % n.n[1]), (int)(t / ((long)n.n[0] * n.n[1]))};
  if (!sommer_point(sg, n.n, p[0], p[1], p[2])) return;
  const double S[3] = {S0, S1, S2};
  auto rd = [&](int a, int m) {
    int q[3] = {p[0], p[1], p[2]};
    q[a] += m;
    double sgn = 1.0;
    if (q[a] < 0) {  // the one mirrored layer below a symmetry plane
      q[a] = -q[a] - 1;
      sgn = S[a];
    }
    return f0[q[0] + q[1] * n.sy + q[2] * n.sz] * sgn;
  };
  frhs[t] = sommer_rhs(sg, n.n, p, vel, rd);
}

// sommerfeld_rout with precor == 0 (sommerfeld_rout.f90:141-217) on one
// field: each point of the face layers takes the outgoing wave
// (1 - r) f0(x - r x), r = alpha sqrt(chi) dT / |x|, f0 interpolated over the
// 6^3 points around x - r x (shifted inside the array where it would leave
// it, mirrored with the parity S below a symmetry plane: decide3d,
// fmisc.f90:798-874), Neville along z, then y, then x (polin3).  A thread per
// array point; a point whose block would still leave the array is left alone
// and raises *bad.
__global__ void __launch_bounds__(256) sommer_interp_kernel(const double* __restrict__ X, const double* __restrict__ Y,
                                                            const double* __restrict__ Z,
                                                            const double* __restrict__ chi0,
                                                            const double* __restrict__ lap0,
                                                            const double* __restrict__ f0, double* __restrict__ f,
                                                            Shape n, int faces, double dT, double S0, double S1,
                                                            double S2, int symmetry, int* __restrict__ bad) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long N = (long)n.n[0] * n.n[1] * n.n[2];
  if (t >= N) return;
  const int p[3] = {(int)(t % n.n[0]), (int)((t / n.n[0]) % n.n[1]), (int)(t / ((long)n.n[0] * n.n[1]))};
  if (!on_faces(faces, n.n, p[0], p[1], p[2])) return;
  const double xc[3] = {X[p[0]], Y[p[1]], Z[p[2]]};
  const double h[3] = {X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0]};
  const double r = (lap0[t] + 1.0) * sqrt(1.0 + chi0[t]) * dT / sqrt(xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
  // lo[a]: first array index (0-based) of the 6-point block; u[a]: the
  // target's position inside it (block index units)
  int lo[3];
  double u[3];
  for (int a = 0; a < 3; ++a) {
    const double cx = r * xc[a] / h[a];
    const double ti = trunc(cx);
    lo[a] = p[a] - (int)ti - (cx > 0.0 ? 3 : 2);
    u[a] = ti - cx + (cx > 0.0 ? 3.0 : 2.0);
    // a block reaching below the array moves up unless the axis mirrors there
    const bool mirrors = a == 2 ? symmetry != 0 : symmetry == 2;
    if (!mirrors && lo[a] < 0) {
      u[a] += lo[a];
      lo[a] = 0;
    }
  }
  for (int a = 0; a < 3; ++a)
    if (lo[a] + 5 > n.n[a] - 1) {
      const int sh = lo[a] + 5 - (n.n[a] - 1);
      u[a] += sh;
      lo[a] -= sh;
    }
  for (int a = 0; a < 3; ++a)  // decide3d's range checks (mirrored reads included)
    if (lo[a] + 5 > n.n[a] - 1 || -lo[a] - 1 > n.n[a] - 1) {
      *bad = 1;
      return;
    }
  const double S[3] = {S0, S1, S2};
  double ym[6];
#pragma unroll 1
  for (int a = 0; a < 6; ++a) {
    double yz[6];
#pragma unroll 1
    for (int b = 0; b < 6; ++b) {
      double line[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        int q[3] = {lo[0] + a, lo[1] + b, lo[2] + c};
        double sgn = 1.0;
#pragma unroll
        for (int m = 0; m < 3; ++m)
          if (q[m] < 0) {
            q[m] = -q[m] - 1;
            sgn *= S[m];
          }
        line[c] = f0[q[0] + q[1] * n.sy + q[2] * n.sz] * sgn;
      }
      yz[b] = neville6(line, u[2]);
    }
    ym[a] = neville6(yz, u[1]);
  }
  f[t] = neville6(ym, u[0]) * (1.0 - r);
}

void launch_sommer_interp(const double* X, const double* Y, const double* Z, const double* chi0, const double* lap0,
                          const double* f0, double* f, const Shape& n, int faces, double dT, const double S[3],
                          int symmetry, int* bad, hipStream_t s) {
  const long N = (long)n.n[0] * n.n[1] * n.n[2];
  if (N > 0 && faces != 0)
    sommer_interp_kernel<<<blocks_for(N, 256), 256, 0, s>>>(X, Y, Z, chi0, lap0, f0, f, n, faces, dT, S[0], S[1],
                                                            S[2], symmetry, bad);
}

void launch_sommer_bam(const double* f0, double* f_rhs, const Shape& n, const SommerGeom& sg, double vel,
                       const double S[3], hipStream_t s) {
  const long N = (long)n.n[0] * n.n[1] * n.n[2];
  if (N > 0 && sg.faces != 0)
    sommer_bam_kernel<<<blocks_for(N, 256), 256, 0, s>>>(f0, f_rhs, n, sg, vel, S[0], S[1], S[2]);
}

void launch_ghost_fill(const FieldsIn& u0, const FieldsOut& out, const Box& box, const Faces& w, int stage,
                       double tinny, int* flag, hipStream_t s) {
  // slabs are grid-strided; size the grid for the largest possible one (the
  // kernel's 32-bit indexing holds: a Block has fewer than 2^29 points)
  int wmax = 0;
  for (int a = 0; a < 3; ++a) wmax = std::max(wmax, std::max(w.lo[a], w.hi[a]));
  const long face = std::max({(long)box.n[0] * box.n[1], (long)box.n[0] * box.n[2], (long)box.n[1] * box.n[2]});
  const unsigned gx = (unsigned)std::min(std::max(blocks_for(face * wmax, 256), 1u), 2048u);
  if (wmax > 0) ghost_fill_kernel<<<dim3(gx, 6), 256, 0, s>>>(u0, out, box, w, stage, tinny, flag);
}

void launch_seg_copy(const FieldsIn& src, const FieldsOut& dst, const SegDesc& d, hipStream_t s) {
  const long N = (long)d.ext[0] * d.ext[1] * d.ext[2];
  if (N > 0) seg_copy_kernel<<<blocks_for(N, 256), 256, 0, s>>>(src, dst, d);
}

void launch_enforce(const FieldsOut& f, long n, hipStream_t s) {
  enforce_kernel<<<blocks_for(n, 256), 256, 0, s>>>(f, n);
}
void launch_lowerbound(double* chi, long n, double tinny, hipStream_t s) {
  lowerbound_kernel<<<blocks_for(n, 256), 256, 0, s>>>(chi, n, tinny);
}
void launch_rk4_update(const double* f0, double* f1, double* frhs, long n, double dT, int stage, hipStream_t s) {
  rk4_update_kernel<<<blocks_for(n, 256), 256, 0, s>>>(f0, f1, frhs, n, dT, stage);
}
void launch_copy_box(const double* f0, double* f, const Box& box, hipStream_t s) {
  const long N = (long)(box.hi[0] - box.lo[0] + 1) * (box.hi[1] - box.lo[1] + 1) * (box.hi[2] - box.lo[2] + 1);
  if (N > 0) copy_box_kernel<<<blocks_for(N, 256), 256, 0, s>>>(f0, f, box);
}
int launch_sumsq(const double* f, const Box& box, double* part, int maxblocks, hipStream_t s) {
  const long N = (long)(box.hi[0] - box.lo[0]) * (box.hi[1] - box.lo[1]) * (box.hi[2] - box.lo[2]);
  int nb = (int)blocks_for(N, 256);
  if (nb > maxblocks) nb = maxblocks;
  if (nb < 1) nb = 1;
  sumsq_kernel<<<nb, 256, 0, s>>>(f, box, part);
  return nb;
}

void launch_time_interp(const FieldsIn& f1, const FieldsIn& f2, const FieldsIn& f3, const FieldsOut& out, long n,
                        double c1, double c2, double c3, bool three, hipStream_t s) {
  if (n > 0) time_interp_kernel<<<dim3(blocks_for(n, 256), NFIELD), 256, 0, s>>>(f1, f2, f3, out, n, c1, c2, c3, three);
}

void launch_l2_jobs(const L2Job* jobs, int njobs, int maxparts, double* part, double* out, hipStream_t s) {
  if (njobs <= 0) return;
  l2_parts_kernel<<<dim3((unsigned)maxparts, (unsigned)njobs), 256, 0, s>>>(jobs, part);
  l2_final_kernel<<<(unsigned)njobs, 64, 0, s>>>(jobs, part, out);
}

}  // namespace amss
