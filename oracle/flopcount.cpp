// flopcount.cpp -- counts the floating-point operations of the reference
// algorithm for one RHS evaluation at one interior grid point by compiling
// the C restatement (bit-identical to the reference Fortran) with a counting
// scalar type.  add/sub/mul/div each count one operation; sqrt, pow and
// comparisons are counted separately and not included in the FLOP total.
// Test infrastructure: build and run with `make -C oracle flopcount`.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static long g_add, g_mul, g_div, g_other;
struct Flop {
  double v;
  constexpr Flop(double x = 0.0) : v(x) {}
};
#define OPS(OP, CNT)                                                                     \
  inline Flop operator OP(Flop a, Flop b) { ++CNT; return Flop(a.v OP b.v); }            \
  inline Flop operator OP(double a, Flop b) { ++CNT; return Flop(a OP b.v); }            \
  inline Flop operator OP(Flop a, double b) { ++CNT; return Flop(a.v OP b); }            \
  inline Flop operator OP(int a, Flop b) { ++CNT; return Flop(a OP b.v); }               \
  inline Flop operator OP(Flop a, int b) { ++CNT; return Flop(a.v OP b); }
OPS(+, g_add)
OPS(-, g_add)
OPS(*, g_mul)
OPS(/, g_div)
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline Flop& operator+=(Flop& a, Flop b) { ++g_add; a.v += b.v; return a; }
inline Flop& operator*=(Flop& a, Flop b) { ++g_mul; a.v *= b.v; return a; }
#define CMP(OP)                                                                      \
  inline bool operator OP(Flop a, Flop b) { return a.v OP b.v; }                    \
  inline bool operator OP(Flop a, double b) { return a.v OP b; }                    \
  inline bool operator OP(double a, Flop b) { return a OP b.v; }
CMP(<) CMP(>) CMP(<=) CMP(>=) CMP(==) CMP(!=)
inline Flop fabs(Flop a) { return Flop(std::fabs(a.v)); }
inline Flop sqrt(Flop a) { ++g_other; return Flop(std::sqrt(a.v)); }
inline Flop pow(Flop a, Flop b) { ++g_other; return Flop(std::pow(a.v, b.v)); }
inline Flop acos(Flop a) { return Flop(std::acos(a.v)); }

#define double Flop
#include "bssn_oracle.c"
#undef double

extern long orc_only_point;

static long count(int co, int gauge, long* adds, long* muls, long* divs, long* other) {
  const int n = 9;  // the centre point (4,4,4) has +-3 room on every axis
  const int ex[3] = {n, n, n};
  const size_t N = (size_t)n * n * n;
  static Flop buf[(24 + 24 + 18 + 6 + 7) * 729];
  static Flop X[9], Y[9], Z[9];
  for (int i = 0; i < n; ++i) X[i] = Y[i] = Z[i] = Flop(-1.0 + (i + 0.5) * 0.25);
  Flop *st[24], *rhs[24], *gam[18], *ric[6], *res[7];
  Flop* q = buf;
  srand(1);
  for (int v = 0; v < 24; ++v, q += N) {
    st[v] = q;
    for (size_t p = 0; p < N; ++p) q[p] = Flop(0.05 * (rand() / (double)RAND_MAX - 0.5) + (v == 0 ? -0.3 : 0.0));
  }
  for (int v = 0; v < 24; ++v, q += N) rhs[v] = q;
  for (int v = 0; v < 18; ++v, q += N) gam[v] = q;
  for (int v = 0; v < 6; ++v, q += N) ric[v] = q;
  for (int v = 0; v < 7; ++v, q += N) res[v] = q;
  orc_only_point = 4 + 9 * (4 + 9 * 4);
  // the NaN pre-sum is a whole-array reduction, not part of the per-point algebra
  long a0 = 0, m0 = 0, d0 = 0;
  {
    long a = g_add;
    Flop s = 0.0;
    for (int v = 0; v < 21; ++v)
      for (size_t p = 0; p < N; ++p) s = s + st[v][p];
    a0 = g_add - a;
  }
  g_add = g_mul = g_div = g_other = 0;
  orc_compute_rhs_bssn(ex, Flop(0.0), X, Y, Z, st, rhs, nullptr, gam, ric, res, 0, 0, Flop(0.02), co, gauge);
  *adds = g_add - a0;
  *muls = g_mul - m0;
  *divs = g_div - d0;
  *other = g_other;
  return *adds + *muls + *divs;
}

int main() {
  for (int co = 1; co >= 0; --co) {
    long a, m, d, o;
    const long t = count(co, 0, &a, &m, &d, &o);
    std::printf("{\"co\": %d, \"gauge\": 0, \"add_sub\": %ld, \"mul\": %ld, \"div\": %ld, \"flops\": %ld, \"sqrt_pow\": %ld}\n",
                co, a, m, d, t, o);
  }
  return 0;
}
