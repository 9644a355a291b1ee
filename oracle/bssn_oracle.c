/*
 * bssn_oracle.c -- CPU restatement of the AMSS-NCKU BSSN RHS + RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark load this library; the product path
 * never does.
 *
 * Every routine evaluates, point by point, the same arithmetic the reference
 * Fortran evaluates array statement by array statement, in the same
 * operation order (build with -ffp-contract=off), so the restatement can be
 * compared with the reference build to the last bit.  Citations are into
 * amss-ncku-python/AMSS_NCKU_source/.
 */
#include "bssn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NF 24
#define IX(ex, i, j, k) ((size_t)(i) + (size_t)(ex)[0] * ((size_t)(j) + (size_t)(ex)[1] * (size_t)(k)))

/* Parity of each evolved field under x, y, z reflection (bssn_class.C:183-361). */
static const double PAR[NF][3] = {
    {1, 1, 1},   {1, 1, 1},   {1, 1, 1},   {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1},
    {1, -1, -1}, {1, 1, 1},   {1, 1, 1},   {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1},
    {1, -1, -1}, {1, 1, 1},   {-1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {1, 1, 1},
    {-1, 1, 1},  {1, -1, 1},  {1, 1, -1},  {-1, 1, 1},  {1, -1, 1},  {1, 1, -1}};

/* ------------------------------------------------------------------------ */
/* Halo-padded read, as produced by symmetry_bd (fmisc.f90:877-901): an index
 * below 0 mirrors onto -i-1 and picks up the parity of that axis.  `off` is
 * added before mirroring (lopsided on gxx = dxx + 1, bssn_rhs.f90:148,948). */
static inline double fh(const double *f, const int *ex, int i, int j, int k,
                        const double *S, double off)
{
    double s = 1.0;
    if (i < 0) { i = -i - 1; s *= S[0]; }
    if (j < 0) { j = -j - 1; s *= S[1]; }
    if (k < 0) { k = -k - 1; s *= S[2]; }
    double v = f[IX(ex, i, j, k)] + off;
    return s == 1.0 ? v : v * s;
}

typedef struct {
    int ex[3];
    int lo2[3];   /* lowest readable 0-based index for fderivs/fdderivs */
    int lo3[3];   /* ... for lopsided / kodis */
    double dX, dY, dZ;
} grid_t;

static void grid_init(grid_t *g, const int *ex, const double *X, const double *Y,
                      const double *Z, int symmetry, int ko_octant_only)
{
    for (int a = 0; a < 3; a++) g->ex[a] = ex[a];
    g->dX = X[1] - X[0];
    g->dY = Y[1] - Y[0];
    g->dZ = Z[1] - Z[0];
    /* diff_new.f90:975-980, lopsidediff.f90:230-232, kodiss.f90:156-158 */
    int zs = symmetry > 0 && fabs(Z[0]) < g->dZ;
    int xs = symmetry > 1 && fabs(X[0]) < g->dX;
    int ys = symmetry > 1 && fabs(Y[0]) < g->dY;
    (void)ko_octant_only;
    g->lo2[0] = xs ? -2 : 0; g->lo2[1] = ys ? -2 : 0; g->lo2[2] = zs ? -2 : 0;
    g->lo3[0] = xs ? -3 : 0; g->lo3[1] = ys ? -3 : 0; g->lo3[2] = zs ? -3 : 0;
}

static inline int room(int i, int r, int lo, int n) { return i - r >= lo && i + r <= n - 1; }

/* fderivs at one point (diff_new.f90:996-1100, joint three-axis test). */
static void d1_pt(const grid_t *g, const double *f, const double *S, int i, int j, int k,
                  double *fx, double *fy, double *fz)
{
    const int *ex = g->ex;
    *fx = *fy = *fz = 0.0;
    if (i > ex[0] - 2 || j > ex[1] - 2 || k > ex[2] - 2) return;
    const double d12dx = 1.0 / 12.0 / g->dX, d12dy = 1.0 / 12.0 / g->dY, d12dz = 1.0 / 12.0 / g->dZ;
    const double d2dx = 1.0 / 2.0 / g->dX, d2dy = 1.0 / 2.0 / g->dY, d2dz = 1.0 / 2.0 / g->dZ;
#define F(a, b, c) fh(f, ex, a, b, c, S, 0.0)
    if (room(i, 2, g->lo2[0], ex[0]) && room(j, 2, g->lo2[1], ex[1]) && room(k, 2, g->lo2[2], ex[2])) {
        *fx = d12dx * (F(i - 2, j, k) - 8.0 * F(i - 1, j, k) + 8.0 * F(i + 1, j, k) - F(i + 2, j, k));
        *fy = d12dy * (F(i, j - 2, k) - 8.0 * F(i, j - 1, k) + 8.0 * F(i, j + 1, k) - F(i, j + 2, k));
        *fz = d12dz * (F(i, j, k - 2) - 8.0 * F(i, j, k - 1) + 8.0 * F(i, j, k + 1) - F(i, j, k + 2));
    } else if (room(i, 1, g->lo2[0], ex[0]) && room(j, 1, g->lo2[1], ex[1]) && room(k, 1, g->lo2[2], ex[2])) {
        *fx = d2dx * (-F(i - 1, j, k) + F(i + 1, j, k));
        *fy = d2dy * (-F(i, j - 1, k) + F(i, j + 1, k));
        *fz = d2dz * (-F(i, j, k - 1) + F(i, j, k + 1));
    }
#undef F
}

/* fdderivs at one point (diff_new.f90:1397-1524). o = xx xy xz yy yz zz. */
static void d2_pt(const grid_t *g, const double *f, const double *S, int i, int j, int k, double *o)
{
    const int *ex = g->ex;
    for (int a = 0; a < 6; a++) o[a] = 0.0;
    if (i > ex[0] - 2 || j > ex[1] - 2 || k > ex[2] - 2) return;
    const double dX = g->dX, dY = g->dY, dZ = g->dZ;
    const double F1o12 = 1.0 / 12.0, F1o144 = 1.0 / 144.0;
#define F(a, b, c) fh(f, ex, a, b, c, S, 0.0)
    if (room(i, 2, g->lo2[0], ex[0]) && room(j, 2, g->lo2[1], ex[1]) && room(k, 2, g->lo2[2], ex[2])) {
        const double Fdxdx = F1o12 / (dX * dX), Fdydy = F1o12 / (dY * dY), Fdzdz = F1o12 / (dZ * dZ);
        const double Fdxdy = F1o144 / (dX * dY), Fdxdz = F1o144 / (dX * dZ), Fdydz = F1o144 / (dY * dZ);
        o[0] = Fdxdx * (-F(i - 2, j, k) + 16.0 * F(i - 1, j, k) - 30.0 * F(i, j, k) - F(i + 2, j, k) + 16.0 * F(i + 1, j, k));
        o[3] = Fdydy * (-F(i, j - 2, k) + 16.0 * F(i, j - 1, k) - 30.0 * F(i, j, k) - F(i, j + 2, k) + 16.0 * F(i, j + 1, k));
        o[5] = Fdzdz * (-F(i, j, k - 2) + 16.0 * F(i, j, k - 1) - 30.0 * F(i, j, k) - F(i, j, k + 2) + 16.0 * F(i, j, k + 1));
#define ROWX(jj, kk) (F(i - 2, jj, kk) - 8.0 * F(i - 1, jj, kk) + 8.0 * F(i + 1, jj, kk) - F(i + 2, jj, kk))
#define ROWY(ii, kk) (F(ii, j - 2, kk) - 8.0 * F(ii, j - 1, kk) + 8.0 * F(ii, j + 1, kk) - F(ii, j + 2, kk))
        o[1] = Fdxdy * ((ROWX(j - 2, k)) - 8.0 * (ROWX(j - 1, k)) + 8.0 * (ROWX(j + 1, k)) - (ROWX(j + 2, k)));
        o[2] = Fdxdz * ((ROWX(j, k - 2)) - 8.0 * (ROWX(j, k - 1)) + 8.0 * (ROWX(j, k + 1)) - (ROWX(j, k + 2)));
        o[4] = Fdydz * ((ROWY(i, k - 2)) - 8.0 * (ROWY(i, k - 1)) + 8.0 * (ROWY(i, k + 1)) - (ROWY(i, k + 2)));
#undef ROWX
#undef ROWY
    } else if (room(i, 1, g->lo2[0], ex[0]) && room(j, 1, g->lo2[1], ex[1]) && room(k, 1, g->lo2[2], ex[2])) {
        const double Sdxdx = 1.0 / (dX * dX), Sdydy = 1.0 / (dY * dY), Sdzdz = 1.0 / (dZ * dZ);
        const double Sdxdy = 0.25 / (dX * dY), Sdxdz = 0.25 / (dX * dZ), Sdydz = 0.25 / (dY * dZ);
        o[0] = Sdxdx * (F(i - 1, j, k) - 2.0 * F(i, j, k) + F(i + 1, j, k));
        o[3] = Sdydy * (F(i, j - 1, k) - 2.0 * F(i, j, k) + F(i, j + 1, k));
        o[5] = Sdzdz * (F(i, j, k - 1) - 2.0 * F(i, j, k) + F(i, j, k + 1));
        o[1] = Sdxdy * (F(i - 1, j - 1, k) - F(i + 1, j - 1, k) - F(i - 1, j + 1, k) + F(i + 1, j + 1, k));
        o[2] = Sdxdz * (F(i - 1, j, k - 1) - F(i + 1, j, k - 1) - F(i - 1, j, k + 1) + F(i + 1, j, k + 1));
        o[4] = Sdydz * (F(i, j - 1, k - 1) - F(i, j + 1, k - 1) - F(i, j - 1, k + 1) + F(i, j + 1, k + 1));
    }
#undef F
}

/* lopsided at one point, one direction at a time (lopsidediff.f90:321-481).
 * p(m) reads the field m points along the axis; returns the updated rhs. */
typedef struct { const double *f; const int *ex; const double *S; double off; int i, j, k, ax; } line_t;
static inline double lp(const line_t *L, int m)
{
    int i = L->i, j = L->j, k = L->k;
    if (L->ax == 0) i += m; else if (L->ax == 1) j += m; else k += m;
    return fh(L->f, L->ex, i, j, k, L->S, L->off);
}

static double lop_dir(double r, const line_t *L, double beta, double d12, int pos, int lo, int n)
{
    if (beta > 0.0) {
        if (pos + 3 <= n - 1)
            r = r + beta * d12 * (-3.0 * lp(L, -1) - 10.0 * lp(L, 0) + 18.0 * lp(L, 1) - 6.0 * lp(L, 2) + lp(L, 3));
        else if (pos + 2 <= n - 1)
            r = r + beta * d12 * (lp(L, -2) - 8.0 * lp(L, -1) + 8.0 * lp(L, 1) - lp(L, 2));
        else if (pos + 1 <= n - 1)
            r = r - beta * d12 * (-3.0 * lp(L, 1) - 10.0 * lp(L, 0) + 18.0 * lp(L, -1) - 6.0 * lp(L, -2) + lp(L, -3));
    } else if (beta < 0.0) {
        if (pos - 3 >= lo)
            r = r - beta * d12 * (-3.0 * lp(L, 1) - 10.0 * lp(L, 0) + 18.0 * lp(L, -1) - 6.0 * lp(L, -2) + lp(L, -3));
        else if (pos - 2 >= lo)
            r = r + beta * d12 * (lp(L, -2) - 8.0 * lp(L, -1) + 8.0 * lp(L, 1) - lp(L, 2));
        else if (pos - 1 >= lo)
            r = r + beta * d12 * (-3.0 * lp(L, -1) - 10.0 * lp(L, 0) + 18.0 * lp(L, 1) - 6.0 * lp(L, 2) + lp(L, 3));
    }
    return r;
}

static double lop_pt(const grid_t *g, const double *f, double off, const double *S, int i, int j, int k,
                     double bx, double by, double bz, double r)
{
    const int *ex = g->ex;
    if (i > ex[0] - 2 || j > ex[1] - 2 || k > ex[2] - 2) return r;
    line_t L = {f, ex, S, off, i, j, k, 0};
    r = lop_dir(r, &L, bx, 1.0 / 12.0 / g->dX, i, g->lo3[0], ex[0]);
    L.ax = 1;
    r = lop_dir(r, &L, by, 1.0 / 12.0 / g->dY, j, g->lo3[1], ex[1]);
    L.ax = 2;
    r = lop_dir(r, &L, bz, 1.0 / 12.0 / g->dZ, k, g->lo3[2], ex[2]);
    return r;
}

/* kodis at one point (kodiss.f90:162-212). */
static double ko_pt(const grid_t *g, const double *f, const double *S, int i, int j, int k, double eps,
                    double r)
{
    const int *ex = g->ex;
    if (!(room(i, 3, g->lo3[0], ex[0]) && room(j, 3, g->lo3[1], ex[1]) && room(k, 3, g->lo3[2], ex[2])))
        return r;
#define F(a, b, c) fh(f, ex, a, b, c, S, 0.0)
    const double c0 = F(i, j, k);
    const double ax = (F(i - 3, j, k) + F(i + 3, j, k)) - 6.0 * (F(i - 2, j, k) + F(i + 2, j, k)) +
                      15.0 * (F(i - 1, j, k) + F(i + 1, j, k)) - 20.0 * c0;
    const double ay = (F(i, j - 3, k) + F(i, j + 3, k)) - 6.0 * (F(i, j - 2, k) + F(i, j + 2, k)) +
                      15.0 * (F(i, j - 1, k) + F(i, j + 1, k)) - 20.0 * c0;
    const double az = (F(i, j, k - 3) + F(i, j, k + 3)) - 6.0 * (F(i, j, k - 2) + F(i, j, k + 2)) +
                      15.0 * (F(i, j, k - 1) + F(i, j, k + 1)) - 20.0 * c0;
#undef F
    return r + eps / 64.0 * (ax / g->dX + ay / g->dY + az / g->dZ);
}

/* ------------------------------------------------------------------------ */
/* Array forms of the stencil routines, for direct comparison with the
 * reference symbols fderivs_, fdderivs_, lopsided_, kodis_, symmetry_bd_. */
void orc_fderivs(const int *ex, const double *f, double *fx, double *fy, double *fz, const double *X,
                 const double *Y, const double *Z, double S1, double S2, double S3, int symmetry)
{
    grid_t g;
    grid_init(&g, ex, X, Y, Z, symmetry, 0);
    const double S[3] = {S1, S2, S3};
    for (int k = 0; k < ex[2]; k++)
        for (int j = 0; j < ex[1]; j++)
            for (int i = 0; i < ex[0]; i++) {
                size_t p = IX(ex, i, j, k);
                d1_pt(&g, f, S, i, j, k, &fx[p], &fy[p], &fz[p]);
            }
}

void orc_fdderivs(const int *ex, const double *f, double *fxx, double *fxy, double *fxz, double *fyy,
                  double *fyz, double *fzz, const double *X, const double *Y, const double *Z, double S1,
                  double S2, double S3, int symmetry)
{
    grid_t g;
    grid_init(&g, ex, X, Y, Z, symmetry, 0);
    const double S[3] = {S1, S2, S3};
    double o[6];
    for (int k = 0; k < ex[2]; k++)
        for (int j = 0; j < ex[1]; j++)
            for (int i = 0; i < ex[0]; i++) {
                size_t p = IX(ex, i, j, k);
                d2_pt(&g, f, S, i, j, k, o);
                fxx[p] = o[0]; fxy[p] = o[1]; fxz[p] = o[2];
                fyy[p] = o[3]; fyz[p] = o[4]; fzz[p] = o[5];
            }
}

void orc_lopsided(const int *ex, const double *X, const double *Y, const double *Z, const double *f,
                  double *f_rhs, const double *Sfx, const double *Sfy, const double *Sfz, int symmetry,
                  const double *SoA)
{
    grid_t g;
    grid_init(&g, ex, X, Y, Z, symmetry, 0);
    for (int k = 0; k < ex[2]; k++)
        for (int j = 0; j < ex[1]; j++)
            for (int i = 0; i < ex[0]; i++) {
                size_t p = IX(ex, i, j, k);
                f_rhs[p] = lop_pt(&g, f, 0.0, SoA, i, j, k, Sfx[p], Sfy[p], Sfz[p], f_rhs[p]);
            }
}

void orc_kodis(const int *ex, const double *X, const double *Y, const double *Z, const double *f,
               double *f_rhs, const double *SoA, int symmetry, double eps)
{
    grid_t g;
    grid_init(&g, ex, X, Y, Z, symmetry, 1);
    for (int k = 0; k < ex[2]; k++)
        for (int j = 0; j < ex[1]; j++)
            for (int i = 0; i < ex[0]; i++) {
                size_t p = IX(ex, i, j, k);
                f_rhs[p] = ko_pt(&g, f, SoA, i, j, k, eps, f_rhs[p]);
            }
}

void orc_symmetry_bd(int ord, const int *ex, const double *f, double *fhout, const double *SoA)
{
    const int n0 = ex[0] + ord, n1 = ex[1] + ord, n2 = ex[2] + ord;
    for (int k = 0; k < n2; k++)
        for (int j = 0; j < n1; j++)
            for (int i = 0; i < n0; i++)
                fhout[(size_t)i + (size_t)n0 * ((size_t)j + (size_t)n1 * k)] =
                    fh(f, ex, i - ord, j - ord, k - ord, SoA, 0.0);
}

/* ------------------------------------------------------------------------ */
/* The RHS (bssn_rhs.f90:5-1186), GAUGE 0 (:792-799) or GAUGE 2 (:809-820).  */
enum { CHI, TRK, DXX, GXY, GXZ, DYY, GYZ, DZZ, AXX, AXY, AXZ, AYY, AYZ, AZZ,
       GAMX, GAMY, GAMZ, LAP, BETAX, BETAY, BETAZ, DTSFX, DTSFY, DTSFZ };

/* Restrict the point loop to one linear index (operation counting only;
 * -1 = every point). */
long orc_only_point = -1;

int orc_compute_rhs_bssn(const int *ex, double T, const double *X, const double *Y, const double *Z,
                         const double *const *st, double *const *rhs, const double *const *mat,
                         double *const *gamo, double *const *rico, double *const *res, int symmetry,
                         int lev, double eps, int co, int gauge)
{
    (void)T; (void)lev;
    const size_t N = (size_t)ex[0] * ex[1] * ex[2];
    /* NaN sanity check over 21 fields (bssn_rhs.f90:109-138). */
    {
        double s = 0.0;
        for (int v = 0; v < 21; v++)
            for (size_t p = 0; p < N; p++) s += st[v][p];
        if (s != s) return 1;
    }
    grid_t g;
    grid_init(&g, ex, X, Y, Z, symmetry, 0);
    const double ZEO = 0.0, ONE = 1.0, TWO = 2.0, FOUR = 4.0, EIGHT = 8.0, HALF = 0.5;
    const double FF = 0.75, eta = 2.0;
    const double F1o3 = 1.0 / 3.0, F2o3 = 2.0 / 3.0, F3o2 = 1.5, F16 = 16.0, F8 = 8.0;
    const double PI = acos(-ONE);
    (void)ZEO;

    for (int k = 0; k < ex[2]; k++)
    for (int j = 0; j < ex[1]; j++)
    for (int i = 0; i < ex[0]; i++) {
        const size_t p = IX(ex, i, j, k);
        if (orc_only_point >= 0 && (long)p != orc_only_point) continue;
        const double chi = st[CHI][p], trK = st[TRK][p];
        const double dxx = st[DXX][p], gxy = st[GXY][p], gxz = st[GXZ][p];
        const double dyy = st[DYY][p], gyz = st[GYZ][p], dzz = st[DZZ][p];
        const double Axx = st[AXX][p], Axy = st[AXY][p], Axz = st[AXZ][p];
        const double Ayy = st[AYY][p], Ayz = st[AYZ][p], Azz = st[AZZ][p];
        const double Gamx = st[GAMX][p], Gamy = st[GAMY][p], Gamz = st[GAMZ][p];
        const double Lap = st[LAP][p], betax = st[BETAX][p], betay = st[BETAY][p], betaz = st[BETAZ][p];
        const double dtSfx = st[DTSFX][p], dtSfy = st[DTSFY][p], dtSfz = st[DTSFZ][p];
        double rho = 0, Sx = 0, Sy = 0, Sz = 0, Sxx = 0, Sxy = 0, Sxz = 0, Syy = 0, Syz = 0, Szz = 0;
        if (mat) {
            rho = mat[0][p]; Sx = mat[1][p]; Sy = mat[2][p]; Sz = mat[3][p];
            Sxx = mat[4][p]; Sxy = mat[5][p]; Sxz = mat[6][p]; Syy = mat[7][p]; Syz = mat[8][p]; Szz = mat[9][p];
        }
        double dd[6];
#define D1(v, a, b, c) d1_pt(&g, st[v], PAR[v], i, j, k, &(a), &(b), &(c))
#define D2(v, a, b, c, d, e, f_) do { d2_pt(&g, st[v], PAR[v], i, j, k, dd); \
        a = dd[0]; b = dd[1]; c = dd[2]; d = dd[3]; e = dd[4]; f_ = dd[5]; } while (0)

        /* :146-150 */
        const double alpn1 = Lap + ONE, chin1 = chi + ONE;
        const double gxx = dxx + ONE, gyy = dyy + ONE, gzz = dzz + ONE;
        /* :152-160 */
        double betaxx, betaxy, betaxz, betayx, betayy, betayz, betazx, betazy, betazz;
        D1(BETAX, betaxx, betaxy, betaxz);
        D1(BETAY, betayx, betayy, betayz);
        D1(BETAZ, betazx, betazy, betazz);
        const double div_beta = betaxx + betayy + betazz;
        double chix, chiy, chiz;
        D1(CHI, chix, chiy, chiz);
        double chi_rhs = F2o3 * chin1 * (alpn1 * trK - div_beta);
        /* :162-191 */
        double gxxx, gxxy, gxxz, gxyx, gxyy, gxyz, gxzx, gxzy, gxzz;
        double gyyx, gyyy, gyyz, gyzx, gyzy, gyzz, gzzx, gzzy, gzzz;
        D1(DXX, gxxx, gxxy, gxxz);
        D1(GXY, gxyx, gxyy, gxyz);
        D1(GXZ, gxzx, gxzy, gxzz);
        D1(DYY, gyyx, gyyy, gyyz);
        D1(GYZ, gyzx, gyzy, gyzz);
        D1(DZZ, gzzx, gzzy, gzzz);
        double gxx_rhs = -TWO * alpn1 * Axx - F2o3 * gxx * div_beta +
                         TWO * (gxx * betaxx + gxy * betayx + gxz * betazx);
        double gyy_rhs = -TWO * alpn1 * Ayy - F2o3 * gyy * div_beta +
                         TWO * (gxy * betaxy + gyy * betayy + gyz * betazy);
        double gzz_rhs = -TWO * alpn1 * Azz - F2o3 * gzz * div_beta +
                         TWO * (gxz * betaxz + gyz * betayz + gzz * betazz);
        double gxy_rhs = -TWO * alpn1 * Axy + F1o3 * gxy * div_beta + gxx * betaxy + gxz * betazy +
                         gyy * betayx + gyz * betazx - gxy * betazz;
        double gyz_rhs = -TWO * alpn1 * Ayz + F1o3 * gyz * div_beta + gxy * betaxz + gyy * betayz +
                         gxz * betaxy + gzz * betazy - gyz * betaxx;
        double gxz_rhs = -TWO * alpn1 * Axz + F1o3 * gxz * div_beta + gxx * betaxz + gxy * betayz +
                         gyz * betayx + gzz * betazx - gxz * betayy;
        /* :194-201 */
        double gupzz = gxx * gyy * gzz + gxy * gyz * gxz + gxz * gxy * gyz - gxz * gyy * gxz -
                       gxy * gxy * gzz - gxx * gyz * gyz;
        const double gupxx = (gyy * gzz - gyz * gyz) / gupzz;
        const double gupxy = -(gxy * gzz - gyz * gxz) / gupzz;
        const double gupxz = (gxy * gyz - gyy * gxz) / gupzz;
        const double gupyy = (gxx * gzz - gxz * gxz) / gupzz;
        const double gupyz = -(gxx * gyz - gxy * gxz) / gupzz;
        gupzz = (gxx * gyy - gxy * gxy) / gupzz;
        /* :203-232 */
        if (co == 0) {
            res[4][p] = Gamx - (gupxx * (gupxx * gxxx + gupxy * gxyx + gupxz * gxzx) +
                                gupxy * (gupxx * gxyx + gupxy * gyyx + gupxz * gyzx) +
                                gupxz * (gupxx * gxzx + gupxy * gyzx + gupxz * gzzx) +
                                gupxx * (gupxy * gxxy + gupyy * gxyy + gupyz * gxzy) +
                                gupxy * (gupxy * gxyy + gupyy * gyyy + gupyz * gyzy) +
                                gupxz * (gupxy * gxzy + gupyy * gyzy + gupyz * gzzy) +
                                gupxx * (gupxz * gxxz + gupyz * gxyz + gupzz * gxzz) +
                                gupxy * (gupxz * gxyz + gupyz * gyyz + gupzz * gyzz) +
                                gupxz * (gupxz * gxzz + gupyz * gyzz + gupzz * gzzz));
            res[5][p] = Gamy - (gupxx * (gupxy * gxxx + gupyy * gxyx + gupyz * gxzx) +
                                gupxy * (gupxy * gxyx + gupyy * gyyx + gupyz * gyzx) +
                                gupxz * (gupxy * gxzx + gupyy * gyzx + gupyz * gzzx) +
                                gupxy * (gupxy * gxxy + gupyy * gxyy + gupyz * gxzy) +
                                gupyy * (gupxy * gxyy + gupyy * gyyy + gupyz * gyzy) +
                                gupyz * (gupxy * gxzy + gupyy * gyzy + gupyz * gzzy) +
                                gupxy * (gupxz * gxxz + gupyz * gxyz + gupzz * gxzz) +
                                gupyy * (gupxz * gxyz + gupyz * gyyz + gupzz * gyzz) +
                                gupyz * (gupxz * gxzz + gupyz * gyzz + gupzz * gzzz));
            res[6][p] = Gamz - (gupxx * (gupxz * gxxx + gupyz * gxyx + gupzz * gxzx) +
                                gupxy * (gupxz * gxyx + gupyz * gyyx + gupzz * gyzx) +
                                gupxz * (gupxz * gxzx + gupyz * gyzx + gupzz * gzzx) +
                                gupxy * (gupxz * gxxy + gupyz * gxyy + gupzz * gxzy) +
                                gupyy * (gupxz * gxyy + gupyz * gyyy + gupzz * gyzy) +
                                gupyz * (gupxz * gxzy + gupyz * gyzy + gupzz * gzzy) +
                                gupxz * (gupxz * gxxz + gupyz * gxyz + gupzz * gxzz) +
                                gupyz * (gupxz * gxyz + gupyz * gyyz + gupzz * gyzz) +
                                gupzz * (gupxz * gxzz + gupyz * gyzz + gupzz * gzzz));
        }
        /* :235-257 conformal Christoffels of the second kind */
        double Gamxxx = HALF * (gupxx * gxxx + gupxy * (TWO * gxyx - gxxy) + gupxz * (TWO * gxzx - gxxz));
        double Gamyxx = HALF * (gupxy * gxxx + gupyy * (TWO * gxyx - gxxy) + gupyz * (TWO * gxzx - gxxz));
        double Gamzxx = HALF * (gupxz * gxxx + gupyz * (TWO * gxyx - gxxy) + gupzz * (TWO * gxzx - gxxz));
        double Gamxyy = HALF * (gupxx * (TWO * gxyy - gyyx) + gupxy * gyyy + gupxz * (TWO * gyzy - gyyz));
        double Gamyyy = HALF * (gupxy * (TWO * gxyy - gyyx) + gupyy * gyyy + gupyz * (TWO * gyzy - gyyz));
        double Gamzyy = HALF * (gupxz * (TWO * gxyy - gyyx) + gupyz * gyyy + gupzz * (TWO * gyzy - gyyz));
        double Gamxzz = HALF * (gupxx * (TWO * gxzz - gzzx) + gupxy * (TWO * gyzz - gzzy) + gupxz * gzzz);
        double Gamyzz = HALF * (gupxy * (TWO * gxzz - gzzx) + gupyy * (TWO * gyzz - gzzy) + gupyz * gzzz);
        double Gamzzz = HALF * (gupxz * (TWO * gxzz - gzzx) + gupyz * (TWO * gyzz - gzzy) + gupzz * gzzz);
        double Gamxxy = HALF * (gupxx * gxxy + gupxy * gyyx + gupxz * (gxzy + gyzx - gxyz));
        double Gamyxy = HALF * (gupxy * gxxy + gupyy * gyyx + gupyz * (gxzy + gyzx - gxyz));
        double Gamzxy = HALF * (gupxz * gxxy + gupyz * gyyx + gupzz * (gxzy + gyzx - gxyz));
        double Gamxxz = HALF * (gupxx * gxxz + gupxy * (gxyz + gyzx - gxzy) + gupxz * gzzx);
        double Gamyxz = HALF * (gupxy * gxxz + gupyy * (gxyz + gyzx - gxzy) + gupyz * gzzx);
        double Gamzxz = HALF * (gupxz * gxxz + gupyz * (gxyz + gyzx - gxzy) + gupzz * gzzx);
        double Gamxyz = HALF * (gupxx * (gxyz + gxzy - gyzx) + gupxy * gyyz + gupxz * gzzy);
        double Gamyyz = HALF * (gupxy * (gxyz + gxzy - gyzx) + gupyy * gyyz + gupyz * gzzy);
        double Gamzyz = HALF * (gupxz * (gxyz + gxzy - gyzx) + gupyz * gyyz + gupzz * gzzy);
        /* :260-282 A^ij kept in the R slots */
        double Rxx = gupxx * gupxx * Axx + gupxy * gupxy * Ayy + gupxz * gupxz * Azz +
                     TWO * (gupxx * gupxy * Axy + gupxx * gupxz * Axz + gupxy * gupxz * Ayz);
        double Ryy = gupxy * gupxy * Axx + gupyy * gupyy * Ayy + gupyz * gupyz * Azz +
                     TWO * (gupxy * gupyy * Axy + gupxy * gupyz * Axz + gupyy * gupyz * Ayz);
        double Rzz = gupxz * gupxz * Axx + gupyz * gupyz * Ayy + gupzz * gupzz * Azz +
                     TWO * (gupxz * gupyz * Axy + gupxz * gupzz * Axz + gupyz * gupzz * Ayz);
        double Rxy = gupxx * gupxy * Axx + gupxy * gupyy * Ayy + gupxz * gupyz * Azz +
                     (gupxx * gupyy + gupxy * gupxy) * Axy + (gupxx * gupyz + gupxz * gupxy) * Axz +
                     (gupxy * gupyz + gupxz * gupyy) * Ayz;
        double Rxz = gupxx * gupxz * Axx + gupxy * gupyz * Ayy + gupxz * gupzz * Azz +
                     (gupxx * gupyz + gupxy * gupxz) * Axy + (gupxx * gupzz + gupxz * gupxz) * Axz +
                     (gupxy * gupzz + gupxz * gupyz) * Ayz;
        double Ryz = gupxy * gupxz * Axx + gupyy * gupyz * Ayy + gupyz * gupzz * Azz +
                     (gupxy * gupyz + gupyy * gupxz) * Axy + (gupxy * gupzz + gupyz * gupxz) * Axz +
                     (gupyy * gupzz + gupyz * gupyz) * Ayz;
        /* :285-313 */
        double Lapx, Lapy, Lapz, Kx, Ky, Kz;
        D1(LAP, Lapx, Lapy, Lapz);
        D1(TRK, Kx, Ky, Kz);
        double Gamx_rhs = -TWO * (Lapx * Rxx + Lapy * Rxy + Lapz * Rxz) +
            TWO * alpn1 * (-F3o2 / chin1 * (chix * Rxx + chiy * Rxy + chiz * Rxz) -
                           gupxx * (F2o3 * Kx + EIGHT * PI * Sx) - gupxy * (F2o3 * Ky + EIGHT * PI * Sy) -
                           gupxz * (F2o3 * Kz + EIGHT * PI * Sz) + Gamxxx * Rxx + Gamxyy * Ryy +
                           Gamxzz * Rzz + TWO * (Gamxxy * Rxy + Gamxxz * Rxz + Gamxyz * Ryz));
        double Gamy_rhs = -TWO * (Lapx * Rxy + Lapy * Ryy + Lapz * Ryz) +
            TWO * alpn1 * (-F3o2 / chin1 * (chix * Rxy + chiy * Ryy + chiz * Ryz) -
                           gupxy * (F2o3 * Kx + EIGHT * PI * Sx) - gupyy * (F2o3 * Ky + EIGHT * PI * Sy) -
                           gupyz * (F2o3 * Kz + EIGHT * PI * Sz) + Gamyxx * Rxx + Gamyyy * Ryy +
                           Gamyzz * Rzz + TWO * (Gamyxy * Rxy + Gamyxz * Rxz + Gamyyz * Ryz));
        double Gamz_rhs = -TWO * (Lapx * Rxz + Lapy * Ryz + Lapz * Rzz) +
            TWO * alpn1 * (-F3o2 / chin1 * (chix * Rxz + chiy * Ryz + chiz * Rzz) -
                           gupxz * (F2o3 * Kx + EIGHT * PI * Sx) - gupyz * (F2o3 * Ky + EIGHT * PI * Sy) -
                           gupzz * (F2o3 * Kz + EIGHT * PI * Sz) + Gamzxx * Rxx + Gamzyy * Ryy +
                           Gamzzz * Rzz + TWO * (Gamzxy * Rxy + Gamzxz * Rxz + Gamzyz * Ryz));
        /* :315-353 second derivatives of the shift land in the g_ij,k slots */
        D2(BETAX, gxxx, gxyx, gxzx, gyyx, gyzx, gzzx);
        D2(BETAY, gxxy, gxyy, gxzy, gyyy, gyzy, gzzy);
        D2(BETAZ, gxxz, gxyz, gxzz, gyyz, gyzz, gzzz);
        double fxx = gxxx + gxyy + gxzz;
        double fxy = gxyx + gyyy + gyzz;
        double fxz = gxzx + gyzy + gzzz;
        const double Gamxa = gupxx * Gamxxx + gupyy * Gamxyy + gupzz * Gamxzz +
                             TWO * (gupxy * Gamxxy + gupxz * Gamxxz + gupyz * Gamxyz);
        const double Gamya = gupxx * Gamyxx + gupyy * Gamyyy + gupzz * Gamyzz +
                             TWO * (gupxy * Gamyxy + gupxz * Gamyxz + gupyz * Gamyyz);
        const double Gamza = gupxx * Gamzxx + gupyy * Gamzyy + gupzz * Gamzzz +
                             TWO * (gupxy * Gamzxy + gupxz * Gamzxz + gupyz * Gamzyz);
        double Gamxx, Gamxy, Gamxz, Gamyx, Gamyy, Gamyz, Gamzx, Gamzy, Gamzz;
        D1(GAMX, Gamxx, Gamxy, Gamxz);
        D1(GAMY, Gamyx, Gamyy, Gamyz);
        D1(GAMZ, Gamzx, Gamzy, Gamzz);
        Gamx_rhs = Gamx_rhs + F2o3 * Gamxa * div_beta - Gamxa * betaxx - Gamya * betaxy - Gamza * betaxz +
                   F1o3 * (gupxx * fxx + gupxy * fxy + gupxz * fxz) + gupxx * gxxx + gupyy * gyyx +
                   gupzz * gzzx + TWO * (gupxy * gxyx + gupxz * gxzx + gupyz * gyzx);
        Gamy_rhs = Gamy_rhs + F2o3 * Gamya * div_beta - Gamxa * betayx - Gamya * betayy - Gamza * betayz +
                   F1o3 * (gupxy * fxx + gupyy * fxy + gupyz * fxz) + gupxx * gxxy + gupyy * gyyy +
                   gupzz * gzzy + TWO * (gupxy * gxyy + gupxz * gxzy + gupyz * gyzy);
        Gamz_rhs = Gamz_rhs + F2o3 * Gamza * div_beta - Gamxa * betazx - Gamya * betazy - Gamza * betazz +
                   F1o3 * (gupxz * fxx + gupyz * fxy + gupzz * fxz) + gupxx * gxxz + gupyy * gyyz +
                   gupzz * gzzz + TWO * (gupxy * gxyz + gupxz * gxzz + gupyz * gyzz);
        /* :356-375 first-kind Christoffels stored in the g_ij,k slots */
        gxxx = gxx * Gamxxx + gxy * Gamyxx + gxz * Gamzxx;
        gxyx = gxx * Gamxxy + gxy * Gamyxy + gxz * Gamzxy;
        gxzx = gxx * Gamxxz + gxy * Gamyxz + gxz * Gamzxz;
        gyyx = gxx * Gamxyy + gxy * Gamyyy + gxz * Gamzyy;
        gyzx = gxx * Gamxyz + gxy * Gamyyz + gxz * Gamzyz;
        gzzx = gxx * Gamxzz + gxy * Gamyzz + gxz * Gamzzz;
        gxxy = gxy * Gamxxx + gyy * Gamyxx + gyz * Gamzxx;
        gxyy = gxy * Gamxxy + gyy * Gamyxy + gyz * Gamzxy;
        gxzy = gxy * Gamxxz + gyy * Gamyxz + gyz * Gamzxz;
        gyyy = gxy * Gamxyy + gyy * Gamyyy + gyz * Gamzyy;
        gyzy = gxy * Gamxyz + gyy * Gamyyz + gyz * Gamzyz;
        gzzy = gxy * Gamxzz + gyy * Gamyzz + gyz * Gamzzz;
        gxxz = gxz * Gamxxx + gyz * Gamyxx + gzz * Gamzxx;
        gxyz = gxz * Gamxxy + gyz * Gamyxy + gzz * Gamzxy;
        gxzz = gxz * Gamxxz + gyz * Gamyxz + gzz * Gamzxz;
        gyyz = gxz * Gamxyy + gyz * Gamyyy + gzz * Gamzyy;
        gyzz = gxz * Gamxyz + gyz * Gamyyz + gzz * Gamzyz;
        gzzz = gxz * Gamxzz + gyz * Gamyzz + gzz * Gamzzz;
        /* :378-400 */
        double fyy, fyz, fzz;
#define LAPG(v, R) do { D2(v, fxx, fxy, fxz, fyy, fyz, fzz); \
        R = gupxx * fxx + gupyy * fyy + gupzz * fzz + (gupxy * fxy + gupxz * fxz + gupyz * fyz) * TWO; } while (0)
        LAPG(DXX, Rxx);
        LAPG(DYY, Ryy);
        LAPG(DZZ, Rzz);
        LAPG(GXY, Rxy);
        LAPG(GXZ, Rxz);
        LAPG(GYZ, Ryz);
#undef LAPG
        /* :402-601 conformal Ricci */
        Rxx = -HALF * Rxx + gxx * Gamxx + gxy * Gamyx + gxz * Gamzx + Gamxa * gxxx + Gamya * gxyx +
              Gamza * gxzx +
              gupxx * (TWO * (Gamxxx * gxxx + Gamyxx * gxyx + Gamzxx * gxzx) + Gamxxx * gxxx +
                       Gamyxx * gxxy + Gamzxx * gxxz) +
              gupxy * (TWO * (Gamxxx * gxyx + Gamyxx * gyyx + Gamzxx * gyzx + Gamxxy * gxxx +
                              Gamyxy * gxyx + Gamzxy * gxzx) +
                       Gamxxy * gxxx + Gamyxy * gxxy + Gamzxy * gxxz + Gamxxx * gxyx + Gamyxx * gxyy +
                       Gamzxx * gxyz) +
              gupxz * (TWO * (Gamxxx * gxzx + Gamyxx * gyzx + Gamzxx * gzzx + Gamxxz * gxxx +
                              Gamyxz * gxyx + Gamzxz * gxzx) +
                       Gamxxz * gxxx + Gamyxz * gxxy + Gamzxz * gxxz + Gamxxx * gxzx + Gamyxx * gxzy +
                       Gamzxx * gxzz) +
              gupyy * (TWO * (Gamxxy * gxyx + Gamyxy * gyyx + Gamzxy * gyzx) + Gamxxy * gxyx +
                       Gamyxy * gxyy + Gamzxy * gxyz) +
              gupyz * (TWO * (Gamxxy * gxzx + Gamyxy * gyzx + Gamzxy * gzzx + Gamxxz * gxyx +
                              Gamyxz * gyyx + Gamzxz * gyzx) +
                       Gamxxz * gxyx + Gamyxz * gxyy + Gamzxz * gxyz + Gamxxy * gxzx + Gamyxy * gxzy +
                       Gamzxy * gxzz) +
              gupzz * (TWO * (Gamxxz * gxzx + Gamyxz * gyzx + Gamzxz * gzzx) + Gamxxz * gxzx +
                       Gamyxz * gxzy + Gamzxz * gxzz);
        Ryy = -HALF * Ryy + gxy * Gamxy + gyy * Gamyy + gyz * Gamzy + Gamxa * gxyy + Gamya * gyyy +
              Gamza * gyzy +
              gupxx * (TWO * (Gamxxy * gxxy + Gamyxy * gxyy + Gamzxy * gxzy) + Gamxxy * gxyx +
                       Gamyxy * gxyy + Gamzxy * gxyz) +
              gupxy * (TWO * (Gamxxy * gxyy + Gamyxy * gyyy + Gamzxy * gyzy + Gamxyy * gxxy +
                              Gamyyy * gxyy + Gamzyy * gxzy) +
                       Gamxyy * gxyx + Gamyyy * gxyy + Gamzyy * gxyz + Gamxxy * gyyx + Gamyxy * gyyy +
                       Gamzxy * gyyz) +
              gupxz * (TWO * (Gamxxy * gxzy + Gamyxy * gyzy + Gamzxy * gzzy + Gamxyz * gxxy +
                              Gamyyz * gxyy + Gamzyz * gxzy) +
                       Gamxyz * gxyx + Gamyyz * gxyy + Gamzyz * gxyz + Gamxxy * gyzx + Gamyxy * gyzy +
                       Gamzxy * gyzz) +
              gupyy * (TWO * (Gamxyy * gxyy + Gamyyy * gyyy + Gamzyy * gyzy) + Gamxyy * gyyx +
                       Gamyyy * gyyy + Gamzyy * gyyz) +
              gupyz * (TWO * (Gamxyy * gxzy + Gamyyy * gyzy + Gamzyy * gzzy + Gamxyz * gxyy +
                              Gamyyz * gyyy + Gamzyz * gyzy) +
                       Gamxyz * gyyx + Gamyyz * gyyy + Gamzyz * gyyz + Gamxyy * gyzx + Gamyyy * gyzy +
                       Gamzyy * gyzz) +
              gupzz * (TWO * (Gamxyz * gxzy + Gamyyz * gyzy + Gamzyz * gzzy) + Gamxyz * gyzx +
                       Gamyyz * gyzy + Gamzyz * gyzz);
        Rzz = -HALF * Rzz + gxz * Gamxz + gyz * Gamyz + gzz * Gamzz + Gamxa * gxzz + Gamya * gyzz +
              Gamza * gzzz +
              gupxx * (TWO * (Gamxxz * gxxz + Gamyxz * gxyz + Gamzxz * gxzz) + Gamxxz * gxzx +
                       Gamyxz * gxzy + Gamzxz * gxzz) +
              gupxy * (TWO * (Gamxxz * gxyz + Gamyxz * gyyz + Gamzxz * gyzz + Gamxyz * gxxz +
                              Gamyyz * gxyz + Gamzyz * gxzz) +
                       Gamxyz * gxzx + Gamyyz * gxzy + Gamzyz * gxzz + Gamxxz * gyzx + Gamyxz * gyzy +
                       Gamzxz * gyzz) +
              gupxz * (TWO * (Gamxxz * gxzz + Gamyxz * gyzz + Gamzxz * gzzz + Gamxzz * gxxz +
                              Gamyzz * gxyz + Gamzzz * gxzz) +
                       Gamxzz * gxzx + Gamyzz * gxzy + Gamzzz * gxzz + Gamxxz * gzzx + Gamyxz * gzzy +
                       Gamzxz * gzzz) +
              gupyy * (TWO * (Gamxyz * gxyz + Gamyyz * gyyz + Gamzyz * gyzz) + Gamxyz * gyzx +
                       Gamyyz * gyzy + Gamzyz * gyzz) +
              gupyz * (TWO * (Gamxyz * gxzz + Gamyyz * gyzz + Gamzyz * gzzz + Gamxzz * gxyz +
                              Gamyzz * gyyz + Gamzzz * gyzz) +
                       Gamxzz * gyzx + Gamyzz * gyzy + Gamzzz * gyzz + Gamxyz * gzzx + Gamyyz * gzzy +
                       Gamzyz * gzzz) +
              gupzz * (TWO * (Gamxzz * gxzz + Gamyzz * gyzz + Gamzzz * gzzz) + Gamxzz * gzzx +
                       Gamyzz * gzzy + Gamzzz * gzzz);
        Rxy = HALF * (-Rxy + gxx * Gamxy + gxy * Gamyy + gxz * Gamzy + gxy * Gamxx + gyy * Gamyx +
                      gyz * Gamzx + Gamxa * gxyx + Gamya * gyyx + Gamza * gyzx + Gamxa * gxxy +
                      Gamya * gxyy + Gamza * gxzy) +
              gupxx * (Gamxxx * gxxy + Gamyxx * gxyy + Gamzxx * gxzy + Gamxxy * gxxx + Gamyxy * gxyx +
                       Gamzxy * gxzx + Gamxxx * gxyx + Gamyxx * gxyy + Gamzxx * gxyz) +
              gupxy * (Gamxxx * gxyy + Gamyxx * gyyy + Gamzxx * gyzy + Gamxxy * gxyx + Gamyxy * gyyx +
                       Gamzxy * gyzx + Gamxxy * gxyx + Gamyxy * gxyy + Gamzxy * gxyz + Gamxxy * gxxy +
                       Gamyxy * gxyy + Gamzxy * gxzy + Gamxyy * gxxx + Gamyyy * gxyx + Gamzyy * gxzx +
                       Gamxxx * gyyx + Gamyxx * gyyy + Gamzxx * gyyz) +
              gupxz * (Gamxxx * gxzy + Gamyxx * gyzy + Gamzxx * gzzy + Gamxxy * gxzx + Gamyxy * gyzx +
                       Gamzxy * gzzx + Gamxxz * gxyx + Gamyxz * gxyy + Gamzxz * gxyz + Gamxxz * gxxy +
                       Gamyxz * gxyy + Gamzxz * gxzy + Gamxyz * gxxx + Gamyyz * gxyx + Gamzyz * gxzx +
                       Gamxxx * gyzx + Gamyxx * gyzy + Gamzxx * gyzz) +
              gupyy * (Gamxxy * gxyy + Gamyxy * gyyy + Gamzxy * gyzy + Gamxyy * gxyx + Gamyyy * gyyx +
                       Gamzyy * gyzx + Gamxxy * gyyx + Gamyxy * gyyy + Gamzxy * gyyz) +
              gupyz * (Gamxxy * gxzy + Gamyxy * gyzy + Gamzxy * gzzy + Gamxyy * gxzx + Gamyyy * gyzx +
                       Gamzyy * gzzx + Gamxxz * gyyx + Gamyxz * gyyy + Gamzxz * gyyz + Gamxxz * gxyy +
                       Gamyxz * gyyy + Gamzxz * gyzy + Gamxyz * gxyx + Gamyyz * gyyx + Gamzyz * gyzx +
                       Gamxxy * gyzx + Gamyxy * gyzy + Gamzxy * gyzz) +
              gupzz * (Gamxxz * gxzy + Gamyxz * gyzy + Gamzxz * gzzy + Gamxyz * gxzx + Gamyyz * gyzx +
                       Gamzyz * gzzx + Gamxxz * gyzx + Gamyxz * gyzy + Gamzxz * gyzz);
        Rxz = HALF * (-Rxz + gxx * Gamxz + gxy * Gamyz + gxz * Gamzz + gxz * Gamxx + gyz * Gamyx +
                      gzz * Gamzx + Gamxa * gxzx + Gamya * gyzx + Gamza * gzzx + Gamxa * gxxz +
                      Gamya * gxyz + Gamza * gxzz) +
              gupxx * (Gamxxx * gxxz + Gamyxx * gxyz + Gamzxx * gxzz + Gamxxz * gxxx + Gamyxz * gxyx +
                       Gamzxz * gxzx + Gamxxx * gxzx + Gamyxx * gxzy + Gamzxx * gxzz) +
              gupxy * (Gamxxx * gxyz + Gamyxx * gyyz + Gamzxx * gyzz + Gamxxz * gxyx + Gamyxz * gyyx +
                       Gamzxz * gyzx + Gamxxy * gxzx + Gamyxy * gxzy + Gamzxy * gxzz + Gamxxy * gxxz +
                       Gamyxy * gxyz + Gamzxy * gxzz + Gamxyz * gxxx + Gamyyz * gxyx + Gamzyz * gxzx +
                       Gamxxx * gyzx + Gamyxx * gyzy + Gamzxx * gyzz) +
              gupxz * (Gamxxx * gxzz + Gamyxx * gyzz + Gamzxx * gzzz + Gamxxz * gxzx + Gamyxz * gyzx +
                       Gamzxz * gzzx + Gamxxz * gxzx + Gamyxz * gxzy + Gamzxz * gxzz + Gamxxz * gxxz +
                       Gamyxz * gxyz + Gamzxz * gxzz + Gamxzz * gxxx + Gamyzz * gxyx + Gamzzz * gxzx +
                       Gamxxx * gzzx + Gamyxx * gzzy + Gamzxx * gzzz) +
              gupyy * (Gamxxy * gxyz + Gamyxy * gyyz + Gamzxy * gyzz + Gamxyz * gxyx + Gamyyz * gyyx +
                       Gamzyz * gyzx + Gamxxy * gyzx + Gamyxy * gyzy + Gamzxy * gyzz) +
              gupyz * (Gamxxy * gxzz + Gamyxy * gyzz + Gamzxy * gzzz + Gamxyz * gxzx + Gamyyz * gyzx +
                       Gamzyz * gzzx + Gamxxz * gyzx + Gamyxz * gyzy + Gamzxz * gyzz + Gamxxz * gxyz +
                       Gamyxz * gyyz + Gamzxz * gyzz + Gamxzz * gxyx + Gamyzz * gyyx + Gamzzz * gyzx +
                       Gamxxy * gzzx + Gamyxy * gzzy + Gamzxy * gzzz) +
              gupzz * (Gamxxz * gxzz + Gamyxz * gyzz + Gamzxz * gzzz + Gamxzz * gxzx + Gamyzz * gyzx +
                       Gamzzz * gzzx + Gamxxz * gzzx + Gamyxz * gzzy + Gamzxz * gzzz);
        Ryz = HALF * (-Ryz + gxy * Gamxz + gyy * Gamyz + gyz * Gamzz + gxz * Gamxy + gyz * Gamyy +
                      gzz * Gamzy + Gamxa * gxzy + Gamya * gyzy + Gamza * gzzy + Gamxa * gxyz +
                      Gamya * gyyz + Gamza * gyzz) +
              gupxx * (Gamxxy * gxxz + Gamyxy * gxyz + Gamzxy * gxzz + Gamxxz * gxxy + Gamyxz * gxyy +
                       Gamzxz * gxzy + Gamxxy * gxzx + Gamyxy * gxzy + Gamzxy * gxzz) +
              gupxy * (Gamxxy * gxyz + Gamyxy * gyyz + Gamzxy * gyzz + Gamxxz * gxyy + Gamyxz * gyyy +
                       Gamzxz * gyzy + Gamxyy * gxzx + Gamyyy * gxzy + Gamzyy * gxzz + Gamxyy * gxxz +
                       Gamyyy * gxyz + Gamzyy * gxzz + Gamxyz * gxxy + Gamyyz * gxyy + Gamzyz * gxzy +
                       Gamxxy * gyzx + Gamyxy * gyzy + Gamzxy * gyzz) +
              gupxz * (Gamxxy * gxzz + Gamyxy * gyzz + Gamzxy * gzzz + Gamxxz * gxzy + Gamyxz * gyzy +
                       Gamzxz * gzzy + Gamxyz * gxzx + Gamyyz * gxzy + Gamzyz * gxzz + Gamxyz * gxxz +
                       Gamyyz * gxyz + Gamzyz * gxzz + Gamxzz * gxxy + Gamyzz * gxyy + Gamzzz * gxzy +
                       Gamxxy * gzzx + Gamyxy * gzzy + Gamzxy * gzzz) +
              gupyy * (Gamxyy * gxyz + Gamyyy * gyyz + Gamzyy * gyzz + Gamxyz * gxyy + Gamyyz * gyyy +
                       Gamzyz * gyzy + Gamxyy * gyzx + Gamyyy * gyzy + Gamzyy * gyzz) +
              gupyz * (Gamxyy * gxzz + Gamyyy * gyzz + Gamzyy * gzzz + Gamxyz * gxzy + Gamyyz * gyzy +
                       Gamzyz * gzzy + Gamxyz * gyzx + Gamyyz * gyzy + Gamzyz * gyzz + Gamxyz * gxyz +
                       Gamyyz * gyyz + Gamzyz * gyzz + Gamxzz * gxyy + Gamyzz * gyyy + Gamzzz * gyzy +
                       Gamxyy * gzzx + Gamyyy * gzzy + Gamzyy * gzzz) +
              gupzz * (Gamxyz * gxzz + Gamyyz * gyzz + Gamzyz * gzzz + Gamxzz * gxzy + Gamyzz * gyzy +
                       Gamzzz * gzzy + Gamxyz * gzzx + Gamyyz * gzzy + Gamzyz * gzzz);
        /* :603-626 chi part of the Ricci tensor */
        D2(CHI, fxx, fxy, fxz, fyy, fyz, fzz);
        fxx = fxx - Gamxxx * chix - Gamyxx * chiy - Gamzxx * chiz;
        fxy = fxy - Gamxxy * chix - Gamyxy * chiy - Gamzxy * chiz;
        fxz = fxz - Gamxxz * chix - Gamyxz * chiy - Gamzxz * chiz;
        fyy = fyy - Gamxyy * chix - Gamyyy * chiy - Gamzyy * chiz;
        fyz = fyz - Gamxyz * chix - Gamyyz * chiy - Gamzyz * chiz;
        fzz = fzz - Gamxzz * chix - Gamyzz * chiy - Gamzzz * chiz;
        double f = gupxx * (fxx - F3o2 / chin1 * chix * chix) + gupyy * (fyy - F3o2 / chin1 * chiy * chiy) +
                   gupzz * (fzz - F3o2 / chin1 * chiz * chiz) + TWO * gupxy * (fxy - F3o2 / chin1 * chix * chiy) +
                   TWO * gupxz * (fxz - F3o2 / chin1 * chix * chiz) + TWO * gupyz * (fyz - F3o2 / chin1 * chiy * chiz);
        Rxx = Rxx + (fxx - chix * chix / chin1 / TWO + gxx * f) / chin1 / TWO;
        Ryy = Ryy + (fyy - chiy * chiy / chin1 / TWO + gyy * f) / chin1 / TWO;
        Rzz = Rzz + (fzz - chiz * chiz / chin1 / TWO + gzz * f) / chin1 / TWO;
        Rxy = Rxy + (fxy - chix * chiy / chin1 / TWO + gxy * f) / chin1 / TWO;
        Rxz = Rxz + (fxz - chix * chiz / chin1 / TWO + gxz * f) / chin1 / TWO;
        Ryz = Ryz + (fyz - chiy * chiz / chin1 / TWO + gyz * f) / chin1 / TWO;
        /* :629-664 lapse: physical Christoffels, D_i D_j alpha */
        D2(LAP, fxx, fxy, fxz, fyy, fyz, fzz);
        gxxx = (gupxx * chix + gupxy * chiy + gupxz * chiz) / chin1;
        gxxy = (gupxy * chix + gupyy * chiy + gupyz * chiz) / chin1;
        gxxz = (gupxz * chix + gupyz * chiy + gupzz * chiz) / chin1;
        Gamxxx = Gamxxx - ((chix + chix) / chin1 - gxx * gxxx) * HALF;
        Gamyxx = Gamyxx - (-gxx * gxxy) * HALF;
        Gamzxx = Gamzxx - (-gxx * gxxz) * HALF;
        Gamxyy = Gamxyy - (-gyy * gxxx) * HALF;
        Gamyyy = Gamyyy - ((chiy + chiy) / chin1 - gyy * gxxy) * HALF;
        Gamzyy = Gamzyy - (-gyy * gxxz) * HALF;
        Gamxzz = Gamxzz - (-gzz * gxxx) * HALF;
        Gamyzz = Gamyzz - (-gzz * gxxy) * HALF;
        Gamzzz = Gamzzz - ((chiz + chiz) / chin1 - gzz * gxxz) * HALF;
        Gamxxy = Gamxxy - (chiy / chin1 - gxy * gxxx) * HALF;
        Gamyxy = Gamyxy - (chix / chin1 - gxy * gxxy) * HALF;
        Gamzxy = Gamzxy - (-gxy * gxxz) * HALF;
        Gamxxz = Gamxxz - (chiz / chin1 - gxz * gxxx) * HALF;
        Gamyxz = Gamyxz - (-gxz * gxxy) * HALF;
        Gamzxz = Gamzxz - (chix / chin1 - gxz * gxxz) * HALF;
        Gamxyz = Gamxyz - (-gyz * gxxx) * HALF;
        Gamyyz = Gamyyz - (chiz / chin1 - gyz * gxxy) * HALF;
        Gamzyz = Gamzyz - (chiy / chin1 - gyz * gxxz) * HALF;
        fxx = fxx - Gamxxx * Lapx - Gamyxx * Lapy - Gamzxx * Lapz;
        fyy = fyy - Gamxyy * Lapx - Gamyyy * Lapy - Gamzyy * Lapz;
        fzz = fzz - Gamxzz * Lapx - Gamyzz * Lapy - Gamzzz * Lapz;
        fxy = fxy - Gamxxy * Lapx - Gamyxy * Lapy - Gamzxy * Lapz;
        fxz = fxz - Gamxxz * Lapx - Gamyxz * Lapy - Gamzxz * Lapz;
        fyz = fyz - Gamxyz * Lapx - Gamyyz * Lapy - Gamzyz * Lapz;
        double trK_rhs = gupxx * fxx + gupyy * fyy + gupzz * fzz + TWO * (gupxy * fxy + gupxz * fxz + gupyz * fyz);
        /* :666-703 matter trace and the trace-free source (BAM form) */
        double S = chin1 * (gupxx * Sxx + gupyy * Syy + gupzz * Szz + TWO * (gupxy * Sxy + gupxz * Sxz + gupyz * Syz));
        f = F2o3 * trK * trK -
            (gupxx * (gupxx * Axx * Axx + gupyy * Axy * Axy + gupzz * Axz * Axz +
                      TWO * (gupxy * Axx * Axy + gupxz * Axx * Axz + gupyz * Axy * Axz)) +
             gupyy * (gupxx * Axy * Axy + gupyy * Ayy * Ayy + gupzz * Ayz * Ayz +
                      TWO * (gupxy * Axy * Ayy + gupxz * Axy * Ayz + gupyz * Ayy * Ayz)) +
             gupzz * (gupxx * Axz * Axz + gupyy * Ayz * Ayz + gupzz * Azz * Azz +
                      TWO * (gupxy * Axz * Ayz + gupxz * Axz * Azz + gupyz * Ayz * Azz)) +
             TWO * (gupxy * (gupxx * Axx * Axy + gupyy * Axy * Ayy + gupzz * Axz * Ayz +
                             gupxy * (Axx * Ayy + Axy * Axy) + gupxz * (Axx * Ayz + Axz * Axy) +
                             gupyz * (Axy * Ayz + Axz * Ayy)) +
                    gupxz * (gupxx * Axx * Axz + gupyy * Axy * Ayz + gupzz * Axz * Azz +
                             gupxy * (Axx * Ayz + Axy * Axz) + gupxz * (Axx * Azz + Axz * Axz) +
                             gupyz * (Axy * Azz + Axz * Ayz)) +
                    gupyz * (gupxx * Axy * Axz + gupyy * Ayy * Ayz + gupzz * Ayz * Azz +
                             gupxy * (Axy * Ayz + Ayy * Axz) + gupxz * (Axy * Azz + Ayz * Axz) +
                             gupyz * (Ayy * Azz + Ayz * Ayz)))) -
            1.6e1 * PI * rho + EIGHT * PI * S;
        f = -F1o3 * (gupxx * fxx + gupyy * fyy + gupzz * fzz + TWO * (gupxy * fxy + gupxz * fxz + gupyz * fyz) +
                     alpn1 / chin1 * f);
        fxx = alpn1 * (Rxx - EIGHT * PI * Sxx) - fxx;
        fxy = alpn1 * (Rxy - EIGHT * PI * Sxy) - fxy;
        fxz = alpn1 * (Rxz - EIGHT * PI * Sxz) - fxz;
        fyy = alpn1 * (Ryy - EIGHT * PI * Syy) - fyy;
        fyz = alpn1 * (Ryz - EIGHT * PI * Syz) - fyz;
        fzz = alpn1 * (Rzz - EIGHT * PI * Szz) - fzz;
        /* :720-777 */
        double Axx_rhs = fxx - gxx * f;
        double Ayy_rhs = fyy - gyy * f;
        double Azz_rhs = fzz - gzz * f;
        double Axy_rhs = fxy - gxy * f;
        double Axz_rhs = fxz - gxz * f;
        double Ayz_rhs = fyz - gyz * f;
        fxx = gupxx * Axx * Axx + gupyy * Axy * Axy + gupzz * Axz * Axz +
              TWO * (gupxy * Axx * Axy + gupxz * Axx * Axz + gupyz * Axy * Axz);
        fyy = gupxx * Axy * Axy + gupyy * Ayy * Ayy + gupzz * Ayz * Ayz +
              TWO * (gupxy * Axy * Ayy + gupxz * Axy * Ayz + gupyz * Ayy * Ayz);
        fzz = gupxx * Axz * Axz + gupyy * Ayz * Ayz + gupzz * Azz * Azz +
              TWO * (gupxy * Axz * Ayz + gupxz * Axz * Azz + gupyz * Ayz * Azz);
        fxy = gupxx * Axx * Axy + gupyy * Axy * Ayy + gupzz * Axz * Ayz + gupxy * (Axx * Ayy + Axy * Axy) +
              gupxz * (Axx * Ayz + Axz * Axy) + gupyz * (Axy * Ayz + Axz * Ayy);
        fxz = gupxx * Axx * Axz + gupyy * Axy * Ayz + gupzz * Axz * Azz + gupxy * (Axx * Ayz + Axy * Axz) +
              gupxz * (Axx * Azz + Axz * Axz) + gupyz * (Axy * Azz + Axz * Ayz);
        fyz = gupxx * Axy * Axz + gupyy * Ayy * Ayz + gupzz * Ayz * Azz + gupxy * (Axy * Ayz + Ayy * Axz) +
              gupxz * (Axy * Azz + Ayz * Axz) + gupyz * (Ayy * Azz + Ayz * Ayz);
        f = chin1;
        trK_rhs = f * trK_rhs;
        Axx_rhs = f * Axx_rhs + alpn1 * (trK * Axx - TWO * fxx) +
                  TWO * (Axx * betaxx + Axy * betayx + Axz * betazx) - F2o3 * Axx * div_beta;
        Ayy_rhs = f * Ayy_rhs + alpn1 * (trK * Ayy - TWO * fyy) +
                  TWO * (Axy * betaxy + Ayy * betayy + Ayz * betazy) - F2o3 * Ayy * div_beta;
        Azz_rhs = f * Azz_rhs + alpn1 * (trK * Azz - TWO * fzz) +
                  TWO * (Axz * betaxz + Ayz * betayz + Azz * betazz) - F2o3 * Azz * div_beta;
        Axy_rhs = f * Axy_rhs + alpn1 * (trK * Axy - TWO * fxy) + Axx * betaxy + Axz * betazy + Ayy * betayx +
                  Ayz * betazx + F1o3 * Axy * div_beta - Axy * betazz;
        Ayz_rhs = f * Ayz_rhs + alpn1 * (trK * Ayz - TWO * fyz) + Axy * betaxz + Ayy * betayz + Axz * betaxy +
                  Azz * betazy + F1o3 * Ayz * div_beta - Ayz * betaxx;
        Axz_rhs = f * Axz_rhs + alpn1 * (trK * Axz - TWO * fxz) + Axx * betaxz + Axy * betayz + Ayz * betayx +
                  Azz * betazx + F1o3 * Axz * div_beta - Axz * betayy;
        /* :780-787 */
        S = f * (gupxx * Sxx + gupyy * Syy + gupzz * Szz + TWO * (gupxy * Sxy + gupxz * Sxz + gupyz * Syz));
        trK_rhs = -trK_rhs + alpn1 * (F1o3 * trK * trK + gupxx * fxx + gupyy * fyy + gupzz * fzz +
                                      TWO * (gupxy * fxy + gupxz * fxz + gupyz * fyz) + FOUR * PI * (rho + S));
        /* :791-820 gauge */
        double Lap_rhs = -TWO * alpn1 * trK;
        double betax_rhs = FF * dtSfx, betay_rhs = FF * dtSfy, betaz_rhs = FF * dtSfz;
        double dtSfx_rhs, dtSfy_rhs, dtSfz_rhs;
        if (gauge == 2) {
            /* fderivs of chi again into the dtSf_rhs slots (same values as chix..) */
            double cx, cy, cz;
            D1(CHI, cx, cy, cz);
            double reta = gupxx * cx * cx + gupyy * cy * cy + gupzz * cz * cz +
                          TWO * (gupxy * cx * cy + gupxz * cx * cz + gupyz * cy * cz);
            reta = 1.31 / 2 * sqrt(reta / chin1) / ((1 - sqrt(chin1)) * (1 - sqrt(chin1)));
            dtSfx_rhs = Gamx_rhs - reta * dtSfx;
            dtSfy_rhs = Gamy_rhs - reta * dtSfy;
            dtSfz_rhs = Gamz_rhs - reta * dtSfz;
        } else {
            dtSfx_rhs = Gamx_rhs - eta * dtSfx;
            dtSfy_rhs = Gamy_rhs - eta * dtSfy;
            dtSfz_rhs = Gamz_rhs - eta * dtSfz;
        }
        /* :948-980 advection, then :983-1042 KO dissipation */
        double r[NF] = {chi_rhs, trK_rhs, gxx_rhs, gxy_rhs, gxz_rhs, gyy_rhs, gyz_rhs, gzz_rhs,
                        Axx_rhs, Axy_rhs, Axz_rhs, Ayy_rhs, Ayz_rhs, Azz_rhs,
                        Gamx_rhs, Gamy_rhs, Gamz_rhs, Lap_rhs, betax_rhs, betay_rhs, betaz_rhs,
                        dtSfx_rhs, dtSfy_rhs, dtSfz_rhs};
        for (int v = 0; v < NF; v++) {
            const double off = (v == DXX || v == DYY || v == DZZ) ? ONE : 0.0;
            r[v] = lop_pt(&g, st[v], off, PAR[v], i, j, k, betax, betay, betaz, r[v]);
        }
        if (eps > 0)
            for (int v = 0; v < NF; v++) r[v] = ko_pt(&g, st[v], PAR[v], i, j, k, eps, r[v]);
        for (int v = 0; v < NF; v++) rhs[v][p] = r[v];

        const double G2[18] = {Gamxxx, Gamxxy, Gamxxz, Gamxyy, Gamxyz, Gamxzz,
                               Gamyxx, Gamyxy, Gamyxz, Gamyyy, Gamyyz, Gamyzz,
                               Gamzxx, Gamzxy, Gamzxz, Gamzyy, Gamzyz, Gamzzz};
        for (int a = 0; a < 18; a++) gamo[a][p] = G2[a];
        rico[0][p] = Rxx; rico[1][p] = Rxy; rico[2][p] = Rxz;
        rico[3][p] = Ryy; rico[4][p] = Ryz; rico[5][p] = Rzz;

        /* :1044-1134 constraint residuals */
        if (co == 0) {
            double ham = gupxx * Rxx + gupyy * Ryy + gupzz * Rzz + TWO * (gupxy * Rxy + gupxz * Rxz + gupyz * Ryz);
            ham = chin1 * ham + F2o3 * trK * trK -
                  (gupxx * (gupxx * Axx * Axx + gupyy * Axy * Axy + gupzz * Axz * Axz +
                            TWO * (gupxy * Axx * Axy + gupxz * Axx * Axz + gupyz * Axy * Axz)) +
                   gupyy * (gupxx * Axy * Axy + gupyy * Ayy * Ayy + gupzz * Ayz * Ayz +
                            TWO * (gupxy * Axy * Ayy + gupxz * Axy * Ayz + gupyz * Ayy * Ayz)) +
                   gupzz * (gupxx * Axz * Axz + gupyy * Ayz * Ayz + gupzz * Azz * Azz +
                            TWO * (gupxy * Axz * Ayz + gupxz * Axz * Azz + gupyz * Ayz * Azz)) +
                   TWO * (gupxy * (gupxx * Axx * Axy + gupyy * Axy * Ayy + gupzz * Axz * Ayz +
                                   gupxy * (Axx * Ayy + Axy * Axy) + gupxz * (Axx * Ayz + Axz * Axy) +
                                   gupyz * (Axy * Ayz + Axz * Ayy)) +
                          gupxz * (gupxx * Axx * Axz + gupyy * Axy * Ayz + gupzz * Axz * Azz +
                                   gupxy * (Axx * Ayz + Axy * Axz) + gupxz * (Axx * Azz + Axz * Axz) +
                                   gupyz * (Axy * Azz + Axz * Ayz)) +
                          gupyz * (gupxx * Axy * Axz + gupyy * Ayy * Ayz + gupzz * Ayz * Azz +
                                   gupxy * (Axy * Ayz + Ayy * Axz) + gupxz * (Axy * Azz + Ayz * Axz) +
                                   gupyz * (Ayy * Azz + Ayz * Ayz)))) -
                  F16 * PI * rho;
            res[0][p] = ham;
            D1(AXX, gxxx, gxxy, gxxz);
            D1(AXY, gxyx, gxyy, gxyz);
            D1(AXZ, gxzx, gxzy, gxzz);
            D1(AYY, gyyx, gyyy, gyyz);
            D1(AYZ, gyzx, gyzy, gyzz);
            D1(AZZ, gzzx, gzzy, gzzz);
            gxxx = gxxx - (Gamxxx * Axx + Gamyxx * Axy + Gamzxx * Axz + Gamxxx * Axx + Gamyxx * Axy + Gamzxx * Axz) - chix * Axx / chin1;
            gxyx = gxyx - (Gamxxy * Axx + Gamyxy * Axy + Gamzxy * Axz + Gamxxx * Axy + Gamyxx * Ayy + Gamzxx * Ayz) - chix * Axy / chin1;
            gxzx = gxzx - (Gamxxz * Axx + Gamyxz * Axy + Gamzxz * Axz + Gamxxx * Axz + Gamyxx * Ayz + Gamzxx * Azz) - chix * Axz / chin1;
            gyyx = gyyx - (Gamxxy * Axy + Gamyxy * Ayy + Gamzxy * Ayz + Gamxxy * Axy + Gamyxy * Ayy + Gamzxy * Ayz) - chix * Ayy / chin1;
            gyzx = gyzx - (Gamxxz * Axy + Gamyxz * Ayy + Gamzxz * Ayz + Gamxxy * Axz + Gamyxy * Ayz + Gamzxy * Azz) - chix * Ayz / chin1;
            gzzx = gzzx - (Gamxxz * Axz + Gamyxz * Ayz + Gamzxz * Azz + Gamxxz * Axz + Gamyxz * Ayz + Gamzxz * Azz) - chix * Azz / chin1;
            gxxy = gxxy - (Gamxxy * Axx + Gamyxy * Axy + Gamzxy * Axz + Gamxxy * Axx + Gamyxy * Axy + Gamzxy * Axz) - chiy * Axx / chin1;
            gxyy = gxyy - (Gamxyy * Axx + Gamyyy * Axy + Gamzyy * Axz + Gamxxy * Axy + Gamyxy * Ayy + Gamzxy * Ayz) - chiy * Axy / chin1;
            gxzy = gxzy - (Gamxyz * Axx + Gamyyz * Axy + Gamzyz * Axz + Gamxxy * Axz + Gamyxy * Ayz + Gamzxy * Azz) - chiy * Axz / chin1;
            gyyy = gyyy - (Gamxyy * Axy + Gamyyy * Ayy + Gamzyy * Ayz + Gamxyy * Axy + Gamyyy * Ayy + Gamzyy * Ayz) - chiy * Ayy / chin1;
            gyzy = gyzy - (Gamxyz * Axy + Gamyyz * Ayy + Gamzyz * Ayz + Gamxyy * Axz + Gamyyy * Ayz + Gamzyy * Azz) - chiy * Ayz / chin1;
            gzzy = gzzy - (Gamxyz * Axz + Gamyyz * Ayz + Gamzyz * Azz + Gamxyz * Axz + Gamyyz * Ayz + Gamzyz * Azz) - chiy * Azz / chin1;
            gxxz = gxxz - (Gamxxz * Axx + Gamyxz * Axy + Gamzxz * Axz + Gamxxz * Axx + Gamyxz * Axy + Gamzxz * Axz) - chiz * Axx / chin1;
            gxyz = gxyz - (Gamxyz * Axx + Gamyyz * Axy + Gamzyz * Axz + Gamxxz * Axy + Gamyxz * Ayy + Gamzxz * Ayz) - chiz * Axy / chin1;
            gxzz = gxzz - (Gamxzz * Axx + Gamyzz * Axy + Gamzzz * Axz + Gamxxz * Axz + Gamyxz * Ayz + Gamzxz * Azz) - chiz * Axz / chin1;
            gyyz = gyyz - (Gamxyz * Axy + Gamyyz * Ayy + Gamzyz * Ayz + Gamxyz * Axy + Gamyyz * Ayy + Gamzyz * Ayz) - chiz * Ayy / chin1;
            gyzz = gyzz - (Gamxzz * Axy + Gamyzz * Ayy + Gamzzz * Ayz + Gamxyz * Axz + Gamyyz * Ayz + Gamzyz * Azz) - chiz * Ayz / chin1;
            gzzz = gzzz - (Gamxzz * Axz + Gamyzz * Ayz + Gamzzz * Azz + Gamxzz * Axz + Gamyzz * Ayz + Gamzzz * Azz) - chiz * Azz / chin1;
            double movx = gupxx * gxxx + gupyy * gxyy + gupzz * gxzz + gupxy * gxyx + gupxz * gxzx + gupyz * gxzy +
                          gupxy * gxxy + gupxz * gxxz + gupyz * gxyz;
            double movy = gupxx * gxyx + gupyy * gyyy + gupzz * gyzz + gupxy * gyyx + gupxz * gyzx + gupyz * gyzy +
                          gupxy * gxyy + gupxz * gxyz + gupyz * gyyz;
            double movz = gupxx * gxzx + gupyy * gyzy + gupzz * gzzz + gupxy * gyzx + gupxz * gzzx + gupyz * gzzy +
                          gupxy * gxzy + gupxz * gxzz + gupyz * gyzz;
            res[1][p] = movx - F2o3 * Kx - F8 * PI * Sx;
            res[2][p] = movy - F2o3 * Ky - F8 * PI * Sy;
            res[3][p] = movz - F2o3 * Kz - F8 * PI * Sz;
        }
#undef D1
#undef D2
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
void orc_enforce_ga(const int *ex, double *const *g, double *const *a)
{
    const size_t N = (size_t)ex[0] * ex[1] * ex[2];
    const double F1o3 = 1.0 / 3.0, ONE = 1.0, TWO = 2.0;
    for (size_t p = 0; p < N; p++) {
        double gxx = g[0][p] + ONE, gxy = g[1][p], gxz = g[2][p];
        double gyy = g[3][p] + ONE, gyz = g[4][p], gzz = g[5][p] + ONE;
        double gupzz = gxx * gyy * gzz + gxy * gyz * gxz + gxz * gxy * gyz - gxz * gyy * gxz -
                       gxy * gxy * gzz - gxx * gyz * gyz;
        gupzz = ONE / pow(gupzz, F1o3);
        gxx = gxx * gupzz; gxy = gxy * gupzz; gxz = gxz * gupzz;
        gyy = gyy * gupzz; gyz = gyz * gupzz; gzz = gzz * gupzz;
        g[0][p] = gxx - ONE; g[1][p] = gxy; g[2][p] = gxz;
        g[3][p] = gyy - ONE; g[4][p] = gyz; g[5][p] = gzz - ONE;
        const double gupxx = (gyy * gzz - gyz * gyz);
        const double gupxy = -(gxy * gzz - gyz * gxz);
        const double gupxz = (gxy * gyz - gyy * gxz);
        const double gupyy = (gxx * gzz - gxz * gxz);
        const double gupyz = -(gxx * gyz - gxy * gxz);
        gupzz = (gxx * gyy - gxy * gxy);
        const double trA = gupxx * a[0][p] + gupyy * a[3][p] + gupzz * a[5][p] +
                           TWO * (gupxy * a[1][p] + gupxz * a[2][p] + gupyz * a[4][p]);
        a[0][p] = a[0][p] - F1o3 * gxx * trA;
        a[1][p] = a[1][p] - F1o3 * gxy * trA;
        a[2][p] = a[2][p] - F1o3 * gxz * trA;
        a[3][p] = a[3][p] - F1o3 * gyy * trA;
        a[4][p] = a[4][p] - F1o3 * gyz * trA;
        a[5][p] = a[5][p] - F1o3 * gzz * trA;
    }
}

int orc_rungekutta4_rout(const int *ex, double dT, const double *f0, double *f1, double *f_rhs, int RK4)
{
    const size_t N = (size_t)ex[0] * ex[1] * ex[2];
    const double F1o6 = 1.0 / 6.0, HLF = 0.5, TWO = 2.0;
    if (RK4 < 0 || RK4 > 3) return 1;
    for (size_t p = 0; p < N; p++) {
        if (RK4 == 0) {
            f1[p] = f0[p] + HLF * dT * f_rhs[p];
        } else if (RK4 == 1) {
            f_rhs[p] = f_rhs[p] + TWO * f1[p];
            f1[p] = f0[p] + HLF * dT * f1[p];
        } else if (RK4 == 2) {
            f_rhs[p] = f_rhs[p] + TWO * f1[p];
            f1[p] = f0[p] + dT * f1[p];
        } else {
            f1[p] = f0[p] + F1o6 * dT * (f1[p] + f_rhs[p]);
        }
    }
    return 0;
}

void orc_lowerboundset(const int *ex, double *chi, double tinny)
{
    const size_t N = (size_t)ex[0] * ex[1] * ex[2];
    for (size_t p = 0; p < N; p++)
        if (chi[p] < tinny) chi[p] = tinny;
}

double orc_l2normhelper(const int *ex, const double *X, const double *Y, const double *Z, double xmin,
                        double ymin, double zmin, double xmax, double ymax, double zmax, const double *f,
                        int gw)
{
    const double dX = X[1] - X[0], dY = Y[1] - Y[0], dZ = Z[1] - Z[0];
    int imin = gw, jmin = gw, kmin = gw;
    int imax = ex[0] - gw - 1, jmax = ex[1] - gw - 1, kmax = ex[2] - gw - 1;
    if (fabs(X[ex[0] - 1] - xmax) < dX) imax = ex[0] - 1;
    if (fabs(Y[ex[1] - 1] - ymax) < dY) jmax = ex[1] - 1;
    if (fabs(Z[ex[2] - 1] - zmax) < dZ) kmax = ex[2] - 1;
    if (fabs(X[0] - xmin) < dX) imin = 0;
    if (fabs(Y[0] - ymin) < dY) jmin = 0;
    if (fabs(Z[0] - zmin) < dZ) kmin = 0;
    double s = 0.0;
    for (int k = kmin; k <= kmax; k++)
        for (int j = jmin; j <= jmax; j++)
            for (int i = imin; i <= imax; i++) {
                const double v = f[IX(ex, i, j, k)];
                s += v * v;
            }
    return s * dX * dY * dZ;
}

void orc_sommerfeld_copy(const int *ex, const double *X, const double *Y, const double *Z, const double *bbox,
                         const double *f0, double *f, int symmetry)
{
    const double dX = X[1] - X[0], dY = Y[1] - Y[0], dZ = Z[1] - Z[0];
    /* layer boxes (0-based, inclusive) for the six faces: sommerfeld_rout.f90:61-119 */
    int lay[6][6];
    int on[6] = {0, 0, 0, 0, 0, 0};
    const int n0 = ex[0] - 1, n1 = ex[1] - 1, n2 = ex[2] - 1;
    if (fabs(X[n0] - bbox[3]) < dX) { on[0] = 1; int t[6] = {n0, 0, 0, n0, n1, n2}; memcpy(lay[0], t, sizeof t); }
    if (fabs(Y[n1] - bbox[4]) < dY) { on[1] = 1; int t[6] = {0, n1, 0, n0, n1, n2}; memcpy(lay[1], t, sizeof t); }
    if (fabs(Z[n2] - bbox[5]) < dZ) { on[2] = 1; int t[6] = {0, 0, n2, n0, n1, n2}; memcpy(lay[2], t, sizeof t); }
    if (fabs(X[0] - bbox[0]) < dX && !(symmetry == 2 && fabs(bbox[0]) < dX / 2)) {
        on[3] = 1; int t[6] = {0, 0, 0, 0, n1, n2}; memcpy(lay[3], t, sizeof t); }
    if (fabs(Y[0] - bbox[1]) < dY && !(symmetry == 2 && fabs(bbox[1]) < dY / 2)) {
        on[4] = 1; int t[6] = {0, 0, 0, n0, 0, n2}; memcpy(lay[4], t, sizeof t); }
    if (fabs(Z[0] - bbox[2]) < dZ && !(symmetry > 0 && fabs(bbox[2]) < dZ / 2)) {
        on[5] = 1; int t[6] = {0, 0, 0, n0, n1, 0}; memcpy(lay[5], t, sizeof t); }
    for (int gp = 0; gp < 6; gp++) {
        if (!on[gp]) continue;
        for (int k = lay[gp][2]; k <= lay[gp][5]; k++)
            for (int j = lay[gp][1]; j <= lay[gp][4]; j++)
                for (int i = lay[gp][0]; i <= lay[gp][3]; i++) f[IX(ex, i, j, k)] = f0[IX(ex, i, j, k)];
    }
}

/* getnp4.f90:10-452 at every point, statement by statement (the Fortran's
 * whole-array statements only combine values of one point, except the
 * fderivs calls, which are taken first). */
void orc_getnp4(const int *ex, const double *X, const double *Y, const double *Z, const double *chi,
                const double *trK, const double *const *g, const double *const *A, const double *const *Gam,
                const double *const *R, double *Rpsi4, double *Ipsi4, int symmetry)
{
    const size_t N = (size_t)ex[0] * ex[1] * ex[2];
    const double SS[3] = {1.0, 1.0, 1.0};
    /* parities of A_ij (getnp4.f90:343-348) */
    const double PA[6][3] = {{1, 1, 1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}, {1, -1, -1}, {1, 1, 1}};
    double *d = malloc(24 * N * sizeof(double));
    double *dA[6][3], *dchi[3], *dK[3];
    for (int s = 0; s < 6; s++)
        for (int a = 0; a < 3; a++) dA[s][a] = d + (3 * s + a) * N;
    for (int a = 0; a < 3; a++) { dchi[a] = d + (18 + a) * N; dK[a] = d + (21 + a) * N; }
    for (int s = 0; s < 6; s++)
        orc_fderivs(ex, A[s], dA[s][0], dA[s][1], dA[s][2], X, Y, Z, PA[s][0], PA[s][1], PA[s][2], symmetry);
    orc_fderivs(ex, chi, dchi[0], dchi[1], dchi[2], X, Y, Z, SS[0], SS[1], SS[2], symmetry);
    orc_fderivs(ex, trK, dK[0], dK[1], dK[2], X, Y, Z, SS[0], SS[1], SS[2], symmetry);
    const double TINYRR = 1.e-14, F1o3 = 1.0 / 3.0, TWO = 2.0;
    for (int k = 0; k < ex[2]; k++)
        for (int j = 0; j < ex[1]; j++)
            for (int i = 0; i < ex[0]; i++) {
                const size_t p = IX(ex, i, j, k);
                const double gxx = g[0][p] + 1.0, gxy = g[1][p], gxz = g[2][p];
                const double gyy = g[3][p] + 1.0, gyz = g[4][p], gzz = g[5][p] + 1.0;
                const double chipn1 = chi[p] + 1.0;
                const double Axx = A[0][p], Axy = A[1][p], Axz = A[2][p], Ayy = A[3][p], Ayz = A[4][p],
                             Azz = A[5][p];
                const double tK = trK[p];
                /* Christoffels in compute_rhs_bssn's argument order */
                const double Gamxxx = Gam[0][p], Gamxxy = Gam[1][p], Gamxxz = Gam[2][p], Gamxyy = Gam[3][p],
                             Gamxyz = Gam[4][p], Gamxzz = Gam[5][p], Gamyxx = Gam[6][p], Gamyxy = Gam[7][p],
                             Gamyxz = Gam[8][p], Gamyyy = Gam[9][p], Gamyyz = Gam[10][p], Gamyzz = Gam[11][p],
                             Gamzxx = Gam[12][p], Gamzxy = Gam[13][p], Gamzxz = Gam[14][p], Gamzyy = Gam[15][p],
                             Gamzyz = Gam[16][p], Gamzzz = Gam[17][p];
                /* invert the conformal metric (:97-105) */
                double gupzz = gxx * gyy * gzz + gxy * gyz * gxz + gxz * gxy * gyz - gxz * gyy * gxz -
                               gxy * gxy * gzz - gxx * gyz * gyz;
                const double gupxx = (gyy * gzz - gyz * gyz) / gupzz;
                const double gupxy = -(gxy * gzz - gyz * gxz) / gupzz;
                const double gupxz = (gxy * gyz - gyy * gxz) / gupzz;
                const double gupyy = (gxx * gzz - gxz * gxz) / gupzz;
                const double gupyz = -(gxx * gyz - gxy * gxz) / gupzz;
                gupzz = (gxx * gyy - gxy * gxy) / gupzz;
                /* the radial tetrad (:259-341) */
                double vx, vy, vz, ux, uy, uz, wx, wy, wz, fx, fy, fz, f;
                if (fabs(X[i]) < TINYRR && fabs(Y[j]) < TINYRR && fabs(Z[k]) < TINYRR) {
                    vx = TINYRR; vy = TINYRR; vz = TINYRR;
                } else {
                    vx = X[i]; vy = Y[j]; vz = Z[k];
                }
                if (fabs(X[i]) < TINYRR && fabs(Y[j]) < TINYRR) {
                    ux = -TINYRR; uy = TINYRR; uz = 0.0;
                    wx = TINYRR * Z[k]; wy = TINYRR * Z[k]; wz = -2 * TINYRR * TINYRR;
                } else {
                    ux = -Y[j]; uy = X[i]; uz = 0.0;
                    wx = X[i] * Z[k]; wy = Y[j] * Z[k]; wz = -(X[i] * X[i] + Y[j] * Y[j]);
                }
                fx = vx; fy = vy; fz = vz;
                vx = gupxx * fx + gupxy * fy + gupxz * fz;
                vy = gupxy * fx + gupyy * fy + gupyz * fz;
                vz = gupxz * fx + gupyz * fy + gupzz * fz;
                f = 1.0 / chipn1;
                fx = gxx * vx * vx + gyy * vy * vy + gzz * vz * vz + (gxy * vx * vy + gxz * vx * vz + gyz * vy * vz) * TWO;
                fx = sqrt(fx * f);
                vx = vx / fx; vy = vy / fx; vz = vz / fx;
                fx = gxx * vx * ux + gxy * vx * uy + gxz * vx * uz + gxy * vy * ux + gyy * vy * uy + gyz * vy * uz +
                     gxz * vz * ux + gyz * vz * uy + gzz * vz * uz;
                fx = fx * f;
                ux = ux - fx * vx; uy = uy - fx * vy; uz = uz - fx * vz;
                fx = gxx * ux * ux + gyy * uy * uy + gzz * uz * uz + (gxy * ux * uy + gxz * ux * uz + gyz * uy * uz) * TWO;
                fx = sqrt(fx * f);
                ux = ux / fx; uy = uy / fx; uz = uz / fx;
                fx = gxx * vx * wx + gxy * vx * wy + gxz * vx * wz + gxy * vy * wx + gyy * vy * wy + gyz * vy * wz +
                     gxz * vz * wx + gyz * vz * wy + gzz * vz * wz;
                fx = fx * f;
                wx = wx - fx * vx; wy = wy - fx * vy; wz = wz - fx * vz;
                fx = gxx * ux * wx + gxy * ux * wy + gxz * ux * wz + gxy * uy * wx + gyy * uy * wy + gyz * uy * wz +
                     gxz * uz * wx + gyz * uz * wy + gzz * uz * wz;
                fx = fx * f;
                wx = wx - fx * ux; wy = wy - fx * uy; wz = wz - fx * uz;
                fx = gxx * wx * wx + gyy * wy * wy + gzz * wz * wz + (gxy * wx * wy + gxz * wx * wz + gyz * wy * wz) * TWO;
                fx = sqrt(fx * f);
                wx = wx / fx; wy = wy / fx; wz = wz / fx;
                /* D_k K_ij up to chi^-1 (:343-385) */
                const double chix = dchi[0][p], chiy = dchi[1][p], chiz = dchi[2][p];
                fx = dK[0][p]; fy = dK[1][p]; fz = dK[2][p];
                double Axxx = dA[0][0][p], Axxy = dA[0][1][p], Axxz = dA[0][2][p];
                double Axyx = dA[1][0][p], Axyy = dA[1][1][p], Axyz = dA[1][2][p];
                double Axzx = dA[2][0][p], Axzy = dA[2][1][p], Axzz = dA[2][2][p];
                double Ayyx = dA[3][0][p], Ayyy = dA[3][1][p], Ayyz = dA[3][2][p];
                double Ayzx = dA[4][0][p], Ayzy = dA[4][1][p], Ayzz = dA[4][2][p];
                double Azzx = dA[5][0][p], Azzy = dA[5][1][p], Azzz = dA[5][2][p];
                Axxx = Axxx - (Gamxxx * Axx + Gamyxx * Axy + Gamzxx * Axz) * TWO - chix / chipn1 * Axx + F1o3 * gxx * fx;
                Axxy = Axxy - (Gamxxy * Axx + Gamyxy * Axy + Gamzxy * Axz) * TWO - chiy / chipn1 * Axx + F1o3 * gxx * fy;
                Axxz = Axxz - (Gamxxz * Axx + Gamyxz * Axy + Gamzxz * Axz) * TWO - chiz / chipn1 * Axx + F1o3 * gxx * fz;
                Ayyx = Ayyx - (Gamxxy * Axy + Gamyxy * Ayy + Gamzxy * Ayz) * TWO - chix / chipn1 * Ayy + F1o3 * gyy * fx;
                Ayyy = Ayyy - (Gamxyy * Axy + Gamyyy * Ayy + Gamzyy * Ayz) * TWO - chiy / chipn1 * Ayy + F1o3 * gyy * fy;
                Ayyz = Ayyz - (Gamxyz * Axy + Gamyyz * Ayy + Gamzyz * Ayz) * TWO - chiz / chipn1 * Ayy + F1o3 * gyy * fz;
                Azzx = Azzx - (Gamxxz * Axz + Gamyxz * Ayz + Gamzxz * Azz) * TWO - chix / chipn1 * Azz + F1o3 * gzz * fx;
                Azzy = Azzy - (Gamxyz * Axz + Gamyyz * Ayz + Gamzyz * Azz) * TWO - chiy / chipn1 * Azz + F1o3 * gzz * fy;
                Azzz = Azzz - (Gamxzz * Axz + Gamyzz * Ayz + Gamzzz * Azz) * TWO - chiz / chipn1 * Azz + F1o3 * gzz * fz;
                Axyx = Axyx - (Gamxxy * Axx + Gamyxy * Axy + Gamzxy * Axz + Gamxxx * Axy + Gamyxx * Ayy + Gamzxx * Ayz) -
                       chix / chipn1 * Axy + F1o3 * gxy * fx;
                Axyy = Axyy - (Gamxyy * Axx + Gamyyy * Axy + Gamzyy * Axz + Gamxxy * Axy + Gamyxy * Ayy + Gamzxy * Ayz) -
                       chiy / chipn1 * Axy + F1o3 * gxy * fy;
                Axyz = Axyz - (Gamxyz * Axx + Gamyyz * Axy + Gamzyz * Axz + Gamxxz * Axy + Gamyxz * Ayy + Gamzxz * Ayz) -
                       chiz / chipn1 * Axy + F1o3 * gxy * fz;
                Axzx = Axzx - (Gamxxz * Axx + Gamyxz * Axy + Gamzxz * Axz + Gamxxx * Axz + Gamyxx * Ayz + Gamzxx * Azz) -
                       chix / chipn1 * Axz + F1o3 * gxz * fx;
                Axzy = Axzy - (Gamxyz * Axx + Gamyyz * Axy + Gamzyz * Axz + Gamxxy * Axz + Gamyxy * Ayz + Gamzxy * Azz) -
                       chiy / chipn1 * Axz + F1o3 * gxz * fy;
                Axzz = Axzz - (Gamxzz * Axx + Gamyzz * Axy + Gamzzz * Axz + Gamxxz * Axz + Gamyxz * Ayz + Gamzxz * Azz) -
                       chiz / chipn1 * Axz + F1o3 * gxz * fz;
                Ayzx = Ayzx - (Gamxxz * Axy + Gamyxz * Ayy + Gamzxz * Ayz + Gamxxy * Axz + Gamyxy * Ayz + Gamzxy * Azz) -
                       chix / chipn1 * Ayz + F1o3 * gyz * fx;
                Ayzy = Ayzy - (Gamxyz * Axy + Gamyyz * Ayy + Gamzyz * Ayz + Gamxyy * Axz + Gamyyy * Ayz + Gamzyy * Azz) -
                       chiy / chipn1 * Ayz + F1o3 * gyz * fy;
                Ayzz = Ayzz - (Gamxzz * Axy + Gamyzz * Ayy + Gamzzz * Ayz + Gamxyz * Axz + Gamyyz * Ayz + Gamzyz * Azz) -
                       chiz / chipn1 * Ayz + F1o3 * gyz * fz;
                /* B_ij = v^k (D_k K_ij - D_j K_ik), symmetrised (:386-392) */
                const double Bxx = (vy * (Axxy - Axyx) + vz * (Axxz - Axzx)) * f;
                const double Byy = (vx * (Ayyx - Axyy) + vz * (Ayyz - Ayzy)) * f;
                const double Bzz = (vx * (Azzx - Axzz) + vy * (Azzy - Ayzz)) * f;
                const double Bxy = (vx * (Axyx - (Axxy + Axyx) / TWO) + vy * (Axyy - Ayyx) / TWO +
                                    vz * (Axyz - (Axzy + Ayzx) / TWO)) * f;
                const double Bxz = (vx * (Axzx - (Axxz + Axzx) / TWO) + vy * (Axzy - (Axyz + Ayzx) / TWO) +
                                    vz * (Axzz - Azzx) / TWO) * f;
                const double Byz = (vx * (Ayzx - (Axyz + Axzy) / TWO) + vy * (Ayzy - (Ayyz + Ayzy) / TWO) +
                                    vz * (Ayzz - Azzy) / TWO) * f;
                /* K_ij up to chi^-1, E_ij = R_ij - K_ik K^k_j + K K_ij - B_ij (:393-440) */
                Axxx = Axx + F1o3 * tK * gxx;
                Axyx = Axy + F1o3 * tK * gxy;
                Axzx = Axz + F1o3 * tK * gxz;
                Ayyx = Ayy + F1o3 * tK * gyy;
                Ayzx = Ayz + F1o3 * tK * gyz;
                Azzx = Azz + F1o3 * tK * gzz;
                double Exx = gupxx * Axxx * Axxx + gupyy * Axyx * Axyx + gupzz * Axzx * Axzx +
                             TWO * (gupxy * Axxx * Axyx + gupxz * Axxx * Axzx + gupyz * Axyx * Axzx);
                double Eyy = gupxx * Axyx * Axyx + gupyy * Ayyx * Ayyx + gupzz * Ayzx * Ayzx +
                             TWO * (gupxy * Axyx * Ayyx + gupxz * Axyx * Ayzx + gupyz * Ayyx * Ayzx);
                double Ezz = gupxx * Axzx * Axzx + gupyy * Ayzx * Ayzx + gupzz * Azzx * Azzx +
                             TWO * (gupxy * Axzx * Ayzx + gupxz * Axzx * Azzx + gupyz * Ayzx * Azzx);
                double Exy = gupxx * Axxx * Axyx + gupyy * Axyx * Ayyx + gupzz * Axzx * Ayzx +
                             gupxy * (Axxx * Ayyx + Axyx * Axyx) + gupxz * (Axxx * Ayzx + Axzx * Axyx) +
                             gupyz * (Axyx * Ayzx + Axzx * Ayyx);
                double Exz = gupxx * Axxx * Axzx + gupyy * Axyx * Ayzx + gupzz * Axzx * Azzx +
                             gupxy * (Axxx * Ayzx + Axyx * Axzx) + gupxz * (Axxx * Azzx + Axzx * Axzx) +
                             gupyz * (Axyx * Azzx + Axzx * Ayzx);
                double Eyz = gupxx * Axyx * Axzx + gupyy * Ayyx * Ayzx + gupzz * Ayzx * Azzx +
                             gupxy * (Axyx * Ayzx + Ayyx * Axzx) + gupxz * (Axyx * Azzx + Ayzx * Axzx) +
                             gupyz * (Ayyx * Azzx + Ayzx * Ayzx);
                Exx = R[0][p] - (Exx - Axxx * tK) * f - Bxx;
                Exy = R[1][p] - (Exy - Axyx * tK) * f - Bxy;
                Exz = R[2][p] - (Exz - Axzx * tK) * f - Bxz;
                Eyy = R[3][p] - (Eyy - Ayyx * tK) * f - Byy;
                Eyz = R[4][p] - (Eyz - Ayzx * tK) * f - Byz;
                Ezz = R[5][p] - (Ezz - Azzx * tK) * f - Bzz;
                /* m = (u - i w) / sqrt(2) (:441-452) */
                const double uuwwxx = ux * ux - wx * wx, uuwwxy = ux * uy - wx * wy, uuwwxz = ux * uz - wx * wz;
                const double uuwwyy = uy * uy - wy * wy, uuwwyz = uy * uz - wy * wz, uuwwzz = uz * uz - wz * wz;
                const double uwxx = ux * wx + wx * ux, uwxy = ux * wy + wx * uy, uwxz = ux * wz + wx * uz;
                const double uwyy = uy * wy + wy * uy, uwyz = uy * wz + wy * uz, uwzz = uz * wz + wz * uz;
                double rp = Exx * uuwwxx + Eyy * uuwwyy + Ezz * uuwwzz + (Exy * uuwwxy + Exz * uuwwxz + Eyz * uuwwyz) * TWO;
                double ip = Exx * uwxx + Eyy * uwyy + Ezz * uwzz + (Exy * uwxy + Exz * uwxz + Eyz * uwyz) * TWO;
                Rpsi4[p] = -rp / TWO;
                Ipsi4[p] = -ip / TWO;
            }
    free(d);
}

/* polint (fmisc.f90:1121-1179): Neville's tableau on xa(1..n), ya(1..n)
 * (1-based arrays of n + 1 entries), starting from the sample nearest x. */
static double neville(const double *xa, const double *ya, double x, int n)
{
    double c[8], d[8], den[8], ho[8];
    for (int i = 1; i <= n; i++) { c[i] = ya[i]; d[i] = ya[i]; ho[i] = xa[i] - x; }
    int ns = 1;
    double dif = fabs(x - xa[1]);
    for (int m = 1; m <= n; m++) {
        const double dift = fabs(x - xa[m]);
        if (dift < dif) { ns = m; dif = dift; }
    }
    double y = ya[ns], dy;
    ns = ns - 1;
    for (int m = 1; m <= n - 1; m++) {
        for (int i = 1; i <= n - m; i++) den[i] = ho[i] - ho[i + m];
        for (int i = 1; i <= n - m; i++) den[i] = (c[i + 1] - d[i]) / den[i];
        for (int i = 1; i <= n - m; i++) {
            d[i] = ho[i + m] * den[i];
            c[i] = ho[i] * den[i];
        }
        if (2 * ns < n - m) dy = c[ns + 1];
        else { dy = d[ns]; ns = ns - 1; }
        y = y + dy;
    }
    return y;
}

/* sommerfeld_rout.f90:141-217 at one face point (1-based i, j, k as in the
 * Fortran); returns 0, or 1 when decide3d rejects the block. */
static int sinterp_point(const int *ex, const double *X, const double *Y, const double *Z, double dX, double dY,
                         double dZ, double dT, const double *chi0, const double *Lap0, const double *f0, double *f,
                         const double *SoA, int symmetry, int i, int j, int k)
{
    enum { ORDN = 6 };
    const size_t p = IX(ex, i - 1, j - 1, k - 1);
    double r = (Lap0[p] + 1.0) * sqrt(1.0 + chi0[p]) * dT /
               sqrt(X[i - 1] * X[i - 1] + Y[j - 1] * Y[j - 1] + Z[k - 1] * Z[k - 1]);
    const double fac = 1.0 - r;
    double cx[3] = {r * X[i - 1] / dX, r * Y[j - 1] / dY, r * Z[k - 1] / dZ};
    const int idx[3] = {i, j, k};
    int cxB[3], cxT[3];
    for (int m = 0; m < 3; m++) {
        if (cx[m] > 0.0) cxB[m] = (int)((double)idx[m] - trunc(cx[m]) - ORDN / 2);
        else cxB[m] = (int)((double)idx[m] - trunc(cx[m]) - ORDN / 2 + 1);
    }
    for (int m = 0; m < 3; m++) {
        if (cx[m] > 0.0) cx[m] = trunc(cx[m]) - cx[m] + ORDN / 2;
        else cx[m] = trunc(cx[m]) - cx[m] + ORDN / 2 - 1;
        cxT[m] = cxB[m] + ORDN - 1;
    }
    if (symmetry == 0 && cxB[2] < 1) { cx[2] = cx[2] + (cxB[2] - 1); cxT[2] = cxT[2] - (cxB[2] - 1); cxB[2] = 1; }
    if (symmetry < 2 && cxB[1] < 1) { cx[1] = cx[1] + (cxB[1] - 1); cxT[1] = cxT[1] - (cxB[1] - 1); cxB[1] = 1; }
    if (symmetry < 2 && cxB[0] < 1) { cx[0] = cx[0] + (cxB[0] - 1); cxT[0] = cxT[0] - (cxB[0] - 1); cxB[0] = 1; }
    for (int m = 0; m < 3; m++)
        if (cxT[m] > ex[m]) {
            cx[m] = cx[m] + (cxT[m] - ex[m]);
            cxB[m] = cxB[m] - (cxT[m] - ex[m]);
            cxT[m] = ex[m];
        }
    /* decide3d: indices >= 1 read f0, indices <= 0 read f0(1 - i) * SoA */
    for (int m = 0; m < 3; m++) {
        const int fmin1 = cxB[m] > 1 ? cxB[m] : 1, fmax1 = cxT[m];
        const int fmin2 = cxB[m], fmax2 = cxT[m] < 0 ? cxT[m] : 0;
        if (fmin1 <= fmax1 && (fmin1 < 1 || fmax1 > ex[m])) return 1;
        if (fmin2 <= fmax2 && (1 - fmax2 < 1 || 1 - fmin2 > ex[m])) return 1;
    }
    double ya[ORDN][ORDN][ORDN]; /* [x][y][z], local index = global - cxB */
    for (int c = 0; c < ORDN; c++)
        for (int b = 0; b < ORDN; b++)
            for (int a = 0; a < ORDN; a++) {
                const int g[3] = {cxB[0] + a, cxB[1] + b, cxB[2] + c};
                double s = 1.0;
                int q[3];
                for (int m = 0; m < 3; m++) {
                    if (g[m] >= 1) q[m] = g[m];
                    else { q[m] = 1 - g[m]; s *= SoA[m]; }
                }
                const double v = f0[IX(ex, q[0] - 1, q[1] - 1, q[2] - 1)];
                ya[a][b][c] = (s == 1.0) ? v : v * s;
            }
    /* polin3 (fmisc.f90:1221-1260): z, then y, then x */
    double xa[ORDN + 1], line[ORDN + 1], ym[ORDN + 1], yat[ORDN];
    for (int a = 1; a <= ORDN; a++) xa[a] = (double)(a - 1);
    for (int a = 0; a < ORDN; a++) {
        for (int b = 0; b < ORDN; b++) {
            for (int c = 0; c < ORDN; c++) line[c + 1] = ya[a][b][c];
            yat[b] = neville(xa, line, cx[2], ORDN);
        }
        for (int b = 0; b < ORDN; b++) line[b + 1] = yat[b];
        ym[a + 1] = neville(xa, line, cx[1], ORDN);
    }
    r = neville(xa, ym, cx[0], ORDN);
    f[p] = r * fac;
    return 0;
}

int orc_sommerfeld_interp(const int *ex, const double *X, const double *Y, const double *Z, const double *bbox,
                          double dT, const double *chi0, const double *Lap0, const double *f0, double *f,
                          const double *SoA, int symmetry)
{
    const double dX = X[1] - X[0], dY = Y[1] - Y[0], dZ = Z[1] - Z[0];
    /* the six faces' layers, 1-based inclusive (sommerfeld_rout.f90:58-119) */
    int lay[6][6], on[6] = {0, 0, 0, 0, 0, 0};
    const int n0 = ex[0], n1 = ex[1], n2 = ex[2];
    if (fabs(X[n0 - 1] - bbox[3]) < dX) { on[0] = 1; int t[6] = {n0, 1, 1, n0, n1, n2}; memcpy(lay[0], t, sizeof t); }
    if (fabs(Y[n1 - 1] - bbox[4]) < dY) { on[1] = 1; int t[6] = {1, n1, 1, n0, n1, n2}; memcpy(lay[1], t, sizeof t); }
    if (fabs(Z[n2 - 1] - bbox[5]) < dZ) { on[2] = 1; int t[6] = {1, 1, n2, n0, n1, n2}; memcpy(lay[2], t, sizeof t); }
    if (fabs(X[0] - bbox[0]) < dX && !(symmetry == 2 && fabs(bbox[0]) < dX / 2)) {
        on[3] = 1; int t[6] = {1, 1, 1, 1, n1, n2}; memcpy(lay[3], t, sizeof t); }
    if (fabs(Y[0] - bbox[1]) < dY && !(symmetry == 2 && fabs(bbox[1]) < dY / 2)) {
        on[4] = 1; int t[6] = {1, 1, 1, n0, 1, n2}; memcpy(lay[4], t, sizeof t); }
    if (fabs(Z[0] - bbox[2]) < dZ && !(symmetry > 0 && fabs(bbox[2]) < dZ / 2)) {
        on[5] = 1; int t[6] = {1, 1, 1, n0, n1, 1}; memcpy(lay[5], t, sizeof t); }
    int bad = 0;
    for (int gp = 0; gp < 6; gp++) {
        if (!on[gp]) continue;
        for (int k = lay[gp][2]; k <= lay[gp][5]; k++)
            for (int j = lay[gp][1]; j <= lay[gp][4]; j++)
                for (int i = lay[gp][0]; i <= lay[gp][3]; i++)
                    bad += sinterp_point(ex, X, Y, Z, dX, dY, dZ, dT, chi0, Lap0, f0, f, SoA, symmetry, i, j, k);
    }
    return bad;
}

/* global_interp, the Cell branch (fmisc.f90:712-794): the 6-point Lagrange
 * interpolant of f at (x1, y1, z1) through decide3d's block (:798-874,
 * mirrored with SoA below a symmetry plane) and polin3 (Neville along z,
 * then y, then x).  Returns 0, or 1 where decide3d rejects the block (the
 * reference stops there). */
int orc_global_interp(const int *ex, const double *X, const double *Y, const double *Z, const double *f,
                      double x1, double y1, double z1, const double *SoA, int symmetry, double *out)
{
    enum { ORDN = 6 };
    const double F04f = (double)0.4f; /* the Fortran literal 0.4 is single precision */
    const double dX = X[1] - X[0], dY = Y[1] - Y[0], dZ = Z[1] - Z[0];
    const double d[3] = {dX, dY, dZ}, p[3] = {x1, y1, z1};
    const double *C[3] = {X, Y, Z};
    int cxB[3], cxT[3], cmin[3] = {1, 1, 1}, cmax[3] = {ex[0], ex[1], ex[2]};
    for (int m = 0; m < 3; m++) {
        const int cxI = (int)((p[m] - C[m][0]) / d[m] + F04f) + 1; /* idint truncates */
        cxB[m] = cxI - ORDN / 2 + 1;
        cxT[m] = cxB[m] + ORDN - 1;
    }
    if (symmetry == 2 && fabs(X[0]) < dX) cmin[0] = -ORDN / 2 + 1;
    if (symmetry == 2 && fabs(Y[0]) < dY) cmin[1] = -ORDN / 2 + 1;
    if (symmetry != 0 && fabs(Z[0]) < dZ) cmin[2] = -ORDN / 2 + 1;
    double cx[3];
    for (int m = 0; m < 3; m++) {
        if (cxB[m] < cmin[m]) { cxB[m] = cmin[m]; cxT[m] = cxB[m] + ORDN - 1; }
        if (cxT[m] > cmax[m]) { cxT[m] = cmax[m]; cxB[m] = cxT[m] + 1 - ORDN; }
        if (cxB[m] > 0) cx[m] = (p[m] - C[m][cxB[m] - 1]) / d[m];
        else cx[m] = (p[m] + C[m][1 - cxB[m] - 1]) / d[m];
    }
    /* decide3d: indices >= 1 read f, indices <= 0 read f(1 - i) * SoA */
    for (int m = 0; m < 3; m++) {
        const int fmin1 = cxB[m] > 1 ? cxB[m] : 1, fmax1 = cxT[m];
        const int fmin2 = cxB[m], fmax2 = cxT[m] < 0 ? cxT[m] : 0;
        if (fmin1 <= fmax1 && (fmin1 < 1 || fmax1 > ex[m])) return 1;
        if (fmin2 <= fmax2 && (1 - fmax2 < 1 || 1 - fmin2 > ex[m])) return 1;
    }
    double ya[ORDN][ORDN][ORDN];
    for (int c = 0; c < ORDN; c++)
        for (int b = 0; b < ORDN; b++)
            for (int a = 0; a < ORDN; a++) {
                const int g[3] = {cxB[0] + a, cxB[1] + b, cxB[2] + c};
                double sg = 1.0;
                int q[3];
                for (int m = 0; m < 3; m++) {
                    if (g[m] >= 1) q[m] = g[m];
                    else { q[m] = 1 - g[m]; sg *= SoA[m]; }
                }
                const double v = f[IX(ex, q[0] - 1, q[1] - 1, q[2] - 1)];
                ya[a][b][c] = (sg == 1.0) ? v : v * sg;
            }
    double xa[ORDN + 1], line[ORDN + 1], ym[ORDN + 1], yat[ORDN];
    for (int a = 1; a <= ORDN; a++) xa[a] = (double)(a - 1);
    for (int a = 0; a < ORDN; a++) {
        for (int b = 0; b < ORDN; b++) {
            for (int c = 0; c < ORDN; c++) line[c + 1] = ya[a][b][c];
            yat[b] = neville(xa, line, cx[2], ORDN);
        }
        for (int b = 0; b < ORDN; b++) line[b + 1] = yat[b];
        ym[a + 1] = neville(xa, line, cx[1], ORDN);
    }
    *out = neville(xa, ym, cx[0], ORDN);
    return 0;
}

/* misc::fact and misc::Wigner_d_function (misc.C:271-309): d^l_{m s} of
 * cos(theta), the alternating sum over t. */
static double wfact(int n) { return n <= 0 ? 1.0 : n * wfact(n - 1); }
static double wigner_d(int l, int m, int s, double costheta)
{
    const int C1 = m - s > 0 ? m - s : 0, C2 = l + m < l - s ? l + m : l - s;
    double vv = 0.0;
    const double sinht = sqrt((1 - costheta) / 2.0), cosht = sqrt((1 + costheta) / 2.0);
    const double first = (C1 % 2 == 0) ? 1.0 : -1.0;
    for (int t = C1; t < C2 + 1; t += 2)
        vv = vv + first * pow(cosht, 2 * l + m - s - 2 * t) * pow(sinht, 2 * t + s - m) /
                      (wfact(l + m - t) * wfact(l - s - t) * wfact(t) * wfact(t + s - m));
    for (int t = C1 + 1; t < C2 + 1; t += 2)
        vv = vv - first * pow(cosht, 2 * l + m - s - 2 * t) * pow(sinht, 2 * t + s - m) /
                      (wfact(l + m - t) * wfact(l - s - t) * wfact(t) * wfact(t + s - m));
    return vv * sqrt(wfact(l + m) * wfact(l - m) * wfact(l + s) * wfact(l - s));
}

/* misc::gaulegf (misc.C): Gauss-Legendre nodes and weights on [x1, x2]. */
static void gaulegf(double x1, double x2, double *x, double *w, int n)
{
    const double eps = 1.2E-16, PI = acos(-1.0);
    const int m = (n + 1) / 2;
    const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 0; i < m; i++) {
        double z = cos(PI * ((double)i + 0.75) / ((double)n + 0.5)), z1, pp, p1, p2, p3;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 0; j < n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2 * (double)j + 1) * z * p2 - (double)j * p3) / ((double)j + 1);
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > eps);
        x[i] = xm - xl * z;
        x[n - 1 - i] = xm + xl * z;
        w[i] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - 1 - i] = w[i];
    }
}

/* surface_integral::surface_integral (surface_integral.C:39-181, GaussInt as
 * macrodef.h:16 sets it) and surf_Wave on one block (:197-377): Psi4 is
 * interpolated (global_interp, order 6) at the n_tot points of the sphere of
 * radius rex -- Gauss-Legendre in cos(theta) (2N nodes over [-1, 1], upper
 * hemisphere first), the midpoint rule in phi (4N points; a quarter of them
 * for Symmetry 2, half the theta nodes for Symmetry 1 and 2, the other
 * octants or the other hemisphere added through the parities) -- and
 * RP + i IP = rex dphi sum_n w_theta sqrt((2l+1)/4pi) d^l_{m,spinw}(theta)
 * Psi4 e^{-i m phi} for l = spinw..maxl, m = -l..l, in that order.  The
 * points must lie in the block's owned cells (all but 3 ghost cells at each
 * end of an axis; none at a lower end on a symmetry plane).
 * Returns 0; 2 for a point outside; 1 where decide3d rejects a block. */
int orc_surf_wave(const int *ex, const double *X, const double *Y, const double *Z, const double *rpsi,
                  const double *ipsi, const double *soa_r, const double *soa_i, int symmetry, double rex, int N,
                  int spinw, int maxl, double *RP, double *IP)
{
    const double PI = acos(-1.0);
    int N_phi = 4 * N, N_theta = 2 * N;
    double *arcos = malloc(sizeof(double) * 2 * N), *wt = malloc(sizeof(double) * 2 * N);
    gaulegf(-1.0, 1.0, arcos, wt, N_theta);
    for (int i = 0; i < (N_theta + 1) / 2; i++) { /* misc::inversearray */
        double r = arcos[i]; arcos[i] = arcos[N_theta - 1 - i]; arcos[N_theta - 1 - i] = r;
        r = wt[i]; wt[i] = wt[N_theta - 1 - i]; wt[N_theta - 1 - i] = r;
    }
    double dphi;
    if (symmetry == 2) { N_phi /= 4; N_theta /= 2; dphi = PI / (2.0 * N_phi); }
    else if (symmetry == 1) { N_theta /= 2; dphi = 2.0 * PI / N_phi; }
    else { dphi = 2.0 * PI / N_phi; }
    const int n_tot = N_theta * N_phi;
    double *sf = malloc(sizeof(double) * 2 * n_tot);
    int rc = 0;
    /* the block's owned cells: three ghost cells at each end, none below a
     * symmetry plane (the array starts there) */
    const double *Cs[3] = {X, Y, Z};
    double lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        const double h = Cs[a][1] - Cs[a][0];
        const int plane = (a == 2 ? symmetry > 0 : symmetry > 1) && fabs(Cs[a][0]) < h;
        lo[a] = Cs[a][plane ? 0 : 3] - h / 2;
        hi[a] = Cs[a][ex[a] - 4] + h / 2;
    }
    for (int i = 0, n = 0; i < N_theta && !rc; ++i) {
        const double ct = arcos[i], st = sqrt(1.0 - ct * ct);
        for (int j = 0; j < N_phi && !rc; ++j, ++n) {
            const double ph = (j + 0.5) * dphi;
            const double q[3] = {rex * (st * cos(ph)), rex * (st * sin(ph)), rex * ct};
            for (int a = 0; a < 3; a++)
                if (q[a] < lo[a] || q[a] > hi[a]) rc = 2;
            if (rc) break;
            rc = orc_global_interp(ex, X, Y, Z, rpsi, q[0], q[1], q[2], soa_r, symmetry, &sf[2 * n]);
            if (!rc) rc = orc_global_interp(ex, X, Y, Z, ipsi, q[0], q[1], q[2], soa_i, symmetry, &sf[2 * n + 1]);
        }
    }
    if (!rc) {
        int NN = 0;
        for (int l = spinw; l <= maxl; l++) NN += 2 * l + 1;
        for (int k = 0; k < NN; k++) RP[k] = IP[k] = 0.0;
        const int lpsy = symmetry == 0 ? 1 : (symmetry == 1 ? 2 : 8);
        for (int n = 0; n < n_tot; n++) {
            const int i = n / N_phi, j = n - i * N_phi;
            int k = 0;
            for (int pl = spinw; pl <= maxl; pl++)
                for (int pm = -pl; pm <= pl; pm++, k++)
                    for (int lp = 0; lp < lpsy; lp++) {
                        /* (theta or pi - theta) x (phi, 2 pi - phi, pi - phi, pi + phi), the
                         * parities of the reflected octant (surface_integral.C:265-326) */
                        const int flip_t = lp & 1;
                        double arg, sgn_sin = 1.0, sr = 1.0, si = 1.0;
                        switch (lp >> 1) {
                        case 0: arg = pm * (j + 0.5) * dphi; break;
                        case 1: arg = pm * (j + 0.5) * dphi; sgn_sin = -1.0; sr = soa_r[1]; si = soa_i[1]; break;
                        case 2: arg = pm * (PI - (j + 0.5) * dphi); sr = soa_r[0]; si = soa_i[0]; break;
                        default: arg = pm * (PI + (j + 0.5) * dphi); sr = soa_r[1] * soa_r[0]; si = soa_i[1] * soa_i[0]; break;
                        }
                        if (flip_t) { sr *= soa_r[2]; si *= soa_i[2]; }
                        const double costheta = flip_t ? -arcos[i] : arcos[i];
                        const double cm = cos(arg), sm = sgn_sin * sin(arg);
                        const double R = sr * sf[2 * n], I = si * sf[2 * n + 1];
                        const double thetap = sqrt((2 * pl + 1.0) / 4.0 / PI) * wigner_d(pl, pm, spinw, costheta);
                        RP[k] = RP[k] + thetap * (R * cm + I * sm) * wt[i];
                        IP[k] = IP[k] + thetap * (I * cm - R * sm) * wt[i];
                    }
        }
        for (int k = 0; k < NN; k++) { RP[k] = RP[k] * rex * dphi; IP[k] = IP[k] * rex * dphi; }
    }
    free(arcos);
    free(wt);
    free(sf);
    return rc;
}

/* One axis of sommerfeld_routbam's derivative (sommerfeld_rout.f90:383-436):
 * second-order differences taken from the side the outgoing wave arrives
 * from (w = v x_a / r > 0: backward), centred or reversed where the array
 * (lowest index lo, -1 below a symmetry plane) leaves no room.  The reference
 * leaves fx unset when w == 0, i.e. x_a == 0, where fx x_a vanishes anyway;
 * here it is 0. */
static double sbam_dir(const line_t *L, double w, int pos, int lo, int hi, double d2)
{
    if (w > 0.0) {
        if (pos - 2 >= lo) return d2 * (3.0 * lp(L, 0) - 4.0 * lp(L, -1) + lp(L, -2));
        if (pos - 1 >= lo) return d2 * (-lp(L, -1) + lp(L, 1));
        return d2 * (-lp(L, 2) + 4.0 * lp(L, 1) - 3.0 * lp(L, 0));
    } else if (w < 0.0) {
        if (pos + 2 <= hi) return d2 * (-lp(L, 2) + 4.0 * lp(L, 1) - 3.0 * lp(L, 0));
        if (pos + 1 <= hi) return d2 * (-lp(L, -1) + lp(L, 1));
        return d2 * (3.0 * lp(L, 0) - 4.0 * lp(L, -1) + lp(L, -2));
    }
    return 0.0;
}

void orc_sommerfeld_routbam(const int *ex, const double *X, const double *Y, const double *Z,
                            const double *bbox, double *f_rhs, const double *f0, double velocity,
                            const double *SoA, int symmetry)
{
    const double dX = X[1] - X[0], dY = Y[1] - Y[0], dZ = Z[1] - Z[0];
    const double d2[3] = {1.0 / 2.0 / dX, 1.0 / 2.0 / dY, 1.0 / 2.0 / dZ};
    /* lowest readable index: one mirrored layer below a symmetry plane
     * (sommerfeld_rout.f90:262-268, symmetry_bd(1, ...)) */
    const int lo[3] = {symmetry > 1 && fabs(X[0]) < dX ? -1 : 0, symmetry > 1 && fabs(Y[0]) < dY ? -1 : 0,
                       symmetry > 0 && fabs(Z[0]) < dZ ? -1 : 0};
    const int n0 = ex[0] - 1, n1 = ex[1] - 1, n2 = ex[2] - 1;
    const int hi[3] = {n0, n1, n2};
    /* the outermost layer of each physical face (:272-325) */
    int lay[6][6], on[6] = {0, 0, 0, 0, 0, 0};
    if (fabs(X[n0] - bbox[3]) < dX) { on[0] = 1; int t[6] = {n0, 0, 0, n0, n1, n2}; memcpy(lay[0], t, sizeof t); }
    if (fabs(Y[n1] - bbox[4]) < dY) { on[1] = 1; int t[6] = {0, n1, 0, n0, n1, n2}; memcpy(lay[1], t, sizeof t); }
    if (fabs(Z[n2] - bbox[5]) < dZ) { on[2] = 1; int t[6] = {0, 0, n2, n0, n1, n2}; memcpy(lay[2], t, sizeof t); }
    if (fabs(X[0] - bbox[0]) < dX && !(symmetry == 2 && fabs(bbox[0]) < dX / 2)) {
        on[3] = 1; int t[6] = {0, 0, 0, 0, n1, n2}; memcpy(lay[3], t, sizeof t); }
    if (fabs(Y[0] - bbox[1]) < dY && !(symmetry == 2 && fabs(bbox[1]) < dY / 2)) {
        on[4] = 1; int t[6] = {0, 0, 0, n0, 0, n2}; memcpy(lay[4], t, sizeof t); }
    if (fabs(Z[0] - bbox[2]) < dZ && !(symmetry > 0 && fabs(bbox[2]) < dZ / 2)) {
        on[5] = 1; int t[6] = {0, 0, 0, n0, n1, 0}; memcpy(lay[5], t, sizeof t); }
    for (int gp = 0; gp < 6; gp++) {
        if (!on[gp]) continue;
        for (int k = lay[gp][2]; k <= lay[gp][5]; k++)
            for (int j = lay[gp][1]; j <= lay[gp][4]; j++)
                for (int i = lay[gp][0]; i <= lay[gp][3]; i++) {
                    const double x = X[i], y = Y[j], z = Z[k];
                    const double R = sqrt(x * x + y * y + z * z);
                    const double w[3] = {velocity * x / R, velocity * y / R, velocity * z / R};
                    const int pos[3] = {i, j, k};
                    double fd[3];
                    line_t L = {f0, ex, SoA, 0.0, i, j, k, 0};
                    for (int a = 0; a < 3; a++) {
                        L.ax = a;
                        fd[a] = sbam_dir(&L, w[a], pos[a], lo[a], hi[a], d2[a]);
                    }
                    f_rhs[IX(ex, i, j, k)] =
                        -(velocity * (fd[0] * x + fd[1] * y + fd[2] * z + f0[IX(ex, i, j, k)])) / R;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Inter-level transfer of cell-centred grids, 2:1 (prolongrestrict_cell.f90,
 * the ghost_width 3 branch: prolong3 :1915-2164, restrict3 :2330-2468).
 * Index bookkeeping in the reference's 1-based cell numbering from a common
 * base point; the Fortran literal 0.4 is single precision. */
#define F04 ((double)0.4f)

typedef struct {
    int lo[3], hi[3]; /* target loop bounds (1-based array indices) */
    int lbf[3], lbc[3];
} xfer_t;

/* base, lbf, lbc and the target/source ranges of a transfer (:1948-1990,
 * 2361-2405).  ll/uu: bounds of the wanted region; tgt_fine: the target is
 * the fine array (prolong).  Returns 0, or 1 where the reference reports an
 * error (prolong3 prints and returns, restrict3 stops). */
static int xfer_setup(const double *llbc, const double *uubc, const int *extc, const double *llbf,
                      const double *uubf, const int *extf, const double *ll, const double *uu, int tgt_fine,
                      xfer_t *x)
{
    double CD[3], FD[3], base[3];
    for (int a = 0; a < 3; a++) {
        CD[a] = (uubc[a] - llbc[a]) / extc[a];
        FD[a] = (uubf[a] - llbf[a]) / extf[a];
        if (fabs(CD[a] - 2 * FD[a]) > 1.e-10) return 1;
    }
    for (int a = 0; a < 3; a++) {
        if (llbc[a] <= llbf[a]) {
            base[a] = llbc[a];
        } else {
            const int j = (int)((llbc[a] - llbf[a]) / FD[a] + F04);
            base[a] = (j / 2 * 2 == j) ? llbf[a] : llbf[a] - CD[a] / 2;
        }
    }
    for (int a = 0; a < 3; a++) {
        x->lbf[a] = (int)((llbf[a] - base[a]) / FD[a] + F04) + 1;
        x->lbc[a] = (int)((llbc[a] - base[a]) / CD[a] + F04) + 1;
        const int lt = (int)((ll[a] - base[a]) / (tgt_fine ? FD[a] : CD[a]) + F04) + 1;
        const int ut = (int)((uu[a] - base[a]) / (tgt_fine ? FD[a] : CD[a]) + F04);
        const int ls = (int)((ll[a] - base[a]) / (tgt_fine ? CD[a] : FD[a]) + F04) + 1;
        const int us = (int)((uu[a] - base[a]) / (tgt_fine ? CD[a] : FD[a]) + F04);
        const int tl = tgt_fine ? x->lbf[a] : x->lbc[a], sl = tgt_fine ? x->lbc[a] : x->lbf[a];
        const int next = tgt_fine ? extf[a] : extc[a], nsrc = tgt_fine ? extc[a] : extf[a];
        x->lo[a] = lt - tl + 1;
        x->hi[a] = ut - tl + 1;
        const int smin = ls - sl + 1, smax = us - sl + 1;
        if (x->lo[a] < 1 || smin < 1 || x->hi[a] > next || smax > nsrc - 2) return 1;
    }
    return 0;
}

/* the source value at 1-based index (i, j, k), with symmetry_bd's mirror
 * below index 1 (fmisc.f90:877-901) */
static inline double xf_at(const double *f, const int *ex, int i, int j, int k, const double *S)
{
    return fh(f, ex, i - 1, j - 1, k - 1, S, 0.0);
}

int orc_prolong3(int wei, const double *llbc, const double *uubc, const int *extc, const double *func,
                 const double *llbf, const double *uubf, const int *extf, double *funf, const double *llbp,
                 const double *uubp, const double *SoA, int symmetry)
{
    (void)symmetry;
    if (wei != 3) return 2;
    xfer_t x;
    if (xfer_setup(llbc, uubc, extc, llbf, uubf, extf, llbp, uubp, 1, &x)) return 1;
    /* the stencils must stay inside funcc(-2:extc); the reference only prints
     * "error in prolong" (:2038) and reads on */
    for (int a = 0; a < 3; a++) {
        const int c0 = (x.lo[a] + x.lbf[a] - 1) / 2 - x.lbc[a] + 1, c1 = (x.hi[a] + x.lbf[a] - 1) / 2 - x.lbc[a] + 1;
        if (c0 < 0 || c1 + 3 > extc[a]) return 1;
    }
    static const double C1 = 7.7e1 / 8.192e3, C2 = -6.93e2 / 8.192e3, C3 = 3.465e3 / 4.096e3;
    static const double C6 = 6.3e1 / 8.192e3, C5 = -4.95e2 / 8.192e3, C4 = 1.155e3 / 4.096e3;
    const double Wf[6] = {C1, C2, C3, C4, C5, C6}, Wr[6] = {C6, C5, C4, C3, C2, C1};
    for (int k = x.lo[2]; k <= x.hi[2]; k++)
        for (int j = x.lo[1]; j <= x.hi[1]; j++)
            for (int i = x.lo[0]; i <= x.hi[0]; i++) {
                const int t[3] = {i, j, k};
                int c[3];
                const double *w[3];
                for (int a = 0; a < 3; a++) {
                    c[a] = (t[a] + x.lbf[a] - 1) / 2 - x.lbc[a] + 1; /* :2033-2036 */
                    const int g = t[a] + x.lbf[a] - 1;
                    w[a] = (g / 2 * 2 == g) ? Wf : Wr;
                }
                double tmp2[6][6], tmp1[6];
                for (int b = 0; b < 6; b++)
                    for (int a = 0; a < 6; a++) {
                        double s = w[2][0] * xf_at(func, extc, c[0] - 2 + a, c[1] - 2 + b, c[2] - 2, SoA);
                        for (int q = 1; q < 6; q++)
                            s = s + w[2][q] * xf_at(func, extc, c[0] - 2 + a, c[1] - 2 + b, c[2] - 2 + q, SoA);
                        tmp2[a][b] = s;
                    }
                for (int a = 0; a < 6; a++) {
                    double s = w[1][0] * tmp2[a][0];
                    for (int b = 1; b < 6; b++) s = s + w[1][b] * tmp2[a][b];
                    tmp1[a] = s;
                }
                double s = w[0][0] * tmp1[0];
                for (int a = 1; a < 6; a++) s = s + w[0][a] * tmp1[a];
                funf[IX(extf, i - 1, j - 1, k - 1)] = s;
            }
    return 0;
}

int orc_restrict3(int wei, const double *llbc, const double *uubc, const int *extc, double *func,
                  const double *llbf, const double *uubf, const int *extf, const double *funf,
                  const double *llbr, const double *uubr, const double *SoA, int symmetry)
{
    (void)symmetry;
    if (wei != 3) return 2;
    xfer_t x;
    if (xfer_setup(llbc, uubc, extc, llbf, uubf, extf, llbr, uubr, 0, &x)) return 1;
    /* the stencils must stay inside funff(-1:extf) (:2455) */
    for (int a = 0; a < 3; a++) {
        const int c0 = 2 * (x.lo[a] + x.lbc[a] - 1) - x.lbf[a], c1 = 2 * (x.hi[a] + x.lbc[a] - 1) - x.lbf[a];
        if (c0 < 1 || c1 + 3 > extf[a]) return 1;
    }
    static const double C1 = 3.e0 / 2.56e2, C2 = -2.5e1 / 2.56e2, C3 = 7.5e1 / 1.28e2;
    for (int k = x.lo[2]; k <= x.hi[2]; k++)
        for (int j = x.lo[1]; j <= x.hi[1]; j++)
            for (int i = x.lo[0]; i <= x.hi[0]; i++) {
                const int t[3] = {i, j, k};
                int c[3];
                for (int a = 0; a < 3; a++) c[a] = 2 * (t[a] + x.lbc[a] - 1) - 1 - x.lbf[a] + 1; /* :2450-2453 */
                double tmp2[6][6], tmp1[6];
#define G(a, b, q) xf_at(funf, extf, c[0] - 2 + (a), c[1] - 2 + (b), c[2] - 2 + (q), SoA)
                for (int b = 0; b < 6; b++)
                    for (int a = 0; a < 6; a++)
                        tmp2[a][b] = C1 * (G(a, b, 0) + G(a, b, 5)) + C2 * (G(a, b, 1) + G(a, b, 4)) +
                                     C3 * (G(a, b, 2) + G(a, b, 3));
#undef G
                for (int a = 0; a < 6; a++)
                    tmp1[a] = C1 * (tmp2[a][0] + tmp2[a][5]) + C2 * (tmp2[a][1] + tmp2[a][4]) +
                              C3 * (tmp2[a][2] + tmp2[a][3]);
                func[IX(extc, i - 1, j - 1, k - 1)] =
                    C1 * (tmp1[0] + tmp1[5]) + C2 * (tmp1[1] + tmp1[4]) + C3 * (tmp1[2] + tmp1[3]);
            }
    return 0;
}

/* Propagation speed of each field in the Sommerfeld condition: sqrt(2) for
 * chi, trK and the lapse (the 1+log gauge speed), 1 for the rest
 * (bssn_class.C:390-410, var.C:30). */
static double propspeed(int v) { return (v == 0 || v == 1 || v == 17) ? sqrt(2.0) : 1.0; }

/* Frozen-ghost boundary: every point within gw of an array face takes f0. */
static void ghost_copy(const int *ex, int gw, const double *f0, double *f)
{
    for (int k = 0; k < ex[2]; k++)
        for (int j = 0; j < ex[1]; j++)
            for (int i = 0; i < ex[0]; i++) {
                if (i < gw || j < gw || k < gw || i >= ex[0] - gw || j >= ex[1] - gw || k >= ex[2] - gw) {
                    const size_t p = IX(ex, i, j, k);
                    f[p] = f0[p];
                }
            }
}

int orc_bssn_step(const int *ex, const double *X, const double *Y, const double *Z, const double *bbox,
                  double *const *U, double dT, double eps, double chitiny, int symmetry, int lev, int gauge,
                  int bc, int gw, double *ham)
{
    const size_t N = (size_t)ex[0] * ex[1] * ex[2];
    double *mem = (double *)calloc(N * (3 * NF + 18 + 6 + 7), sizeof(double));
    if (!mem) return 2;
    double *pre[NF], *cor[NF], *acc[NF], *gam[18], *ric[6], *res[7];
    double *q = mem;
    for (int v = 0; v < NF; v++) { pre[v] = q; q += N; }
    for (int v = 0; v < NF; v++) { cor[v] = q; q += N; }
    for (int v = 0; v < NF; v++) { acc[v] = q; q += N; }
    for (int a = 0; a < 18; a++) { gam[a] = q; q += N; }
    for (int a = 0; a < 6; a++) { ric[a] = q; q += N; }
    for (int a = 0; a < 7; a++) { res[a] = q; q += N; }
    const double TINNY = chitiny - 1.0; /* bssn_class.C:179 */
    int err = 0;
    /* predictor: bssn_class.C:3042-3163 */
    orc_enforce_ga(ex, U + DXX, U + AXX);
    err |= orc_compute_rhs_bssn(ex, 0.0, X, Y, Z, (const double *const *)U, acc, NULL, gam, ric, res, symmetry,
                                lev, eps, 0, gauge);
    for (int v = 0; v < NF; v++) {
        /* level-0 BAM Sommerfeld on the RHS before the update (bssn_class.C:3106-3117) */
        if (bc == 3) orc_sommerfeld_routbam(ex, X, Y, Z, bbox, acc[v], U[v], propspeed(v), PAR[v], symmetry);
        orc_rungekutta4_rout(ex, dT, U[v], pre[v], acc[v], 0);
        if (bc == 0) ghost_copy(ex, gw, U[v], pre[v]);
        else if (bc == 1) orc_sommerfeld_copy(ex, X, Y, Z, bbox, U[v], pre[v], symmetry);
    }
    orc_lowerboundset(ex, pre[CHI], TINNY);
    if (ham) memcpy(ham, res[0], N * sizeof(double));
    /* correctors: bssn_class.C:3401-3781 */
    for (int it = 1; it < 4; it++) {
        orc_enforce_ga(ex, pre + DXX, pre + AXX);
        err |= orc_compute_rhs_bssn(ex, 0.0, X, Y, Z, (const double *const *)pre, cor, NULL, gam, ric, res,
                                    symmetry, lev, eps, 1, gauge);
        for (int v = 0; v < NF; v++) {
            /* bssn_class.C:3476-3484: the stage's k, from the stage input */
            if (bc == 3) orc_sommerfeld_routbam(ex, X, Y, Z, bbox, cor[v], pre[v], propspeed(v), PAR[v], symmetry);
            orc_rungekutta4_rout(ex, dT, U[v], cor[v], acc[v], it);
            if (bc == 0) ghost_copy(ex, gw, U[v], cor[v]);
            else if (bc == 1) orc_sommerfeld_copy(ex, X, Y, Z, bbox, U[v], cor[v], symmetry);
        }
        orc_lowerboundset(ex, cor[CHI], TINNY);
        if (it < 3)
            for (int v = 0; v < NF; v++) { double *t = pre[v]; pre[v] = cor[v]; cor[v] = t; }
    }
    /* rotation: the corrector output becomes the state (bssn_class.C:3845-3859) */
    for (int v = 0; v < NF; v++) memcpy(U[v], cor[v], N * sizeof(double));
    free(mem);
    return err;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11). */
static inline void mulhilo32(unsigned a, unsigned b, unsigned *hi, unsigned *lo)
{
    unsigned long long p = (unsigned long long)a * b;
    *hi = (unsigned)(p >> 32);
    *lo = (unsigned)p;
}

void orc_philox4x32_10(const unsigned *ctr, const unsigned *key, unsigned *out)
{
    unsigned c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        unsigned hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
