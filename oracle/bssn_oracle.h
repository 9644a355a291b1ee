/*
 * bssn_oracle.h -- CPU restatement of the AMSS-NCKU BSSN RHS + RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links this library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the checker (or as the timed CPU baseline).
 *
 * Conventions: arrays are contiguous FP64 in Fortran order (i fastest),
 * ex[0]*ex[1]*ex[2] points, indices 0-based.  Field order of the 24 evolved
 * fields follows the reference registry (bssn_class.C:437-460):
 *   0 chi 1 trK 2 dxx 3 gxy 4 gxz 5 dyy 6 gyz 7 dzz 8 Axx 9 Axy 10 Axz
 *   11 Ayy 12 Ayz 13 Azz 14 Gamx 15 Gamy 16 Gamz 17 Lap 18 betax 19 betay
 *   20 betaz 21 dtSfx 22 dtSfy 23 dtSfz
 * chi, Lap and the diagonal metric are stored as offsets (value - 1).
 */
#ifndef BSSN_ORACLE_H
#define BSSN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* diff_new.f90:946-1104 (ghost_width 3 branch) */
void orc_fderivs(const int *ex, const double *f, double *fx, double *fy, double *fz,
                 const double *X, const double *Y, const double *Z,
                 double S1, double S2, double S3, int symmetry);
/* diff_new.f90:1338-1528 */
void orc_fdderivs(const int *ex, const double *f, double *fxx, double *fxy, double *fxz,
                  double *fyy, double *fyz, double *fzz,
                  const double *X, const double *Y, const double *Z,
                  double S1, double S2, double S3, int symmetry);
/* lopsidediff.f90:186-488 (active branch 321-481) */
void orc_lopsided(const int *ex, const double *X, const double *Y, const double *Z,
                  const double *f, double *f_rhs, const double *Sfx, const double *Sfy,
                  const double *Sfz, int symmetry, const double *SoA);
/* kodiss.f90:120-216 */
void orc_kodis(const int *ex, const double *X, const double *Y, const double *Z,
               const double *f, double *f_rhs, const double *SoA, int symmetry, double eps);
/* fmisc.f90:877-901 ; fh has (ex+ord)^3 points, offset ord per axis */
void orc_symmetry_bd(int ord, const int *ex, const double *f, double *fh, const double *SoA);

/* bssn_rhs.f90:5-1186.  st/rhs: 24 arrays; mat: 10 arrays or NULL (vacuum);
 * gam: 18, ric: 6, res: 7 (ham, movx, movy, movz, Gmx, Gmy, Gmz).
 * gauge: 0 (B^i driver, eta=2) or 2 (the reference build's configured gauge).
 * Returns 1 when the input sum is NaN (nothing written), else 0. */
int orc_compute_rhs_bssn(const int *ex, double T, const double *X, const double *Y,
                         const double *Z, const double *const *st, double *const *rhs,
                         const double *const *mat, double *const *gam, double *const *ric,
                         double *const *res, int symmetry, int lev, double eps, int co,
                         int gauge);

/* enforce_algebra.f90:72-127; g: dxx gxy gxz dyy gyz dzz, a: Axx..Azz */
void orc_enforce_ga(const int *ex, double *const *g, double *const *a);
/* rungekutta4_rout.f90:94-139; returns 1 on a bad stage index */
int orc_rungekutta4_rout(const int *ex, double dT, const double *f0, double *f1,
                         double *f_rhs, int RK4);
/* fmisc.f90:1644-1657 */
void orc_lowerboundset(const int *ex, double *chi, double tinny);
/* fmisc.f90:1263-1309 */
double orc_l2normhelper(const int *ex, const double *X, const double *Y, const double *Z,
                        double xmin, double ymin, double zmin, double xmax, double ymax,
                        double zmax, const double *f, int gw);
/* sommerfeld_rout.f90:9-221, precor == 1 (copy) branch */
void orc_sommerfeld_copy(const int *ex, const double *X, const double *Y, const double *Z,
                         const double *bbox, const double *f0, double *f, int symmetry);
/* getnp4.f90:10-452 (Cartesian, tetradtype 2 and ABV 0 of the reference's
 * macrodef.fh): the Newman-Penrose Weyl scalar Psi4 = -(E - B)_ij m^i m^j
 * with the radial tetrad, from the state, the physical Christoffels (18, in
 * compute_rhs_bssn's order) and Ricci (6) of the same state. */
void orc_getnp4(const int *ex, const double *X, const double *Y, const double *Z, const double *chi,
                const double *trK, const double *const *g, const double *const *A, const double *const *Gam,
                const double *const *R, double *Rpsi4, double *Ipsi4, int symmetry);
/* sommerfeld_rout.f90:9-121, 141-217: precor == 0 (the interpolating
 * predictor): on the outermost layer of each physical face of bbox,
 * f = (1 - r) f0(x - r x), r = alpha sqrt(chi) dT / |x| (alpha, chi from
 * Lap0, chi0), f0 interpolated by 6-point Neville polynomials (polin3,
 * fmisc.f90:1121-1260) over a 6^3 block (decide3d, fmisc.f90:798-874:
 * mirrored with SoA below index 1).  Returns the number of points whose
 * block decide3d rejects (left unchanged; the reference prints an error and
 * interpolates undefined data there). */
/* fmisc.f90:712-794 (global_interp, Cell): 6-point interpolation of f at
 * (x1, y1, z1); returns 0, or 1 where decide3d rejects the block. */
int orc_global_interp(const int *ex, const double *X, const double *Y, const double *Z, const double *f,
                      double x1, double y1, double z1, const double *SoA, int symmetry, double *out);
/* surface_integral.C:39-181 + 197-377 (GaussInt): the spin-weighted modes of
 * Psi4 on the sphere of radius rex, N points per quarter sphere; RP, IP hold
 * sum_{l=spinw..maxl} (2l+1) values.  Returns 0, 2 for a sphere point outside
 * the block's owned cells, 1 where decide3d rejects a block. */
int orc_surf_wave(const int *ex, const double *X, const double *Y, const double *Z, const double *rpsi,
                  const double *ipsi, const double *soa_r, const double *soa_i, int symmetry, double rex, int N,
                  int spinw, int maxl, double *RP, double *IP);
int orc_sommerfeld_interp(const int *ex, const double *X, const double *Y, const double *Z, const double *bbox,
                          double dT, const double *chi0, const double *Lap0, const double *f0, double *f,
                          const double *SoA, int symmetry);
/* sommerfeld_rout.f90:223-449 (sommerfeld_routbam): on the outermost layer
 * of each physical face of bbox, f_rhs = -velocity (x.grad f0 + f0) / r. */
void orc_sommerfeld_routbam(const int *ex, const double *X, const double *Y, const double *Z,
                            const double *bbox, double *f_rhs, const double *f0, double velocity,
                            const double *SoA, int symmetry);

/* prolongrestrict_cell.f90 (Cell, ghost_width 3): 6-point 5th-order
 * prolongation onto the fine region [llbp, uubp] and restriction onto the
 * coarse region [llbr, uubr].  Return 0, 1 where the reference reports a bad
 * region (and writes nothing), 2 for wei != 3. */
int orc_prolong3(int wei, const double *llbc, const double *uubc, const int *extc, const double *func,
                 const double *llbf, const double *uubf, const int *extf, double *funf, const double *llbp,
                 const double *uubp, const double *SoA, int symmetry);
int orc_restrict3(int wei, const double *llbc, const double *uubc, const int *extc, double *func,
                  const double *llbf, const double *uubf, const int *extf, const double *funf,
                  const double *llbr, const double *uubr, const double *SoA, int symmetry);

/* One RK4 step of one Block in the order of bssn_class::Step
 * (bssn_class.C:2981-3901) for a single rank and a single patch.
 * U: 24 arrays, overwritten with the new state.  bc: 0 = copy the U0 value
 * onto every point within gw of the array faces after each sub-step
 * (frozen ghost zones), 1 = the reference's one-layer precor copy
 * (sommerfeld_rout with cor) against bbox, 2 = none, 3 = the level-0 BAM
 * Sommerfeld condition on each stage's RHS (sommerfeld_routbam, with each
 * field's propagation speed and parity).  ham (optional): the
 * Hamiltonian residual written by the predictor's co=0 RHS call.
 * Returns the OR of the RHS NaN flags. */
int orc_bssn_step(const int *ex, const double *X, const double *Y, const double *Z,
                  const double *bbox, double *const *U, double dT, double eps,
                  double chitiny, int symmetry, int lev, int gauge, int bc, int gw,
                  double *ham);

/* Philox4x32-10 (Salmon et al., SC'11): out[4] from ctr[4], key[2]. */
void orc_philox4x32_10(const unsigned *ctr, const unsigned *key, unsigned *out);

#ifdef __cplusplus
}
#endif
#endif
