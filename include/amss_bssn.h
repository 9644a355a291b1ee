/*
 * amss_bssn.h -- C ABI of the MI355X (gfx950) BSSN RHS + RK4 library
 * (libamssbssn.so).
 *
 * Two levels of entry points:
 *
 * 1. Link-compatible replacements of the reference's Fortran entry points
 *    (trailing-underscore mangling of -Dfortran3, every argument by
 *    reference, host arrays in Fortran order, i fastest).  They stage the
 *    host arrays through the GPU and return the same results.  Each cites
 *    the reference declaration it replaces (paths relative to
 *    amss-ncku-python/AMSS_NCKU_source/).
 *
 * 2. A device-resident API: device pointers, stream-ordered, no hidden
 *    host synchronisation except where documented.  Blocks, patches and the
 *    Step driver live behind an opaque level handle.
 *
 * Return codes: 0 ok, 1 non-finite input (the reference's gont = 1),
 * 2 invalid argument, 3 HIP runtime error, 4 unsupported configuration.
 * amss_last_error() describes the last non-zero code of the calling thread.
 * No entry point aborts the process or throws.
 */
#ifndef AMSS_BSSN_H
#define AMSS_BSSN_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define AMSS_OK 0
#define AMSS_ERR_NONFINITE 1
#define AMSS_ERR_ARG 2
#define AMSS_ERR_HIP 3
#define AMSS_ERR_UNSUPPORTED 4

/* Library version (major*10000 + minor*100 + patch). */
int amss_version(void);
/* Text of the calling thread's last error (empty when none). */
const char *amss_last_error(void);
/* Gauge used by the Fortran-ABI replacements: 0 (B^i driver, eta = 2;
 * bssn_rhs.f90:792-799, the default) or 2 (bssn_rhs.f90:809-820).  The
 * reference fixes it at compile time (macrodef.fh GAUGE). */
int amss_set_gauge(int gauge);

/* ---------------------------------------------------------------------------
 * 1. Fortran-ABI replacements (host pointers)
 * ------------------------------------------------------------------------- */

/* Replaces f_compute_rhs_bssn (bssn_rhs.h:37-54; bssn_rhs.f90:5-1186).
 * 98 arguments: ex, T, X, Y, Z, the 24 state arrays, the 24 RHS arrays,
 * rho, Sx..Sz, Sxx..Szz, the 18 physical Christoffels, the 6 physical Ricci
 * components, ham/mom x3/Gam x3 residuals (written only when co == 0),
 * Symmetry, Lev, eps, co.  Returns 1 (nothing written) when the sum of the
 * first 21 state arrays is NaN.  Symmetry 1 (equatorial: z) and 2 (octant:
 * x, y, z) mirror the stencils across a lower face whose first coordinate is
 * within one spacing of 0 (diff_new.f90:975-980), with each field's parity;
 * other values return 2, and so does an axis of fewer than 5 points without a
 * symmetry plane (the upwind stencil would read past it). */
int compute_rhs_bssn_(int *ex, double *T, double *X, double *Y, double *Z,
                      double *chi, double *trK, double *dxx, double *gxy, double *gxz, double *dyy,
                      double *gyz, double *dzz, double *Axx, double *Axy, double *Axz, double *Ayy,
                      double *Ayz, double *Azz, double *Gamx, double *Gamy, double *Gamz, double *Lap,
                      double *betax, double *betay, double *betaz, double *dtSfx, double *dtSfy,
                      double *dtSfz, double *chi_rhs, double *trK_rhs, double *gxx_rhs, double *gxy_rhs,
                      double *gxz_rhs, double *gyy_rhs, double *gyz_rhs, double *gzz_rhs, double *Axx_rhs,
                      double *Axy_rhs, double *Axz_rhs, double *Ayy_rhs, double *Ayz_rhs, double *Azz_rhs,
                      double *Gamx_rhs, double *Gamy_rhs, double *Gamz_rhs, double *Lap_rhs,
                      double *betax_rhs, double *betay_rhs, double *betaz_rhs, double *dtSfx_rhs,
                      double *dtSfy_rhs, double *dtSfz_rhs, double *rho, double *Sx, double *Sy, double *Sz,
                      double *Sxx, double *Sxy, double *Sxz, double *Syy, double *Syz, double *Szz,
                      double *Gamxxx, double *Gamxxy, double *Gamxxz, double *Gamxyy, double *Gamxyz,
                      double *Gamxzz, double *Gamyxx, double *Gamyxy, double *Gamyxz, double *Gamyyy,
                      double *Gamyyz, double *Gamyzz, double *Gamzxx, double *Gamzxy, double *Gamzxz,
                      double *Gamzyy, double *Gamzyz, double *Gamzzz, double *Rxx, double *Rxy,
                      double *Rxz, double *Ryy, double *Ryz, double *Rzz, double *ham_Res,
                      double *movx_Res, double *movy_Res, double *movz_Res, double *Gmx_Res,
                      double *Gmy_Res, double *Gmz_Res, int *Symmetry, int *Lev, double *eps, int *co);

/* Replaces f_rungekutta4_rout (rungekutta4_rout.h:34-36; rungekutta4_rout.f90:94-139).
 * A stage outside 0..3 returns 2 instead of stopping the program. */
int rungekutta4_rout_(int *ex, double *dT, double *f0, double *f1, double *f_rhs, int *RK4);

/* Replaces f_enforce_ga (enforce_algebra.h:26-28; enforce_algebra.f90:72-127). */
void enforce_ga_(int *ex, double *dxx, double *gxy, double *gxz, double *dyy, double *gyz, double *dzz,
                 double *Axx, double *Axy, double *Axz, double *Ayy, double *Ayz, double *Azz);

/* Replaces f_lowerboundset (fmisc.h:210; fmisc.f90:1644-1657). */
void lowerboundset_(int *ex, double *chi, double *tinny);

/* Replaces f_l2normhelper (fmisc.h:161-164; fmisc.f90:1263-1309). */
void l2normhelper_(int *ex, double *X, double *Y, double *Z, double *xmin, double *ymin, double *zmin,
                   double *xmax, double *ymax, double *zmax, double *f, double *f_out, int *gw);

/* Replaces f_sommerfeld_rout (sommerfeld_rout.h:26-29; sommerfeld_rout.f90:9-221).
 * precor == 1: copy f0 onto the outermost layer of every physical face.
 * Any other precor (the interpolating predictor of SommerType 1): each point
 * of those layers becomes the outgoing wave (1 - r) f0(x - r x) with
 * r = (Lap0 + 1) sqrt(1 + chi0) dT / |x|, f0 interpolated by 6-point
 * polynomials along z, y, x over a 6^3 block kept inside the array (mirrored
 * with SoA below a symmetry plane).  Where the array is too small for a
 * block (the reference's decide3d error) those points are left unchanged.
 * Errors (that one, Symmetry outside 0..2, a HIP failure) are written to
 * stderr and to amss_last_error(). */
void sommerfeld_rout_(int *ex, double *X, double *Y, double *Z, double *xmin, double *ymin, double *zmin,
                      double *xmax, double *ymax, double *zmax, double *dT, double *chi0, double *Lap0,
                      double *f0, double *f, double *SoA, int *Symmetry, int *precor);

/* Replaces f_sommerfeld_routbam (sommerfeld_rout.h:34-36; sommerfeld_rout.f90:223-449):
 * on the outermost layer of every physical face of the bbox, f_rhs becomes
 * -velocity (x.grad f0 + f0) / r (second-order one-sided differences from
 * the side the outgoing wave arrives from); SoA: f0's parity per axis, used
 * for the mirrored layer below a symmetry plane.  Shapes below 3 or Symmetry
 * outside 0..2 leave f_rhs untouched and set amss_last_error(). */
void sommerfeld_routbam_(int *ex, double *X, double *Y, double *Z, double *xmin, double *ymin, double *zmin,
                         double *xmax, double *ymax, double *zmax, double *f_rhs, double *f0, double *velocity,
                         double *SoA, int *Symmetry);

/* Replaces f_getnp4 (getnp4.h:97-105; getnp4.f90:10-452 with tetradtype 2 and
 * ABV 0, called at bssn_class.C:6247): the Newman-Penrose Weyl scalar Psi4 =
 * Rpsi4 + i Ipsi4 of the BSSN state in the radial tetrad, from the physical
 * Christoffels and Ricci tensor compute_rhs_bssn_ returned for that state
 * (read only).  The first derivatives of A_ij, chi and trK are fderivs'
 * (diff_new.f90:946-1104): fourth order with +-2 room on every axis, second
 * order with +-1, zero at the edges and on the last index of each axis.
 * Returns 0, or 2 for a shape below 2 or Symmetry outside 0..2. */
int getnp4_(int *ex, double *X, double *Y, double *Z, double *chi, double *trK, double *dxx, double *gxy,
            double *gxz, double *dyy, double *gyz, double *dzz, double *Axx, double *Axy, double *Axz,
            double *Ayy, double *Ayz, double *Azz, double *Gamxxx, double *Gamxxy, double *Gamxxz,
            double *Gamxyy, double *Gamxyz, double *Gamxzz, double *Gamyxx, double *Gamyxy, double *Gamyxz,
            double *Gamyyy, double *Gamyyz, double *Gamyzz, double *Gamzxx, double *Gamzxy, double *Gamzxz,
            double *Gamzyy, double *Gamzyz, double *Gamzzz, double *Rxx, double *Rxy, double *Rxz,
            double *Ryy, double *Ryz, double *Rzz, double *Rpsi4, double *Ipsi4, int *Symmetry);

/* Replace f_prolong3 / f_restrict3 (prolongrestrict.h:26-37) with the
 * Cell, ghost_width 3 routines of prolongrestrict_cell.f90 (prolong3
 * :1915-2164, restrict3 :2330-2468): 6-point, 5th-order interpolation from a
 * coarse array (bbox llbc..uubc, extc cells) to the fine cells of the region
 * llbp..uubp of a 2:1 fine array (llbf..uubf, extf), and back onto the
 * coarse cells of llbr..uubr.  SoA: the field's parity (mirrored halo below
 * the first cell).  A region the reference rejects, or whose stencils would
 * leave the source and its mirrored halo, writes nothing: prolong3_ returns
 * 1 (restrict3_ sets amss_last_error; the reference stops there); grids that
 * are not 2:1 or wei != 3: 2. */
int prolong3_(int *wei, double *llbc, double *uubc, int *extc, double *func, double *llbf, double *uubf,
              int *extf, double *funf, double *llbp, double *uubp, double *SoA, int *Symmetry);
void restrict3_(int *wei, double *llbc, double *uubc, int *extc, double *func, double *llbf, double *uubf,
                int *extf, double *funf, double *llbr, double *uubr, double *SoA, int *Symmetry);

/* ---------------------------------------------------------------------------
 * 2. Device-resident API
 * ------------------------------------------------------------------------- */

/* RHS of one block on device arrays (same outputs as compute_rhs_bssn_).
 * ex: shape with ghosts; h: dX, dY, dZ (= X[1]-X[0] ...); state/rhs: 24
 * device pointers in the reference field order; matter: 10 pointers or NULL
 * (vacuum); gam (18), ric (6): required; res (7): required when co == 0.
 * symmetry: the reference's Symmetry (0, 1, 2), as in compute_rhs_bssn_;
 * origin: X[0], Y[0], Z[0] of the block (read only when symmetry != 0).
 * stream: a hipStream_t or NULL.  flag (device int, may be NULL) is set to 1
 * when an input is NaN; the call itself does not synchronise. */
int amss_rhs_bssn_dev(const int *ex, const double *h, const double *const *state, double *const *rhs,
                      const double *const *matter, double *const *gam, double *const *ric, double *const *res,
                      int symmetry, const double *origin, double eps, int co, int gauge, int *flag,
                      void *stream);

/* amss_rhs_bssn_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as amss_rhs_bssn_dev; where that entry
 * point allocates (a zeroed matter array for a vacuum call, matter NULL; a
 * spare NaN flag when flag is NULL), this one uses a fixed device array of
 * the library (one per GPU, held until the process ends): a vacuum call
 * accepts at most 2,500,000 points (ghosts included) and returns 2 with a
 * message beyond that (pass zeroed matter arrays instead).  The first call on
 * a device looks the array up and clears it (a kernel on stream).  Calls that
 * overlap on one GPU share the array; only the spare flag is written, so such
 * calls race on that flag alone. */
int amss_rhs_bssn_dev_noalloc(const int *ex, const double *h, const double *const *state, double *const *rhs,
                              const double *const *matter, double *const *gam, double *const *ric,
                              double *const *res, int symmetry, const double *origin, double eps, int co,
                              int gauge, int *flag, void *stream);

/* prolong3_ / restrict3_ on device arrays, ordered on stream (no
 * synchronisation); same arguments and return codes. */
int amss_prolong3_dev(const double *llbc, const double *uubc, const int *extc, const double *func,
                      const double *llbf, const double *uubf, const int *extf, double *funf, const double *llbp,
                      const double *uubp, const double *SoA, void *stream);
int amss_restrict3_dev(const double *llbc, const double *uubc, const int *extc, double *func, const double *llbf,
                       const double *uubf, const int *extf, const double *funf, const double *llbr,
                       const double *uubr, const double *SoA, void *stream);

/* enforce_ga on device arrays (12 pointers: dxx gxy gxz dyy gyz dzz Axx..Azz). */
int amss_enforce_ga_dev(const int *ex, double *const *ga, void *stream);

/* getnp4_ on device arrays, ordered on stream (no synchronisation).
 * h: dX, dY, dZ; fields: 14 device pointers chi trK dxx gxy gxz dyy gyz dzz
 * Axx Axy Axz Ayy Ayz Azz; gam (18), ric (6): as compute_rhs_bssn_ wrote
 * them; X, Y, Z: device coordinates (ex[0], ex[1], ex[2] values); rpsi,
 * ipsi: outputs; origin: X[0], Y[0], Z[0] on the host (read only when
 * symmetry != 0). */
int amss_psi4_dev(const int *ex, const double *h, const double *const *fields, const double *const *gam,
                  const double *const *ric, const double *X, const double *Y, const double *Z, double *rpsi,
                  double *ipsi, int symmetry, const double *origin, void *stream);

/* The spin-weighted modes of Psi4 on the sphere of radius rex about the
 * origin (surface_integral::surf_Wave, surface_integral.C:197-377, with the
 * Gauss-Legendre rule of :39-181 that macrodef.h's GaussInt selects):
 * Rpsi4 / Ipsi4 of one block (device arrays of shape ex, e.g. from
 * amss_psi4_dev) are interpolated by global_interp's 6-point rule
 * (fmisc.f90:712-794) at 2N x 4N points (N = nquarter; a hemisphere for
 * Symmetry 1, an octant for 2, unfolded through the parities soa_r /
 * soa_i), and RP[k] + i IP[k] = rex int Psi4 sqrt((2l+1)/4pi)
 * d^l_{m,spinw}(theta) e^{-i m phi} dOmega for l = spinw..maxl, m = -l..l
 * (k in that order; sum (2l+1) entries).  X, Y, Z: the block's coordinates on
 * the host.  The sphere must lie in the block's owned cells (all but three
 * ghost cells at each end of an axis, none below a symmetry plane), else 2.
 * Synchronous: RP, IP are host arrays filled on return. */
int amss_surf_wave_dev(const int *ex, const double *X, const double *Y, const double *Z, const double *rpsi,
                       const double *ipsi, const double *soa_r, const double *soa_i, int symmetry, double rex,
                       int nquarter, int spinw, int maxl, double *RP, double *IP, void *stream);

/* ---- levels: Blocks + the Step driver ------------------------------------ */
typedef struct amss_level amss_level;

/* Boundary handling after each RK4 sub-step:
 *  AMSS_BC_FROZEN_GHOSTS: every point within gw of a patch face that no
 *    other patch of the level owns keeps the step's base state (the copy
 *    outer boundary of the benchmark configurations; gw 3 is the fast case,
 *    wider zones overwrite the first interior layers after each sub-step);
 *  AMSS_BC_COPY_LAYER: the reference's sommerfeld_rout(precor = 1) one-layer
 *    copy on the patch faces, with the reference's stencils up to the edge
 *    (the copied face points skip the RHS in the correctors and the
 *    residual-free predictor, so AMSS_LIST_RHS holds no RHS there);
 *  AMSS_BC_NONE: no boundary step;
 *  AMSS_BC_SOMMERFELD: the reference's level-0 outer boundary: each stage's
 *    RHS on the outermost layer of the patch faces is replaced by
 *    sommerfeld_routbam's before the update (bssn_class.C:3106-3117,
 *    3476-3484), with the reference's stencils up to the edge. */
#define AMSS_BC_FROZEN_GHOSTS 0
#define AMSS_BC_COPY_LAYER 1
#define AMSS_BC_NONE 2
#define AMSS_BC_SOMMERFELD 3

/* Grid-function lists of a Block (bssn_class.C:412-535). */
#define AMSS_LIST_STATE 0 /* StateList     (t^n)            */
#define AMSS_LIST_OLD 1   /* OldStateList  (t^{n-1})        */
#define AMSS_LIST_PRE 2   /* SynchList_pre (stage input)    */
#define AMSS_LIST_COR 3   /* SynchList_cor (stage output/k) */
#define AMSS_LIST_RHS 4   /* RHSList       (RK4 accumulator) */
#define AMSS_LIST_RES 5   /* Cons_Ham, Cons_Px..Pz, Cons_Gx..Gz (7 functions) */

/* One refinement level of npatch patches (one Block each).  shapes: 3 ints
 * per patch (points, ghosts included); bboxes: 6 doubles per patch (xmin
 * ymin zmin xmax ymax zmax of the cell-centred array, X = xmin + (i+1/2) h).
 * All patches share one grid spacing.  symmetry: the reference's Symmetry
 * (0; 1 equatorial, 2 octant), whose planes are the lower faces with a first
 * coordinate within one spacing of 0: stencils there read mirror points with
 * each field's parity and the one-layer boundaries skip those faces; it
 * needs bc 1, 2 or 3.  Returns NULL on error. */
amss_level *amss_level_create(int npatch, const int *shapes, const double *bboxes, int gauge, double eps,
                              double chitiny, int bc, int gw, int symmetry);
void amss_level_destroy(amss_level *L);
/* Device pointer of grid function `field` of `list` on `patch`. */
double *amss_level_field(amss_level *L, int patch, int list, int field);
/* Copy a whole grid function in (dir 0) or out (dir 1); host_or_dev is any
 * pointer hipMemcpy accepts.  Synchronous. */
int amss_level_copy(amss_level *L, int patch, int list, int field, double *host_or_dev, int dir);
/* Copy the 24 grid functions of a list of one patch from (dir 0) or to
 * (dir 1) 24 device arrays, ordered on the level's stream (no
 * synchronisation; order other streams after amss_level_stream). */
int amss_level_copy_list(amss_level *L, int patch, int list, double *const *ptrs, int dir);
/* Advance nsteps RK4 steps of size dT (bssn_class::Step order).
 * mode 0: one launch per patch and sub-step; mode 1: batched work queue over
 * all patches; mode 2: batched, replayed from a captured hipGraph.
 * want_ham: 0 no constraint residuals; 1 the residuals (Cons_Ham, Cons_P*,
 * Cons_G*) in the last step's predictor only; 2 in every step's predictor,
 * as the reference's Step does (co = 0 there, bssn_class.C:3086-3092), and
 * the seven L2 norms of every step recorded (Constraint_Out,
 * bssn_class.C:7475-7481; read them with amss_level_constraint_series).
 * Synchronises at the end; returns 1 if any RHS saw a NaN input. */
int amss_level_step(amss_level *L, double dT, int nsteps, int mode, int want_ham);
/* The L2 norm of Cons_Ham over all patches, from the last predictor that
 * computed residuals (Parallel::L2Norm, Parallel.C:4214-4243); the same as
 * amss_level_constraint_norms(L, -1, v) then v[0]. */
int amss_level_ham_l2(amss_level *L, double *out);
/* The seven constraint norms (Ham, Px, Py, Pz, Gx, Gy, Gz: Parallel::L2Norm
 * of Cons_Ham .. Cons_Gz, bssn_class.C:7475-7481) of one patch, or of all
 * patches together (patch -1: the square root of the summed squares), from
 * the last predictor that computed residuals.  The points summed are those of
 * l2normhelper (fmisc.f90:1263-1309): a level's Block always reaches its
 * patch's bbox, so under the reference's boundary modes (copy layer, none,
 * Sommerfeld) every point of the Block counts; with frozen ghost zones, whose
 * residuals are never evaluated, the points [gw, n - gw) do.  The reference's
 * Constraint_Out reports the first patch of a level (patch 0). */
int amss_level_constraint_norms(amss_level *L, int patch, double *out7);
/* The norms recorded by the last amss_level_step with want_ham 2: out[7 s + c]
 * for step s < min(maxsteps, *nsteps), norm c as above; *nsteps receives the
 * number of steps that call recorded. */
int amss_level_constraint_series(amss_level *L, int patch, double *out, int maxsteps, int *nsteps);
/* The stream the level's work is ordered on (a hipStream_t). */
void *amss_level_stream(amss_level *L);

/* ---- a hierarchy of refinement levels ------------------------------------ */
/* Replaces bssn_class::RecursiveStep with RestrictProlong (bssn_class.C:2323-2403,
 * 5724-5904; Parallel::Restrict, OutBdLow2Hi, prepare_inter_time_level,
 * Parallel.C:4329-4441, 4487-4640) for a hierarchy of nlev levels on one GPU.
 * npatch[l]: patches of level l; shapes / bboxes: all patches, level after
 * level, as amss_level_create.  Level 0 is a level of boundary mode bc0 and
 * ghost width 3; level l > 0 has half the spacing of level l - 1, and each of
 * its patch arrays holds the reference's buffer zone of 6 cells
 * (buffer_width, macrodef.h:88) on every face but a lower face on a symmetry
 * plane; such a patch must sit on the coarse cells and its buffer zone inside
 * the coarse patches less their own buffer zones and 3 ghost cells.  A fine
 * level's Step copies the base state onto the outermost layer of its faces
 * (sommerfeld_rout with cor, bssn_class.C:3124-3135), i.e. boundary mode
 * AMSS_BC_COPY_LAYER.  Returns NULL on error. */
typedef struct amss_amr amss_amr;
amss_amr *amss_amr_create(int nlev, const int *npatch, const int *shapes, const double *bboxes, int gauge,
                          double eps, double chitiny, int bc0, int symmetry);
void amss_amr_destroy(amss_amr *A);
/* Level lev of the hierarchy as a level handle (copy in / out, fields,
 * constraint norms, timing); owned by the hierarchy: amss_level_destroy on
 * it does nothing. */
amss_level *amss_amr_level(amss_amr *A, int lev);
/* nsteps coarse steps of size dT: level l takes 2^l steps of dT / 2^l, each
 * followed by the restriction of its state onto level l - 1 and the
 * prolongation of level l - 1 into its buffer zone, with level l - 1
 * interpolated in time to the intermediate time (average of two time levels
 * on the first coarse step, average2 of three after it).  mode / want_ham:
 * as amss_level_step, for every level step. */
int amss_amr_step(amss_amr *A, double dT, int nsteps, int mode, int want_ham);
/* The hierarchy's physical time (the finest level's, PhysTime). */
double amss_amr_time(amss_amr *A);

/* ---- on-disk formats of the reference ------------------------------------ */
/* One grid function of one patch in the layout of Parallel::writefile
 * (Parallel.C:1049-1071, read by plot_binary_data.py:37-43): time (double),
 * nx ny nz (int32), xmin xmax ymin ymax zmin zmax (double), then nx*ny*nz
 * doubles with x fastest.  The reference names such files
 * "<dir>/Lev%02d-%02d_<var>_%05d.bin" (Parallel.C:1361-1364); path is the
 * caller's.  Synchronous. */
int amss_level_write_bin(amss_level *L, const char *path, int patch, int list, int field, double time);
/* The level as a one-level checkpoint in the layout of
 * checkpoint::writecheck_cgh (checkpoint.C:132-191): time, levels = 1,
 * movls, BH_num_in = nbh (int32), grids = npatch (int32), Lt = time, per
 * patch bbox (6 doubles), shape (3 int32), handle (3 doubles; handles or
 * zeros), nbh puncture positions porg (3 doubles each), then per patch the
 * 24 StateList and the 24 OldStateList grid functions (the reference's
 * CheckList, bssn_class.C:682-683).  The reference's file name is
 * "<name>_cgh.CHK".  Synchronous. */
int amss_level_checkpoint_write(amss_level *L, const char *path, double time, int movls, int nbh,
                                const double *porg, const double *handles);
/* Restore StateList and OldStateList from such a file (checkpoint.C:193-292)
 * after checking that it holds one level whose patches have this level's
 * shapes and bboxes.  time, movls, nbh: out; porg: room for 3*maxbh doubles
 * (may be NULL when maxbh is 0; more punctures than maxbh is an error). */
int amss_level_checkpoint_read(amss_level *L, const char *path, double *time, int *movls, int *nbh, double *porg,
                               int maxbh);
/* The hierarchy as a multi-level checkpoint (checkpoint.C:132-191, every
 * level): time (PhysTime), levels, movls, BH_num_in = nbh, grids[levels],
 * Lt[levels] (each level's time), then per level its patch table (bbox,
 * shape, handle: handles holds 3 doubles per patch, level after level, or is
 * NULL for zeros) and its nbh puncture positions (porg: 3 doubles per
 * puncture, level after level), then per level and patch the 24 StateList
 * and 24 OldStateList grid functions.  Synchronous. */
int amss_amr_checkpoint_write(amss_amr *A, const char *path, int movls, int nbh, const double *porg,
                              const double *handles);
/* Restore such a file into a hierarchy of the same levels, patch shapes and
 * bboxes (checkpoint.C:193-292): both lists of every patch, PhysTime and each
 * level's time (the start time, which chooses between two- and three-level
 * time interpolation, is kept, as bssn_class.C:699-731 keeps StartTime).
 * porg: room for 3 * maxbh doubles per level. */
int amss_amr_checkpoint_read(amss_amr *A, const char *path, int *movls, int *nbh, double *porg, int maxbh);
/* Interior kernel of the fused RHS + RK4 sub-step: 0 reads every stencil
 * point through the vector L1/L2 (raw buffer loads); 1 stages a halo tile of
 * each field group in LDS (8x8x4-point tiles, double-buffered, filled by
 * direct global->LDS loads; the default).  Both give the same results to
 * rounding. */
int amss_level_set_variant(amss_level *L, int variant);
/* The interior kernel variant in use (-1 for a null level). */
int amss_level_variant(amss_level *L);
/* HIP-event timing of the fused RHS + RK4 interior kernel launches: enable
 * (resets the totals) or disable; then read the summed milliseconds and the
 * launch counts per sub-step 0..3 since the last enable. */
int amss_level_timing(amss_level *L, int enable);
int amss_level_kernel_ms(amss_level *L, double *ms4, long *launches4);
/* Batched step (modes 1, 2): the interior of every patch is cut into tiles of
 * one of AMSS_NUM_TILE_SHAPES shapes (0: 8x8x4, 1: 4x8x8, 2: 8x4x8, 3: 8x8x8 points);
 * one launch per sub-step walks the tile list of all patches with `grid`
 * persistent workgroups (0: one workgroup per tile), in list order or
 * (remap 1) with each XCD given a contiguous run of tiles.  autotune times
 * every (shape, grid in {0, 1, 2} x the resident workgroups, remap) candidate on sub-step 0 of
 * the current state (writes only the RHS and stage lists), keeps the fastest
 * and returns it in choice[3] (shape, grid, remap); ms (nms entries, may be
 * NULL) receives each candidate's time in that loop order.  The first
 * batched step autotunes by itself unless set_tune was called. */
#define AMSS_NUM_TILE_SHAPES 4
int amss_level_autotune(amss_level *L, double *ms, int nms, int *choice);
int amss_level_set_tune(amss_level *L, int shape, int grid, int remap);
int amss_level_get_tune(amss_level *L, int *choice);

/* ---- discrete adjoint of one RK4 step (GAUGE 0, vacuum, frozen ghosts) ----
 * No counterpart in the reference: the reverse of bssn_class::Step
 * (bssn_class.C:3042-3163, 3401-3526) with ghost width 3 and frozen ghost
 * zones, for gradients of a scalar of the evolved state; a level of several
 * patches includes the reverse of the inter-patch exchange (Parallel::Sync).
 * amss_adjoint_create: one Block of shape ex and spacing h;
 * amss_level_adjoint_create: npatch patches as in amss_level_create.
 * Array arguments are 24 device arrays per patch (patch-major: patch p,
 * field v at index 24 p + v) of the patch's full shape, in the reference
 * field order; calls are ordered on `stream` (a hipStream_t, NULL for the
 * default stream) and do not synchronise.  The object holds the stage
 * states and work arrays (about 520 grid functions per patch;
 * amss_adjoint_bytes).
 *  forward: U -> Unext (one RK4 step, the adjoint's own evaluation of it);
 *  step:    lam = dJ/dU^{n+1} in, dJ/dU^n out, U = the state at t^n;
 *  rhs:     out = dJ/dU of sum_p lam(p) . rhs(U)(p) over the points with
 *           +-3 room on every axis (the vector-Jacobian product of
 *           compute_rhs_bssn_ there). */
typedef struct amss_adjoint amss_adjoint;
amss_adjoint *amss_adjoint_create(const int *ex, const double *h, double eps, double chitiny, int gauge);
amss_adjoint *amss_level_adjoint_create(int npatch, const int *shapes, const double *bboxes, double eps,
                                        double chitiny, int gauge);
size_t amss_adjoint_bytes(amss_adjoint *A);
/* A tape of `slots` steps' stage data (192 grid functions per patch each):
 * forward_slot keeps the step's stage data in slot `slot`, and step_slot
 * then reverses that step from it without re-evaluating it (U must be the
 * state that forward_slot stepped). */
int amss_adjoint_reserve(amss_adjoint *A, int slots);
int amss_adjoint_forward_slot(amss_adjoint *A, double dT, const double *const *U, double *const *Unext, int slot,
                              void *stream);
int amss_adjoint_step_slot(amss_adjoint *A, double dT, const double *const *U, double *const *lam, int slot,
                           void *stream);
void amss_adjoint_destroy(amss_adjoint *A);
int amss_adjoint_forward(amss_adjoint *A, double dT, const double *const *U, double *const *Unext, void *stream);
int amss_adjoint_step(amss_adjoint *A, double dT, const double *const *U, double *const *lam, void *stream);
int amss_adjoint_rhs(amss_adjoint *A, const double *const *U, const double *const *lam, double *const *out,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif /* AMSS_BSSN_H */
