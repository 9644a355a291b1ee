"""ctypes front end to the CPU oracle of the BSSN RHS + RK4 path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the product path.

Two back ends with the same Python signatures:

* ``Orc``  -- the C restatement in ``oracle/liborc.so`` (``bssn_oracle.c``);
* ``Ref``  -- the reference Fortran itself, built by ``build_ref.sh`` into
  ``oracle/_ref/libamssref.so`` (only where the reference sources were
  present at build time).  Its configured gauge is GAUGE 2.

Arrays are numpy float64 of shape (nz, ny, nx), C-contiguous, which is the
reference's Fortran (nx, ny, nz) layout with i fastest.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NF = 24
FIELDS = ["chi", "trK", "dxx", "gxy", "gxz", "dyy", "gyz", "dzz",
          "Axx", "Axy", "Axz", "Ayy", "Ayz", "Azz", "Gamx", "Gamy", "Gamz",
          "Lap", "betax", "betay", "betaz", "dtSfx", "dtSfy", "dtSfz"]
# Reflection parity of each field (bssn_class.C:183-361).
PARITY = [(1, 1, 1), (1, 1, 1), (1, 1, 1), (-1, -1, 1), (-1, 1, -1), (1, 1, 1),
          (1, -1, -1), (1, 1, 1), (1, 1, 1), (-1, -1, 1), (-1, 1, -1), (1, 1, 1),
          (1, -1, -1), (1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1), (1, 1, 1),
          (-1, 1, 1), (1, -1, 1), (1, 1, -1), (-1, 1, 1), (1, -1, 1), (1, 1, -1)]
# Propagation speed in the Sommerfeld condition: sqrt(2) for chi, trK and the
# lapse (bssn_class.C:390-410), 1 otherwise (var.C:30).
PROPSPEED = [2 ** 0.5 if v in (0, 1, 17) else 1.0 for v in range(NF)]
GAM_NAMES = ["Gamxxx", "Gamxxy", "Gamxxz", "Gamxyy", "Gamxyz", "Gamxzz",
             "Gamyxx", "Gamyxy", "Gamyxz", "Gamyyy", "Gamyyz", "Gamyzz",
             "Gamzxx", "Gamzxy", "Gamzxz", "Gamzyy", "Gamzyz", "Gamzzz"]
RIC_NAMES = ["Rxx", "Rxy", "Rxz", "Ryy", "Ryz", "Rzz"]
RES_NAMES = ["ham_Res", "movx_Res", "movy_Res", "movz_Res", "Gmx_Res", "Gmy_Res", "Gmz_Res"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _ex(shape):
    nz, ny, nx = shape
    return (C.c_int * 3)(nx, ny, nz)


def _arr_ptrs(arrs):
    return (_dp * len(arrs))(*[_ptr(a) for a in arrs])


def have_ref() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libamssref.so"))


class Orc:
    """C restatement (liborc.so)."""

    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        self.lib = L = C.CDLL(path)
        L.orc_compute_rhs_bssn.restype = C.c_int
        L.orc_rungekutta4_rout.restype = C.c_int
        L.orc_l2normhelper.restype = C.c_double
        L.orc_bssn_step.restype = C.c_int
        self.gauge_default = 0

    # stencils ------------------------------------------------------------
    def fderivs(self, f, X, Y, Z, sym, symmetry=0):
        fx, fy, fz = (np.zeros_like(f) for _ in range(3))
        self.lib.orc_fderivs(_ex(f.shape), _ptr(f), _ptr(fx), _ptr(fy), _ptr(fz), _ptr(X), _ptr(Y),
                             _ptr(Z), C.c_double(sym[0]), C.c_double(sym[1]), C.c_double(sym[2]),
                             C.c_int(symmetry))
        return fx, fy, fz

    def fdderivs(self, f, X, Y, Z, sym, symmetry=0):
        o = [np.zeros_like(f) for _ in range(6)]
        self.lib.orc_fdderivs(_ex(f.shape), _ptr(f), *[_ptr(a) for a in o], _ptr(X), _ptr(Y), _ptr(Z),
                              C.c_double(sym[0]), C.c_double(sym[1]), C.c_double(sym[2]),
                              C.c_int(symmetry))
        return o

    def lopsided(self, f, f_rhs, bx, by, bz, X, Y, Z, sym, symmetry=0):
        r = f_rhs.copy()
        soa = (C.c_double * 3)(*sym)
        self.lib.orc_lopsided(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(f), _ptr(r), _ptr(bx),
                              _ptr(by), _ptr(bz), C.c_int(symmetry), soa)
        return r

    def kodis(self, f, f_rhs, X, Y, Z, sym, eps, symmetry=0):
        r = f_rhs.copy()
        soa = (C.c_double * 3)(*sym)
        self.lib.orc_kodis(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(f), _ptr(r), soa,
                           C.c_int(symmetry), C.c_double(eps))
        return r

    # RHS -----------------------------------------------------------------
    def compute_rhs(self, st, X, Y, Z, eps, co=1, symmetry=0, matter=None, gauge=None, res_init=None):
        gauge = self.gauge_default if gauge is None else gauge
        shape = st[0].shape
        rhs = [np.zeros(shape) for _ in range(NF)]
        gam = [np.zeros(shape) for _ in range(18)]
        ric = [np.zeros(shape) for _ in range(6)]
        res = [np.zeros(shape) if res_init is None else res_init[a].copy() for a in range(7)]
        mat = _arr_ptrs(matter) if matter is not None else None
        gont = self.lib.orc_compute_rhs_bssn(
            _ex(shape), C.c_double(0.0), _ptr(X), _ptr(Y), _ptr(Z), _arr_ptrs(st), _arr_ptrs(rhs), mat,
            _arr_ptrs(gam), _arr_ptrs(ric), _arr_ptrs(res), C.c_int(symmetry), C.c_int(0),
            C.c_double(eps), C.c_int(co), C.c_int(gauge))
        return dict(gont=gont, rhs=rhs, gam=gam, ric=ric, res=res)

    # pointwise / RK4 -------------------------------------------------------
    def enforce_ga(self, st):
        """In place on st[2:8] (metric) and st[8:14] (A)."""
        self.lib.orc_enforce_ga(_ex(st[0].shape), _arr_ptrs(st[2:8]), _arr_ptrs(st[8:14]))

    def rk4(self, f0, f1, frhs, dT, stage):
        return self.lib.orc_rungekutta4_rout(_ex(f0.shape), C.c_double(dT), _ptr(f0), _ptr(f1), _ptr(frhs),
                                             C.c_int(stage))

    def lowerboundset(self, chi, tinny):
        self.lib.orc_lowerboundset(_ex(chi.shape), _ptr(chi), C.c_double(tinny))

    def l2norm(self, f, X, Y, Z, bbox, gw):
        return self.lib.orc_l2normhelper(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), *[C.c_double(b) for b in bbox],
                                         _ptr(f), C.c_int(gw))

    def sommerfeld_copy(self, f0, f, X, Y, Z, bbox, symmetry=0):
        bb = (C.c_double * 6)(*bbox)
        self.lib.orc_sommerfeld_copy(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), bb, _ptr(f0), _ptr(f),
                                     C.c_int(symmetry))

    def sommerfeld_interp(self, f0, f, chi0, lap0, X, Y, Z, bbox, dT, soa, symmetry=0):
        """sommerfeld_rout with precor == 0 on f in place; returns the number of
        face points whose interpolation block decide3d rejects."""
        bb = (C.c_double * 6)(*bbox)
        sa = (C.c_double * 3)(*soa)
        return self.lib.orc_sommerfeld_interp(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), bb, C.c_double(dT),
                                              _ptr(chi0), _ptr(lap0), _ptr(f0), _ptr(f), sa, C.c_int(symmetry))

    def getnp4(self, st, gam, ric, X, Y, Z, symmetry=0):
        """Psi4 (getnp4, tetradtype 2) of state st from compute_rhs's gam and
        ric; returns (Rpsi4, Ipsi4)."""
        shape = st[0].shape
        rp, ip = np.zeros(shape), np.zeros(shape)
        self.lib.orc_getnp4(_ex(shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(st[0]), _ptr(st[1]),
                            _arr_ptrs(st[2:8]), _arr_ptrs(st[8:14]), _arr_ptrs(gam), _arr_ptrs(ric), _ptr(rp),
                            _ptr(ip), C.c_int(symmetry))
        return rp, ip

    def global_interp(self, f, X, Y, Z, pt, soa=(1.0, 1.0, 1.0), symmetry=0):
        """global_interp (Cell, order 6) of f at pt; returns (status, value)."""
        out = C.c_double(0.0)
        rc = self.lib.orc_global_interp(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(f), C.c_double(pt[0]),
                                        C.c_double(pt[1]), C.c_double(pt[2]), (C.c_double * 3)(*soa),
                                        C.c_int(symmetry), C.byref(out))
        return rc, out.value

    def surf_wave(self, rpsi, ipsi, X, Y, Z, rex, nquarter, maxl, spinw=2, soa_r=(1.0, 1.0, 1.0),
                  soa_i=(1.0, 1.0, 1.0), symmetry=0):
        """surface_integral::surf_Wave on one block; returns (status, RP, IP)."""
        nn = sum(2 * l + 1 for l in range(spinw, maxl + 1))
        rp, ip = np.zeros(nn), np.zeros(nn)
        rc = self.lib.orc_surf_wave(_ex(rpsi.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(rpsi), _ptr(ipsi),
                                    (C.c_double * 3)(*soa_r), (C.c_double * 3)(*soa_i), C.c_int(symmetry),
                                    C.c_double(rex), C.c_int(nquarter), C.c_int(spinw), C.c_int(maxl), _ptr(rp),
                                    _ptr(ip))
        return rc, rp, ip

    def prolong3(self, func, llbc, uubc, funf, llbf, uubf, llbp, uubp, soa, symmetry=0):
        """prolong3 into funf in place; returns the status (0 ok, 1 bad region)."""
        d3 = lambda v: (C.c_double * 3)(*v)  # noqa: E731
        return self.lib.orc_prolong3(C.c_int(3), d3(llbc), d3(uubc), _ex(func.shape), _ptr(func), d3(llbf),
                                     d3(uubf), _ex(funf.shape), _ptr(funf), d3(llbp), d3(uubp), d3(soa),
                                     C.c_int(symmetry))

    def restrict3(self, func, llbc, uubc, funf, llbf, uubf, llbr, uubr, soa, symmetry=0):
        """restrict3 into func in place; returns the status (0 ok, 1 bad region)."""
        d3 = lambda v: (C.c_double * 3)(*v)  # noqa: E731
        return self.lib.orc_restrict3(C.c_int(3), d3(llbc), d3(uubc), _ex(func.shape), _ptr(func), d3(llbf),
                                      d3(uubf), _ex(funf.shape), _ptr(funf), d3(llbr), d3(uubr), d3(soa),
                                      C.c_int(symmetry))

    def sommerfeld_bam(self, f_rhs, f0, X, Y, Z, bbox, velocity, soa, symmetry=0):
        """sommerfeld_routbam on f_rhs in place."""
        bb = (C.c_double * 6)(*bbox)
        sa = (C.c_double * 3)(*soa)
        self.lib.orc_sommerfeld_routbam(_ex(f0.shape), _ptr(X), _ptr(Y), _ptr(Z), bb, _ptr(f_rhs), _ptr(f0),
                                        C.c_double(velocity), sa, C.c_int(symmetry))

    def step(self, U, X, Y, Z, bbox, dT, eps, chitiny=1e-5, symmetry=0, gauge=None, bc=0, gw=3, want_ham=False):
        """One RK4 step in place on U (list of 24 arrays). Returns (gont, ham)."""
        gauge = self.gauge_default if gauge is None else gauge
        ham = np.zeros(U[0].shape) if want_ham else None
        bb = (C.c_double * 6)(*bbox)
        gont = self.lib.orc_bssn_step(_ex(U[0].shape), _ptr(X), _ptr(Y), _ptr(Z), bb, _arr_ptrs(U),
                                      C.c_double(dT), C.c_double(eps), C.c_double(chitiny), C.c_int(symmetry),
                                      C.c_int(0), C.c_int(gauge), C.c_int(bc), C.c_int(gw),
                                      _ptr(ham) if ham is not None else None)
        return gont, ham

    def philox(self, ctr, key):
        c = (C.c_uint * 4)(*ctr)
        k = (C.c_uint * 2)(*key)
        o = (C.c_uint * 4)()
        self.lib.orc_philox4x32_10(c, k, o)
        return list(o)


class Ref:
    """The reference Fortran (Fortran calling convention: every argument by reference)."""

    gauge_default = 2

    def __init__(self, opt: str = ""):
        path = os.path.join(HERE, "_ref", f"libamssref{opt}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference build absent: {path}")
        self.lib = C.CDLL(path)
        self.lib.compute_rhs_bssn_.restype = C.c_int

    @staticmethod
    def _d(x):
        return C.byref(C.c_double(x))

    @staticmethod
    def _i(x):
        return C.byref(C.c_int(x))

    def fderivs(self, f, X, Y, Z, sym, symmetry=0):
        fx, fy, fz = (np.zeros_like(f) for _ in range(3))
        self.lib.fderivs_(_ex(f.shape), _ptr(f), _ptr(fx), _ptr(fy), _ptr(fz), _ptr(X), _ptr(Y), _ptr(Z),
                          self._d(sym[0]), self._d(sym[1]), self._d(sym[2]), self._i(symmetry), self._i(0))
        return fx, fy, fz

    def fdderivs(self, f, X, Y, Z, sym, symmetry=0):
        o = [np.zeros_like(f) for _ in range(6)]
        self.lib.fdderivs_(_ex(f.shape), _ptr(f), *[_ptr(a) for a in o], _ptr(X), _ptr(Y), _ptr(Z),
                           self._d(sym[0]), self._d(sym[1]), self._d(sym[2]), self._i(symmetry), self._i(0))
        return o

    def lopsided(self, f, f_rhs, bx, by, bz, X, Y, Z, sym, symmetry=0):
        r = f_rhs.copy()
        soa = (C.c_double * 3)(*sym)
        self.lib.lopsided_(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(f), _ptr(r), _ptr(bx), _ptr(by),
                           _ptr(bz), self._i(symmetry), soa)
        return r

    def kodis(self, f, f_rhs, X, Y, Z, sym, eps, symmetry=0):
        r = f_rhs.copy()
        soa = (C.c_double * 3)(*sym)
        self.lib.kodis_(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(f), _ptr(r), soa, self._i(symmetry),
                        self._d(eps))
        return r

    def compute_rhs(self, st, X, Y, Z, eps, co=1, symmetry=0, matter=None, gauge=None, res_init=None):
        if gauge not in (None, 2):
            raise ValueError("the reference build is configured with GAUGE 2")
        shape = st[0].shape
        rhs = [np.zeros(shape) for _ in range(NF)]
        gam = [np.zeros(shape) for _ in range(18)]
        ric = [np.zeros(shape) for _ in range(6)]
        res = [np.zeros(shape) if res_init is None else res_init[a].copy() for a in range(7)]
        mat = matter if matter is not None else [np.zeros(shape) for _ in range(10)]
        args = [_ex(shape), self._d(0.0), _ptr(X), _ptr(Y), _ptr(Z)]
        args += [_ptr(a) for a in st] + [_ptr(a) for a in rhs] + [_ptr(a) for a in mat]
        args += [_ptr(a) for a in gam] + [_ptr(a) for a in ric] + [_ptr(a) for a in res]
        args += [self._i(symmetry), self._i(0), self._d(eps), self._i(co)]
        gont = self.lib.compute_rhs_bssn_(*args)
        return dict(gont=gont, rhs=rhs, gam=gam, ric=ric, res=res)

    def enforce_ga(self, st):
        self.lib.enforce_ga_(_ex(st[0].shape), *[_ptr(a) for a in st[2:14]])

    def rk4(self, f0, f1, frhs, dT, stage):
        self.lib.rungekutta4_rout_(_ex(f0.shape), self._d(dT), _ptr(f0), _ptr(f1), _ptr(frhs), self._i(stage))
        return 0

    def lowerboundset(self, chi, tinny):
        self.lib.lowerboundset_(_ex(chi.shape), _ptr(chi), self._d(tinny))

    def l2norm(self, f, X, Y, Z, bbox, gw):
        out = C.c_double(0.0)
        self.lib.l2normhelper_(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), *[self._d(b) for b in bbox],
                               _ptr(f), C.byref(out), self._i(gw))
        return out.value

    def sommerfeld_copy(self, f0, f, X, Y, Z, bbox, symmetry=0):
        """sommerfeld_rout with precor == 1 (the copy branch)."""
        soa = (C.c_double * 3)(1.0, 1.0, 1.0)
        dummy = np.zeros_like(f0)
        self.lib.sommerfeld_rout_(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), *[self._d(b) for b in bbox],
                                  self._d(0.0), _ptr(dummy), _ptr(dummy), _ptr(f0), _ptr(f), soa,
                                  self._i(symmetry), self._i(1))

    def sommerfeld_interp(self, f0, f, chi0, lap0, X, Y, Z, bbox, dT, soa, symmetry=0):
        """sommerfeld_rout with precor == 0 (the interpolating predictor)."""
        sa = (C.c_double * 3)(*soa)
        self.lib.sommerfeld_rout_(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), *[self._d(b) for b in bbox],
                                  self._d(dT), _ptr(chi0), _ptr(lap0), _ptr(f0), _ptr(f), sa,
                                  self._i(symmetry), self._i(0))
        return 0

    def getnp4(self, st, gam, ric, X, Y, Z, symmetry=0):
        shape = st[0].shape
        rp, ip = np.zeros(shape), np.zeros(shape)
        self.lib.getnp4_(_ex(shape), _ptr(X), _ptr(Y), _ptr(Z), *[_ptr(a) for a in st[:14]],
                         *[_ptr(a) for a in gam], *[_ptr(a) for a in ric], _ptr(rp), _ptr(ip), self._i(symmetry))
        return rp, ip

    def global_interp(self, f, X, Y, Z, pt, soa=(1.0, 1.0, 1.0), symmetry=0):
        """global_interp_ (stops the program where decide3d fails: call it with
        points whose block stays inside); returns (0, value)."""
        out = C.c_double(0.0)
        self.lib.global_interp_(_ex(f.shape), _ptr(X), _ptr(Y), _ptr(Z), _ptr(f), C.byref(out), self._d(pt[0]),
                                self._d(pt[1]), self._d(pt[2]), self._i(6), (C.c_double * 3)(*soa),
                                self._i(symmetry))
        return 0, out.value

    def prolong3(self, func, llbc, uubc, funf, llbf, uubf, llbp, uubp, soa, symmetry=0):
        """prolong3_ (a subroutine: no status) into funf in place."""
        d3 = lambda v: (C.c_double * 3)(*v)  # noqa: E731
        self.lib.prolong3_(self._i(3), d3(llbc), d3(uubc), _ex(func.shape), _ptr(func), d3(llbf), d3(uubf),
                           _ex(funf.shape), _ptr(funf), d3(llbp), d3(uubp), d3(soa), self._i(symmetry))
        return 0

    def restrict3(self, func, llbc, uubc, funf, llbf, uubf, llbr, uubr, soa, symmetry=0):
        """restrict3_ into func in place."""
        d3 = lambda v: (C.c_double * 3)(*v)  # noqa: E731
        self.lib.restrict3_(self._i(3), d3(llbc), d3(uubc), _ex(func.shape), _ptr(func), d3(llbf), d3(uubf),
                            _ex(funf.shape), _ptr(funf), d3(llbr), d3(uubr), d3(soa), self._i(symmetry))
        return 0

    def sommerfeld_bam(self, f_rhs, f0, X, Y, Z, bbox, velocity, soa, symmetry=0):
        """sommerfeld_routbam_ on f_rhs in place."""
        sa = (C.c_double * 3)(*soa)
        self.lib.sommerfeld_routbam_(_ex(f0.shape), _ptr(X), _ptr(Y), _ptr(Z), *[self._d(b) for b in bbox],
                                     _ptr(f_rhs), _ptr(f0), self._d(velocity), sa, self._i(symmetry))

    def symmetry_bd(self, ord_, f, sym):
        nz, ny, nx = f.shape
        out = np.zeros((nz + ord_, ny + ord_, nx + ord_))
        soa = (C.c_double * 3)(*sym)
        self.lib.symmetry_bd_(self._i(ord_), _ex(f.shape), _ptr(f), _ptr(out), soa)
        return out

    def step(self, U, X, Y, Z, bbox, dT, eps, chitiny=1e-5, symmetry=0, gauge=None, bc=0, gw=3, want_ham=False):
        """One RK4 step driven in the order of bssn_class::Step, calling the
        reference routines (bssn_class.C:2981-3901); boundary as in Orc.step."""
        if gauge not in (None, 2):
            raise ValueError("the reference build is configured with GAUGE 2")
        return step_driver(self, U, X, Y, Z, bbox, dT, eps, chitiny, symmetry, bc, gw, want_ham)


def _ghost_copy(f0, f, gw):
    nz, ny, nx = f.shape
    m = np.ones(f.shape, bool)
    m[gw:nz - gw, gw:ny - gw, gw:nx - gw] = False
    f[m] = f0[m]


def step_driver(be, U, X, Y, Z, bbox, dT, eps, chitiny, symmetry, bc, gw, want_ham):
    """Python statement of bssn_class::Step for one block, over a back end's
    array routines (used with the reference build)."""
    tinny = chitiny - 1.0
    shape = U[0].shape
    pre = [np.zeros(shape) for _ in range(NF)]
    cor = [np.zeros(shape) for _ in range(NF)]

    def bcf(f0, f):
        if bc == 0:
            _ghost_copy(f0, f, gw)
        elif bc == 1:
            be.sommerfeld_copy(f0, f, X, Y, Z, bbox, symmetry)

    def bam(f_rhs, f0, v):
        if bc == 3:  # level-0 Sommerfeld on the stage's RHS (bssn_class.C:3106-3117, 3476-3484)
            be.sommerfeld_bam(f_rhs, f0, X, Y, Z, bbox, PROPSPEED[v], PARITY[v], symmetry)

    be.enforce_ga(U)
    o = be.compute_rhs(U, X, Y, Z, eps, co=0, symmetry=symmetry)
    gont = o["gont"]
    acc = o["rhs"]
    ham = o["res"][0].copy() if want_ham else None
    for v in range(NF):
        bam(acc[v], U[v], v)
        be.rk4(U[v], pre[v], acc[v], dT, 0)
        bcf(U[v], pre[v])
    be.lowerboundset(pre[0], tinny)
    for it in (1, 2, 3):
        be.enforce_ga(pre)
        o = be.compute_rhs(pre, X, Y, Z, eps, co=1, symmetry=symmetry)
        gont |= o["gont"]
        for v in range(NF):
            np.copyto(cor[v], o["rhs"][v])
            bam(cor[v], pre[v], v)
            be.rk4(U[v], cor[v], acc[v], dT, it)
            bcf(U[v], cor[v])
        be.lowerboundset(cor[0], tinny)
        if it < 3:
            pre, cor = cor, pre
    for v in range(NF):
        np.copyto(U[v], cor[v])
    return gont, ham


# ---------------------------------------------------------------------------
# Berger-Oliger sub-cycling over levels of one patch each, in the order of
# bssn_class::RecursiveStep (bssn_class.C:2323-2403) and RestrictProlong
# (:5724-5904, RPB 0, MIXOUTB 0), with the regions of Parallel::Restrict
# (build_owned_gsl2 onto build_complete_gsl, Parallel.C:4487-4530, 2762-2815)
# and Parallel::OutBdLow2Hi (build_buffer_gsl from build_owned_gsl4,
# Parallel.C:4580-4612, 4794-4843, 2922-3013).
BUFFER_WIDTH = 6  # macrodef.h:88


class AmrLevel:
    """One level (one patch): the array's bbox and shape, its buffer widths
    (lli, uui), and the reference's State / Old / cor lists after a step."""

    def __init__(self, U, bbox, lli=(0, 0, 0), uui=(0, 0, 0)):
        nz, ny, nx = U[0].shape
        self.shape = (nx, ny, nz)
        self.bbox = tuple(bbox)
        self.h = tuple((bbox[3 + a] - bbox[a]) / self.shape[a] for a in range(3))
        self.X, self.Y, self.Z = [bbox[a] + (np.arange(self.shape[a]) + 0.5) * self.h[a] for a in range(3)]
        self.lli, self.uui = tuple(lli), tuple(uui)
        self.U = [u.copy() for u in U]
        self.old = [u.copy() for u in U]
        self.cor = [u.copy() for u in U]


def _overlap(a, b, hmin):
    lo = [max(a[0][d], b[0][d]) for d in range(3)]
    hi = [min(a[1][d], b[1][d]) for d in range(3)]
    if any(hi[d] - lo[d] < hmin[d] / 2 for d in range(3)):  # the Cell rule of build_gstl
        return None
    return lo, hi


def _on_plane(axis, symmetry, x, h):
    return (symmetry > 0 if axis == 2 else symmetry > 1) and abs(x) < h / 2


def amr_restrict_prolong(be, C, F, same_time, three, symmetry=0):
    """RestrictProlong between coarse level C and fine level F (one patch each)."""
    if not same_time:  # prepare_inter_time_level into SynchList_pre (average2 / average)
        if three:
            tgt = [3.0 / 8.0 * a + 3.0 / 4.0 * b + -1.0 / 8.0 * c for a, b, c in zip(C.U, C.old, C.cor)]
        else:
            tgt = [0.5 * (a + b) for a, b in zip(C.U, C.old)]
    else:
        tgt = C.U
    owned2 = ([F.bbox[a] + F.lli[a] * F.h[a] for a in range(3)], [F.bbox[3 + a] - F.uui[a] * F.h[a] for a in range(3)])
    complete = (list(C.bbox[:3]), list(C.bbox[3:]))
    r = _overlap(owned2, complete, F.h)
    if r is not None:
        for v in range(NF):
            assert be.restrict3(tgt[v], C.bbox[:3], C.bbox[3:], F.U[v], F.bbox[:3], F.bbox[3:], r[0], r[1],
                                PARITY[v], symmetry) == 0
    # the fine buffer zone: the whole array less the inner box, as slabs
    inner_lo, inner_hi = owned2
    rest_lo, rest_hi = list(F.bbox[:3]), list(F.bbox[3:])
    slabs = []
    for a in (2, 1, 0):
        if F.lli[a] > 0:
            hi = list(rest_hi)
            hi[a] = inner_lo[a]
            slabs.append((list(rest_lo), hi))
        if F.uui[a] > 0:
            lo = list(rest_lo)
            lo[a] = inner_hi[a]
            slabs.append((lo, list(rest_hi)))
        rest_lo[a], rest_hi[a] = inner_lo[a], inner_hi[a]
    owned4 = ([C.bbox[a] if _on_plane(a, symmetry, C.bbox[a], C.h[a]) else C.bbox[a] + (C.lli[a] + 3) * C.h[a]
               for a in range(3)], [C.bbox[3 + a] - (C.uui[a] + 3) * C.h[a] for a in range(3)])
    for s in slabs:
        r = _overlap(s, owned4, F.h)
        if r is None:
            continue
        for v in range(NF):
            assert be.prolong3(tgt[v], C.bbox[:3], C.bbox[3:], F.U[v], F.bbox[:3], F.bbox[3:], r[0], r[1],
                               PARITY[v], symmetry) == 0


def amr_step_driver(be, levels, dT, eps, chitiny=1e-5, gauge=None, bc0=0, symmetry=0, t=(0.0, 0.0)):
    """One coarse step of the hierarchy `levels` (AmrLevel list, coarse first)
    in place; t = (PhysTime, StartTime) before the step.  Returns the new
    PhysTime."""
    state = {"t": t[0]}

    def level_step(L, lev, dTl):
        # Step enforces StateList in place before the predictor
        # (bssn_class.C:3051-3057), so the state that the rotation makes the
        # old one is the enforced U^n
        base = [u.copy() for u in L.U]
        be.enforce_ga(base)
        new = [u.copy() for u in L.U]
        g, _ = be.step(new, L.X, L.Y, L.Z, L.bbox, dTl, eps, chitiny, symmetry=symmetry, gauge=gauge,
                       bc=bc0 if lev == 0 else 1, gw=3)
        assert g == 0
        # the list rotation (bssn_class.C:3845-3859): State <- new, Old <- State, cor <- Old
        L.cor, L.old, L.U = L.old, base, new

    def rec(lev):
        dTl = dT * 0.5 ** lev
        nit = 1 if lev == 0 else 2
        for i in range(nit):
            same_time = i == nit - 1
            level_step(levels[lev], lev, dTl)
            if lev + 1 < len(levels):
                rec(lev + 1)
            else:
                state["t"] += dT * 0.5 ** lev
            if lev > 0:
                three = (state["t"] - dTl) - t[1] > dTl / 2  # fgt(PhysTime - dT_lev, StartTime, dT_lev / 2)
                amr_restrict_prolong(be, levels[lev - 1], levels[lev], same_time, three, symmetry)

    rec(0)
    return state["t"]
