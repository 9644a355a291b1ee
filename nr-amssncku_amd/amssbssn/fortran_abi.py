"""Calls of the library's Fortran-ABI replacements (compute_rhs_bssn_,
rungekutta4_rout_, enforce_ga_, lowerboundset_, l2normhelper_,
sommerfeld_rout_) on host numpy arrays, with the same Python signatures as
the reference build's, so a caller switches by changing the library.
Arrays: float64, shape (nz, ny, nx), C-contiguous.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .grid import NF

_dp = C.POINTER(C.c_double)


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _ex(shape):
    nz, ny, nx = shape
    return (C.c_int * 3)(nx, ny, nz)


def _d(x):
    return C.byref(C.c_double(x))


def _i(x):
    return C.byref(C.c_int(x))


def set_gauge(gauge: int) -> None:
    check(lib().amss_set_gauge(gauge), "amss_set_gauge")


def compute_rhs(st, X, Y, Z, eps, co=1, symmetry=0, matter=None, res_init=None):
    """compute_rhs_bssn_ on the GPU; returns dict(gont, rhs, gam, ric, res)."""
    shape = st[0].shape
    rhs = [np.zeros(shape) for _ in range(NF)]
    gam = [np.zeros(shape) for _ in range(18)]
    ric = [np.zeros(shape) for _ in range(6)]
    res = [np.zeros(shape) if res_init is None else res_init[a].copy() for a in range(7)]
    mat = matter if matter is not None else [np.zeros(shape) for _ in range(10)]
    args = [_ex(shape), _d(0.0), _p(X), _p(Y), _p(Z)]
    args += [_p(a) for a in st] + [_p(a) for a in rhs] + [_p(a) for a in mat]
    args += [_p(a) for a in gam] + [_p(a) for a in ric] + [_p(a) for a in res]
    args += [_i(symmetry), _i(0), _d(eps), _i(co)]
    gont = lib().compute_rhs_bssn_(*args)
    if gont not in (0, 1):
        check(gont, "compute_rhs_bssn_")
    return dict(gont=gont, rhs=rhs, gam=gam, ric=ric, res=res)


def rk4(f0, f1, frhs, dT, stage):
    return lib().rungekutta4_rout_(_ex(f0.shape), _d(dT), _p(f0), _p(f1), _p(frhs), _i(stage))


def enforce_ga(st):
    lib().enforce_ga_(_ex(st[0].shape), *[_p(a) for a in st[2:14]])


def lowerboundset(chi, tinny):
    lib().lowerboundset_(_ex(chi.shape), _p(chi), _d(tinny))


def l2norm(f, X, Y, Z, bbox, gw):
    out = C.c_double(0.0)
    lib().l2normhelper_(_ex(f.shape), _p(X), _p(Y), _p(Z), *[_d(b) for b in bbox], _p(f), C.byref(out), _i(gw))
    return out.value


def sommerfeld_copy(f0, f, X, Y, Z, bbox, symmetry=0):
    soa = (C.c_double * 3)(1.0, 1.0, 1.0)
    dummy = np.zeros_like(f0)
    lib().sommerfeld_rout_(_ex(f.shape), _p(X), _p(Y), _p(Z), *[_d(b) for b in bbox], _d(0.0), _p(dummy),
                           _p(dummy), _p(f0), _p(f), soa, _i(symmetry), _i(1))
