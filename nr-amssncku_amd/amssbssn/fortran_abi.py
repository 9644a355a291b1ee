"""Calls of the library's Fortran-ABI replacements (compute_rhs_bssn_,
rungekutta4_rout_, enforce_ga_, lowerboundset_, l2normhelper_,
sommerfeld_rout_, sommerfeld_routbam_, prolong3_, restrict3_, getnp4_) on host numpy
arrays, with the same Python signatures as the reference build's, so a
caller switches by changing the library.
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


def sommerfeld_interp(f0, f, chi0, lap0, X, Y, Z, bbox, dT, soa, symmetry=0):
    """sommerfeld_rout_ with precor 0 (the interpolating predictor) on f in place."""
    sa = (C.c_double * 3)(*soa)
    lib().sommerfeld_rout_(_ex(f.shape), _p(X), _p(Y), _p(Z), *[_d(b) for b in bbox], _d(dT), _p(chi0),
                           _p(lap0), _p(f0), _p(f), sa, _i(symmetry), _i(0))


def sommerfeld_bam(f_rhs, f0, X, Y, Z, bbox, velocity, soa, symmetry=0):
    """sommerfeld_routbam_ on f_rhs in place."""
    sa = (C.c_double * 3)(*soa)
    lib().sommerfeld_routbam_(_ex(f0.shape), _p(X), _p(Y), _p(Z), *[_d(b) for b in bbox], _p(f_rhs), _p(f0),
                              _d(velocity), sa, _i(symmetry))


def getnp4(st, gam, ric, X, Y, Z, symmetry=0):
    """getnp4_ (Psi4 in the radial tetrad) of the state st (24 or the first 14
    arrays) from compute_rhs's gam (18) and ric (6); returns (Rpsi4, Ipsi4)."""
    shape = st[0].shape
    rp, ip = np.zeros(shape), np.zeros(shape)
    check(lib().getnp4_(_ex(shape), _p(X), _p(Y), _p(Z), *[_p(a) for a in st[:14]], *[_p(a) for a in gam],
                        *[_p(a) for a in ric], _p(rp), _p(ip), _i(symmetry)), "getnp4_")
    return rp, ip


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def prolong3(func, llbc, uubc, funf, llbf, uubf, llbp, uubp, soa, symmetry=0):
    """prolong3_ into funf in place; returns the status (0, 1 bad region, 2 bad grids)."""
    return lib().prolong3_(_i(3), _d3(llbc), _d3(uubc), _ex(func.shape), _p(func), _d3(llbf), _d3(uubf),
                           _ex(funf.shape), _p(funf), _d3(llbp), _d3(uubp), _d3(soa), _i(symmetry))


def restrict3(func, llbc, uubc, funf, llbf, uubf, llbr, uubr, soa, symmetry=0):
    """restrict3_ into func in place (a bad region leaves func untouched and
    sets the library's last error)."""
    lib().restrict3_(_i(3), _d3(llbc), _d3(uubc), _ex(func.shape), _p(func), _d3(llbf), _d3(uubf),
                     _ex(funf.shape), _p(funf), _d3(llbr), _d3(uubr), _d3(soa), _i(symmetry))
    return 0


def compute_rhs_dev(state, h, eps, co=1, symmetry=0, origin=(0.0, 0.0, 0.0), gauge=0, matter=None, stream=None):
    """amss_rhs_bssn_dev, the device-resident form of compute_rhs_bssn_, on
    torch tensors: state is (24, nz, ny, nx) float64 on the GPU (matter: 10
    such arrays or None for vacuum).  Returns dict(flag, rhs, gam, ric, res)
    of device tensors after synchronising the stream."""
    import torch
    assert state.dtype == torch.float64 and state.is_cuda and state.is_contiguous() and state.shape[0] == NF
    shape = tuple(state.shape[1:])
    dev = state.device

    def ptrs(t):
        return (C.c_void_p * t.shape[0])(*[t[i].data_ptr() for i in range(t.shape[0])])

    rhs = torch.zeros((NF,) + shape, dtype=torch.float64, device=dev)
    gam = torch.zeros((18,) + shape, dtype=torch.float64, device=dev)
    ric = torch.zeros((6,) + shape, dtype=torch.float64, device=dev)
    res = torch.zeros((7,) + shape, dtype=torch.float64, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    mat = None
    if matter is not None:
        assert matter.shape[0] == 10 and matter.is_contiguous()
        mat = ptrs(matter)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    hh = (C.c_double * 3)(*h)
    org = (C.c_double * 3)(*origin)
    check(lib().amss_rhs_bssn_dev(_ex(shape), hh, ptrs(state), ptrs(rhs), mat, ptrs(gam), ptrs(ric), ptrs(res),
                                  symmetry, org, eps, co, gauge, C.c_void_p(flag.data_ptr()),
                                  C.c_void_p(s.cuda_stream)), "amss_rhs_bssn_dev")
    s.synchronize()
    return dict(flag=int(flag.item()), rhs=rhs, gam=gam, ric=ric, res=res)


def psi4_dev(state, gam, ric, X, Y, Z, h, symmetry=0, origin=(0.0, 0.0, 0.0), stream=None):
    """amss_psi4_dev on torch tensors: state (24 or 14, nz, ny, nx), gam
    (18, ...), ric (6, ...) and the coordinate vectors X, Y, Z, all float64 on
    the GPU.  Returns (rpsi, ipsi) after synchronising the stream."""
    import torch
    for t in (state, gam, ric, X, Y, Z):
        assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous()
    shape = tuple(state.shape[1:])
    assert state.shape[0] >= 14 and gam.shape == (18,) + shape and ric.shape == (6,) + shape
    assert (X.numel(), Y.numel(), Z.numel()) == shape[::-1]
    dev = state.device

    def ptrs(t, n):
        return (C.c_void_p * n)(*[t[i].data_ptr() for i in range(n)])

    rpsi = torch.empty(shape, dtype=torch.float64, device=dev)
    ipsi = torch.empty(shape, dtype=torch.float64, device=dev)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    vp = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    check(lib().amss_psi4_dev(_ex(shape), (C.c_double * 3)(*h), ptrs(state, 14), ptrs(gam, 18), ptrs(ric, 6),
                              vp(X), vp(Y), vp(Z), vp(rpsi), vp(ipsi), symmetry, (C.c_double * 3)(*origin),
                              C.c_void_p(s.cuda_stream)), "amss_psi4_dev")
    s.synchronize()
    return rpsi, ipsi


def surf_wave_dev(rpsi, ipsi, X, Y, Z, rex, nquarter, maxl, spinw=2, soa_r=(1.0, 1.0, 1.0), soa_i=(1.0, 1.0, 1.0),
                  symmetry=0, stream=None):
    """amss_surf_wave_dev: the spin-weighted modes of Psi4 (torch float64
    device arrays of the block's shape) on the sphere of radius rex; X, Y, Z
    are the block's coordinates (numpy).  Returns (RP, IP) as numpy arrays in
    the order l = spinw..maxl, m = -l..l."""
    import torch
    for t in (rpsi, ipsi):
        assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous()
    shape = tuple(rpsi.shape)
    assert tuple(ipsi.shape) == shape and (len(X), len(Y), len(Z)) == shape[::-1]
    nn = sum(2 * l + 1 for l in range(spinw, maxl + 1))
    rp, ip = np.zeros(nn), np.zeros(nn)
    Xc, Yc, Zc = (np.ascontiguousarray(a, dtype=np.float64) for a in (X, Y, Z))
    s = stream if stream is not None else torch.cuda.current_stream(rpsi.device)
    check(lib().amss_surf_wave_dev(_ex(shape), _p(Xc), _p(Yc), _p(Zc), C.c_void_p(rpsi.data_ptr()),
                                   C.c_void_p(ipsi.data_ptr()), (C.c_double * 3)(*soa_r), (C.c_double * 3)(*soa_i),
                                   symmetry, rex, nquarter, spinw, maxl, _p(rp), _p(ip), C.c_void_p(s.cuda_stream)),
          "amss_surf_wave_dev")
    return rp, ip
