"""ctypes binding of libamssbssn.so (the gfx950 BSSN RHS + RK4 library).

The shared library is built in place next to this file (``make -C
nr-amssncku_amd/csrc``).  There is no CPU fallback: if the library is missing
or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libamssbssn.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None

# Exported C symbols and their argument / result types (include/amss_bssn.h).
SIGNATURES = {
    "amss_version": ([], C.c_int),
    "amss_last_error": ([], C.c_char_p),
    "amss_set_gauge": ([C.c_int], C.c_int),
    "compute_rhs_bssn_": (None, C.c_int),  # 98 by-reference arguments
    "rungekutta4_rout_": ([_ip, _dp, _dp, _dp, _dp, _ip], C.c_int),
    "enforce_ga_": ([_ip] + [_dp] * 12, None),
    "lowerboundset_": ([_ip, _dp, _dp], None),
    "l2normhelper_": ([_ip, _dp, _dp, _dp] + [_dp] * 6 + [_dp, _dp, _ip], None),
    "sommerfeld_rout_": ([_ip, _dp, _dp, _dp] + [_dp] * 6 + [_dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip], None),
    "sommerfeld_routbam_": ([_ip, _dp, _dp, _dp] + [_dp] * 6 + [_dp, _dp, _dp, _dp, _ip], None),
    "prolong3_": ([_ip, _dp, _dp, _ip, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _ip], C.c_int),
    "restrict3_": ([_ip, _dp, _dp, _ip, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _ip], None),
    "amss_prolong3_dev": ([_dp, _dp, _ip, C.c_void_p, _dp, _dp, _ip, C.c_void_p, _dp, _dp, _dp, C.c_void_p],
                          C.c_int),
    "amss_restrict3_dev": ([_dp, _dp, _ip, C.c_void_p, _dp, _dp, _ip, C.c_void_p, _dp, _dp, _dp, C.c_void_p],
                           C.c_int),
    "amss_rhs_bssn_dev": ([_ip, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_int, _dp, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p], C.c_int),
    "amss_rhs_bssn_dev_noalloc": ([_ip, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_int, _dp, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p],
                                  C.c_int),
    "amss_enforce_ga_dev": ([_ip, C.c_void_p, C.c_void_p], C.c_int),
    "getnp4_": ([_ip, _dp, _dp, _dp] + [_dp] * 40 + [_ip], C.c_int),
    "amss_surf_wave_dev": ([_ip, _dp, _dp, _dp, C.c_void_p, C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int,
                            C.c_int, C.c_int, _dp, _dp, C.c_void_p], C.c_int),
    "amss_psi4_dev": ([_ip, _dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_void_p], C.c_int),
    "amss_level_create": ([C.c_int, _ip, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int],
                          C.c_void_p),
    "amss_level_destroy": ([C.c_void_p], None),
    "amss_level_field": ([C.c_void_p, C.c_int, C.c_int, C.c_int], C.c_void_p),
    "amss_level_copy": ([C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int], C.c_int),
    "amss_level_step": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int], C.c_int),
    "amss_level_ham_l2": ([C.c_void_p, _dp], C.c_int),
    "amss_level_constraint_norms": ([C.c_void_p, C.c_int, _dp], C.c_int),
    "amss_level_constraint_series": ([C.c_void_p, C.c_int, _dp, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "amss_level_stream": ([C.c_void_p], C.c_void_p),
    "amss_amr_create": ([C.c_int, _ip, _ip, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int], C.c_void_p),
    "amss_amr_destroy": ([C.c_void_p], None),
    "amss_amr_level": ([C.c_void_p, C.c_int], C.c_void_p),
    "amss_amr_step": ([C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int], C.c_int),
    "amss_amr_time": ([C.c_void_p], C.c_double),
    "amss_level_write_bin": ([C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double], C.c_int),
    "amss_level_checkpoint_write": ([C.c_void_p, C.c_char_p, C.c_double, C.c_int, C.c_int, _dp, _dp], C.c_int),
    "amss_level_checkpoint_read": ([C.c_void_p, C.c_char_p, _dp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp,
                                   C.c_int], C.c_int),
    "amss_amr_checkpoint_write": ([C.c_void_p, C.c_char_p, C.c_int, C.c_int, _dp, _dp], C.c_int),
    "amss_amr_checkpoint_read": ([C.c_void_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp, C.c_int],
                                 C.c_int),
    "amss_level_timing": ([C.c_void_p, C.c_int], C.c_int),
    "amss_level_set_variant": ([C.c_void_p, C.c_int], C.c_int),
    "amss_level_variant": ([C.c_void_p], C.c_int),
    "amss_level_copy_list": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int], C.c_int),
    "amss_level_autotune": ([C.c_void_p, _dp, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "amss_level_set_tune": ([C.c_void_p, C.c_int, C.c_int, C.c_int], C.c_int),
    "amss_level_get_tune": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "amss_adjoint_create": ([_ip, _dp, C.c_double, C.c_double, C.c_int], C.c_void_p),
    "amss_adjoint_destroy": ([C.c_void_p], None),
    "amss_level_adjoint_create": ([C.c_int, _ip, _dp, C.c_double, C.c_double, C.c_int], C.c_void_p),
    "amss_adjoint_bytes": ([C.c_void_p], C.c_size_t),
    "amss_adjoint_reserve": ([C.c_void_p, C.c_int], C.c_int),
    "amss_adjoint_forward_slot": ([C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "amss_adjoint_step_slot": ([C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "amss_adjoint_forward": ([C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "amss_adjoint_step": ([C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "amss_adjoint_rhs": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "amss_level_kernel_ms": ([C.c_void_p, _dp, C.POINTER(C.c_long)], C.c_int),
}


class AmssError(RuntimeError):
    pass


def lib():
    """Load (once) and return the library; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AmssError(f"libamssbssn.so is not built ({LIB_PATH}); run `make -C nr-amssncku_amd/csrc`")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7.  Once
    # torch is loaded, the library's dependency on that soname binds to
    # torch's copy; loaded the other way round, torch would later bring in a
    # second runtime that finds no GPU.  So load torch first when it exists.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        f = getattr(L, name)
        if args is not None:
            f.argtypes = args
        f.restype = res
    _lib = L
    return L


def last_error() -> str:
    msg = lib().amss_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> int:
    if rc not in (0,):
        raise AmssError(f"{what} failed with code {rc}: {last_error()}")
    return rc
