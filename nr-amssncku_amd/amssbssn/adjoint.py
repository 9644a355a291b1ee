"""Discrete adjoint of the RK4 step on one Block (the HIP adjoint kernels
behind amss_adjoint_*), on torch device tensors of shape [24, nz, ny, nx].

torch is the device-memory and stream plumbing here; every operation on the
state runs in libamssbssn.so.
"""
from __future__ import annotations

import ctypes as C

from ._lib import AmssError, check, lib

NF = 24


def _ptrs(t):
    """24 device pointers of a [24, nz, ny, nx] contiguous float64 tensor."""
    import torch
    if t.dtype != torch.float64 or not t.is_contiguous() or t.dim() != 4 or t.shape[0] != NF:
        raise ValueError("expected a contiguous float64 tensor of shape [24, nz, ny, nx]")
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    base, step = t.data_ptr(), t[0].numel() * 8
    return (C.c_void_p * NF)(*[base + v * step for v in range(NF)])


def _level_ptrs(ts):
    """npatch x 24 device pointers (patch-major) of a list of per-patch tensors."""
    out = []
    for t in ts:
        out.extend(_ptrs(t))
    return (C.c_void_p * len(out))(*out)


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class StepAdjoint:
    """Forward step and its adjoint for one Block of shape (nx, ny, nz) (ghosts
    included, width 3, frozen), grid spacing h = (hx, hy, hz), GAUGE 0."""

    def __init__(self, shape, h, eps: float = 0.02, chitiny: float = 1e-5, gauge: int = 0):
        ex = (C.c_int * 3)(*shape)
        hh = (C.c_double * 3)(*h)
        self.shape = tuple(shape)
        self.h = lib().amss_adjoint_create(ex, hh, eps, chitiny, gauge)
        if not self.h:
            raise AmssError("amss_adjoint_create: " + (lib().amss_last_error() or b"").decode())

    def close(self):
        if getattr(self, "h", None):
            lib().amss_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, dt: float, U, Unext) -> None:
        """Unext <- one RK4 step of U (stream-ordered)."""
        check(lib().amss_adjoint_forward(self.h, dt, _ptrs(U), _ptrs(Unext), _stream()), "amss_adjoint_forward")

    def step(self, dt: float, U, lam) -> None:
        """lam: dJ/dU^{n+1} in, dJ/dU^n out; U: the state at t^n."""
        check(lib().amss_adjoint_step(self.h, dt, _ptrs(U), _ptrs(lam), _stream()), "amss_adjoint_step")

    def rhs_vjp(self, U, lam, out) -> None:
        """out <- dJ/dU of sum_p lam(p).rhs(U)(p) over the interior points."""
        check(lib().amss_adjoint_rhs(self.h, _ptrs(U), _ptrs(lam), _ptrs(out), _stream()), "amss_adjoint_rhs")


class LevelAdjoint:
    """Forward step and its adjoint for a level of patches (PatchGeom list,
    ghost width 3, frozen ghosts, GAUGE 0), including the reverse of the
    inter-patch exchange.  States are lists of per-patch [24, nz, ny, nx]
    device tensors."""

    def __init__(self, patches, eps: float = 0.02, chitiny: float = 1e-5, gauge: int = 0):
        shapes = (C.c_int * (3 * len(patches)))(*[s for p in patches for s in p.shape])
        bbs = (C.c_double * (6 * len(patches)))(*[b for p in patches for b in p.bbox])
        self.h = lib().amss_level_adjoint_create(len(patches), shapes, bbs, eps, chitiny, gauge)
        if not self.h:
            raise AmssError("amss_level_adjoint_create: " + (lib().amss_last_error() or b"").decode())
        self.npatch = len(patches)

    def close(self):
        if getattr(self, "h", None):
            lib().amss_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bytes(self) -> int:
        return int(lib().amss_adjoint_bytes(self.h))

    def reserve(self, slots: int) -> None:
        """A tape of `slots` steps' stage data (see forward / step)."""
        check(lib().amss_adjoint_reserve(self.h, slots), "amss_adjoint_reserve")

    def forward(self, dt: float, U, Unext, slot: int = -1) -> None:
        """Unext <- one step of U; slot >= 0 keeps the step's stage data in that tape slot."""
        check(lib().amss_adjoint_forward_slot(self.h, dt, _level_ptrs(U), _level_ptrs(Unext), slot, _stream()),
              "amss_adjoint_forward_slot")

    def step(self, dt: float, U, lam, slot: int = -1) -> None:
        """lam: dJ/dU^{n+1} in, dJ/dU^n out; slot >= 0 reuses that slot's stage data
        (which forward(U, ..., slot) must have written), -1 re-evaluates the step."""
        check(lib().amss_adjoint_step_slot(self.h, dt, _level_ptrs(U), _level_ptrs(lam), slot, _stream()),
              "amss_adjoint_step_slot")
