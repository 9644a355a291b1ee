"""Python driver of one refinement level on the GPU (the C++ BssnLevel behind
the C ABI).  Mirrors how bssn_class drives Step for a level: the state lives
in HBM in each Block's grid functions; Python only moves initial data in and
results out.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, AmssError
from .grid import NF, Config

BC_FROZEN_GHOSTS, BC_COPY_LAYER, BC_NONE, BC_SOMMERFELD = 0, 1, 2, 3
LIST_STATE, LIST_OLD, LIST_PRE, LIST_COR, LIST_RHS, LIST_RES = range(6)
MODE_PLAIN, MODE_BATCHED, MODE_GRAPH = 0, 1, 2
RES_NONE, RES_LAST_STEP, RES_EVERY_STEP = 0, 1, 2


class Level:
    def __init__(self, patches, gauge: int = 0, eps: float = 0.02, chitiny: float = 1e-5,
                 bc: int = BC_FROZEN_GHOSTS, gw: int = 3, symmetry: int = 0):
        L = lib()
        self.patches = list(patches)
        shapes = (C.c_int * (3 * len(patches)))(*[s for p in patches for s in p.shape])
        bbs = (C.c_double * (6 * len(patches)))(*[b for p in patches for b in p.bbox])
        self.h = L.amss_level_create(len(patches), shapes, bbs, gauge, eps, chitiny, bc, gw, symmetry)
        if not self.h:
            raise AmssError("amss_level_create: " + (L.amss_last_error() or b"").decode())

    def close(self):
        if getattr(self, "h", None):
            lib().amss_level_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @classmethod
    def from_config(cls, cfg: Config, gauge: int = 0, bc: int = BC_FROZEN_GHOSTS, symmetry: int = 0):
        lv = cls(cfg.patches, gauge=gauge, eps=cfg.eps, chitiny=cfg.chitiny, bc=bc, gw=cfg.patches[0].gw,
                 symmetry=symmetry)
        for i in range(len(cfg.patches)):
            lv.set_state(i, cfg.initial_state(i))
        return lv

    # -- data movement -------------------------------------------------------
    def set_field(self, patch: int, list_: int, v: int, arr) -> None:
        if hasattr(arr, "data_ptr"):  # a torch tensor (host or device)
            ptr = C.c_void_p(arr.data_ptr())
        else:
            arr = np.ascontiguousarray(arr, dtype=np.float64)
            ptr = arr.ctypes.data_as(C.c_void_p)
        check(lib().amss_level_copy(self.h, patch, list_, v, ptr, 0), "amss_level_copy(in)")

    def get_field(self, patch: int, list_: int, v: int) -> np.ndarray:
        shp = self.patches[patch].shape
        out = np.empty((shp[2], shp[1], shp[0]))
        check(lib().amss_level_copy(self.h, patch, list_, v, out.ctypes.data_as(C.c_void_p), 1),
              "amss_level_copy(out)")
        return out

    def set_state(self, patch: int, st) -> None:
        for v in range(NF):
            self.set_field(patch, LIST_STATE, v, st[v])

    def get_state(self, patch: int):
        return [self.get_field(patch, LIST_STATE, v) for v in range(NF)]

    def field_ptr(self, patch: int, list_: int, v: int) -> int:
        return lib().amss_level_field(self.h, patch, list_, v)

    # -- evolution -------------------------------------------------------------
    def step(self, dt: float, nsteps: int = 1, mode: int = MODE_PLAIN, want_ham=False) -> int:
        """want_ham: False/0 no residuals; True/1 residuals in the last step's
        predictor; RES_EVERY_STEP (2) in every predictor, as the reference's
        Step does, with the seven constraint norms recorded per step."""
        rc = lib().amss_level_step(self.h, dt, nsteps, mode, int(want_ham))
        if rc not in (0, 1):
            check(rc, "amss_level_step")
        return rc

    def ham_l2(self) -> float:
        out = C.c_double(0.0)
        check(lib().amss_level_ham_l2(self.h, C.byref(out)), "amss_level_ham_l2")
        return out.value

    def constraint_norms(self, patch: int = -1) -> np.ndarray:
        """The 7 L2 norms (Ham, Px, Py, Pz, Gx, Gy, Gz) of the last residuals."""
        out = (C.c_double * 7)()
        check(lib().amss_level_constraint_norms(self.h, patch, out), "amss_level_constraint_norms")
        return np.array(out[:])

    def constraint_series(self, patch: int = -1) -> np.ndarray:
        """(steps, 7) norms recorded by the last step() with RES_EVERY_STEP."""
        n = C.c_int(0)
        check(lib().amss_level_constraint_series(self.h, patch, None, 0, C.byref(n)), "amss_level_constraint_series")
        out = (C.c_double * max(7 * n.value, 1))()
        check(lib().amss_level_constraint_series(self.h, patch, out, n.value, C.byref(n)),
              "amss_level_constraint_series")
        return np.array(out[:7 * n.value]).reshape(n.value, 7)

    def stream(self) -> int:
        return lib().amss_level_stream(self.h)

    # -- the reference's on-disk formats ---------------------------------------
    def write_bin(self, path: str, patch: int, list_: int, v: int, time: float) -> None:
        """One grid function in Parallel::writefile's binary layout."""
        check(lib().amss_level_write_bin(self.h, path.encode(), patch, list_, v, time), "amss_level_write_bin")

    def checkpoint_write(self, path: str, time: float, movls: int = 0, porg=(), handles=None) -> None:
        """StateList and OldStateList in checkpoint::writecheck_cgh's layout
        (one level); porg: puncture positions, handles: 3 per patch."""
        pts = [x for p in porg for x in p]
        pa = (C.c_double * max(len(pts), 1))(*pts)
        hd = None
        if handles is not None:
            hl = [x for p in handles for x in p]
            hd = (C.c_double * len(hl))(*hl)
        check(lib().amss_level_checkpoint_write(self.h, path.encode(), time, movls, len(porg), pa, hd),
              "amss_level_checkpoint_write")

    def checkpoint_read(self, path: str, maxbh: int = 16):
        """Restore StateList and OldStateList; returns (time, movls, porg)."""
        t, mv, nb = C.c_double(0.0), C.c_int(0), C.c_int(0)
        pa = (C.c_double * (3 * max(maxbh, 1)))()
        check(lib().amss_level_checkpoint_read(self.h, path.encode(), C.byref(t), C.byref(mv), C.byref(nb), pa,
                                               maxbh), "amss_level_checkpoint_read")
        return t.value, mv.value, [tuple(pa[3 * b:3 * b + 3]) for b in range(nb.value)]

    def set_variant(self, variant: int) -> None:
        """Interior kernel: 0 direct L1/L2 stencil reads, 1 LDS-staged tiles."""
        check(lib().amss_level_set_variant(self.h, variant), "amss_level_set_variant")

    def variant(self) -> int:
        return lib().amss_level_variant(self.h)

    TILE_SHAPES = ((8, 8, 4), (4, 8, 8), (8, 4, 8), (8, 8, 8))

    def autotune(self):
        """Time every batched-kernel candidate (tile shape x persistent grid x
        tile order) on the current state, keep the fastest.  Returns
        (choice dict, list of (shape, grid, remap, ms))."""
        ns = len(self.TILE_SHAPES)
        n = ns * 3 * 2
        ms = (C.c_double * n)()
        ch = (C.c_int * 3)()
        check(lib().amss_level_autotune(self.h, ms, n, ch), "amss_level_autotune")
        cands = [(s, g, r) for s in range(ns) for g in range(3) for r in range(2)]
        return self._tune_dict(ch), [(s, g, r, float(t)) for (s, g, r), t in zip(cands, ms)]

    def set_tune(self, shape: int, grid: int, remap: int) -> None:
        check(lib().amss_level_set_tune(self.h, shape, grid, remap), "amss_level_set_tune")

    def get_tune(self):
        ch = (C.c_int * 3)()
        check(lib().amss_level_get_tune(self.h, ch), "amss_level_get_tune")
        return self._tune_dict(ch)

    def _tune_dict(self, ch):
        return dict(tile="x".join(map(str, self.TILE_SHAPES[ch[0]])), shape=int(ch[0]), grid=int(ch[1]),
                    xcd_order=bool(ch[2]))

    def set_timing(self, on: bool) -> None:
        check(lib().amss_level_timing(self.h, 1 if on else 0), "amss_level_timing")

    def kernel_ms(self):
        """(summed ms, launch count) per RK4 sub-step of the fused interior kernel."""
        ms = (C.c_double * 4)()
        n = (C.c_long * 4)()
        check(lib().amss_level_kernel_ms(self.h, ms, n), "amss_level_kernel_ms")
        return list(ms), list(n)
