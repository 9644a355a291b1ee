"""Python driver of a hierarchy of refinement levels on the GPU (the C++
AmrHierarchy behind amss_amr_*): Berger-Oliger sub-cycling in the order of
bssn_class::RecursiveStep / RestrictProlong (bssn_class.C:2323-2403,
5724-5904).  Level 0 is an ordinary level; a finer level halves the spacing
and its patch arrays carry the reference's 6-cell buffer zone."""
from __future__ import annotations

import ctypes as C

from ._lib import AmssError, check, lib
from .grid import PatchGeom
from .level import BC_FROZEN_GHOSTS, Level

BUFFER_WIDTH = 6  # the reference's buffer_width (macrodef.h:88)


class _LevelView(Level):
    """A level owned by a hierarchy (closing it does not free it)."""

    def __init__(self, handle, patches):  # noqa: super().__init__ would create a level
        self.h = handle
        self.patches = list(patches)

    def close(self):
        self.h = None


class Hierarchy:
    def __init__(self, levels, gauge: int = 0, eps: float = 0.02, chitiny: float = 1e-5,
                 bc0: int = BC_FROZEN_GHOSTS, symmetry: int = 0):
        """levels: a list (coarse first) of lists of PatchGeom."""
        L = lib()
        allp = [p for lev in levels for p in lev]
        npatch = (C.c_int * len(levels))(*[len(lev) for lev in levels])
        shapes = (C.c_int * (3 * len(allp)))(*[s for p in allp for s in p.shape])
        bbs = (C.c_double * (6 * len(allp)))(*[b for p in allp for b in p.bbox])
        self.h = L.amss_amr_create(len(levels), npatch, shapes, bbs, gauge, eps, chitiny, bc0, symmetry)
        if not self.h:
            raise AmssError("amss_amr_create: " + (L.amss_last_error() or b"").decode())
        self.levels = [_LevelView(L.amss_amr_level(self.h, i), lev) for i, lev in enumerate(levels)]

    def level(self, i: int) -> Level:
        return self.levels[i]

    def step(self, dt: float, nsteps: int = 1, mode: int = 0, want_ham=False) -> int:
        rc = lib().amss_amr_step(self.h, dt, nsteps, mode, int(want_ham))
        if rc not in (0, 1):
            check(rc, "amss_amr_step")
        return rc

    @property
    def time(self) -> float:
        return lib().amss_amr_time(self.h)

    def checkpoint_write(self, path: str, movls: int = 0, porg=(), handles=None) -> None:
        """Every level's StateList and OldStateList in checkpoint::writecheck_cgh's
        layout; porg: per level a list of puncture positions (the same count on
        every level), handles: 3 per patch, level after level."""
        nbh = len(porg[0]) if porg else 0
        assert not porg or (len(porg) == len(self.levels) and all(len(p) == nbh for p in porg))
        pts = [x for lev in porg for p in lev for x in p]
        pa = (C.c_double * max(len(pts), 1))(*pts)
        hd = None
        if handles is not None:
            hl = [x for p in handles for x in p]
            hd = (C.c_double * len(hl))(*hl)
        check(lib().amss_amr_checkpoint_write(self.h, path.encode(), movls, nbh, pa, hd),
              "amss_amr_checkpoint_write")

    def checkpoint_read(self, path: str, maxbh: int = 16):
        """Restore both lists of every level and the times; returns (time,
        movls, porg per level)."""
        mv, nb = C.c_int(0), C.c_int(0)
        nl = len(self.levels)
        pa = (C.c_double * (3 * max(maxbh, 1) * nl))()
        check(lib().amss_amr_checkpoint_read(self.h, path.encode(), C.byref(mv), C.byref(nb), pa, maxbh),
              "amss_amr_checkpoint_read")
        n = nb.value
        porg = [[tuple(pa[3 * (l * n + b):3 * (l * n + b) + 3]) for b in range(n)] for l in range(nl)]
        return self.time, mv.value, porg

    def close(self):
        if getattr(self, "h", None):
            for lv in self.levels:
                lv.close()
            lib().amss_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nested_box(coarse: PatchGeom, lo_cell, hi_cell, gw: int = 3) -> PatchGeom:
    """The fine patch (half the spacing) that refines the coarse interior
    cells [lo_cell, hi_cell) per axis (counted from the first interior cell),
    with the 6-cell buffer zone on every face."""
    h = coarse.h
    shape, bb_lo, bb_hi = [], [], []
    for a in range(3):
        x0 = coarse.bbox[a] + (gw + lo_cell[a]) * h[a]
        x1 = coarse.bbox[a] + (gw + hi_cell[a]) * h[a]
        hf = h[a] / 2
        bb_lo.append(x0 - BUFFER_WIDTH * hf)
        bb_hi.append(x1 + BUFFER_WIDTH * hf)
        shape.append(2 * (hi_cell[a] - lo_cell[a]) + 2 * BUFFER_WIDTH)
    return PatchGeom(tuple(shape), tuple(bb_lo) + tuple(bb_hi), gw)
