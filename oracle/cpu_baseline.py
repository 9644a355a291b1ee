"""CPU baseline of the BSSN RK4 step: the reference Fortran (or, where the
reference build is absent, the C restatement) driven in bssn_class::Step
order (bssn_class.C:2981-3901) on all host cores.

TEST INFRASTRUCTURE ONLY: used by the benchmark's cpu_baseline leg.

Parallelism is the reference's own block decomposition
(Parallel.C:373-396): the z range is cut into slabs, each extended by three
ghost planes, and every thread runs compute_rhs_bssn_ on its slab; only the
owned planes of the result are kept, so the owned points are exactly the
whole-array result.  Pointwise routines (enforce_ga_, rungekutta4_rout_,
lowerboundset_) run on disjoint plane ranges.  ctypes releases the GIL, so
threads run the Fortran concurrently.  The reference keeps its ~75
temporaries as automatic (stack) arrays, hence the large thread stacks.
"""
from __future__ import annotations

import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import oracle as O

NF = 24


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


class SlabStepper:
    def __init__(self, backend, geom, dt, eps, chitiny=1e-5, gauge=None, nthreads=None, gw=3):
        self.be = backend
        self.gauge = gauge
        self.X, self.Y, self.Z = geom.coords()
        self.dt, self.eps, self.tinny, self.gw = dt, eps, chitiny - 1.0, gw
        self.nthreads = nthreads or host_cores()
        nz = geom.shape[2]
        cuts = np.linspace(gw, nz - gw, self.nthreads + 1).round().astype(int)
        self.owned = [(a, b) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]
        pcuts = np.linspace(0, nz, self.nthreads + 1).round().astype(int)
        self.planes = [(a, b) for a, b in zip(pcuts[:-1], pcuts[1:]) if b > a]
        threading.stack_size(512 * 1024 * 1024)
        self.pool = ThreadPoolExecutor(self.nthreads)

    def close(self):
        self.pool.shutdown()

    def _rhs(self, src, dst, co):
        def job(ab):
            k0, k1 = ab
            s0, s1 = k0 - 3, k1 + 3
            st = [np.ascontiguousarray(f[s0:s1]) for f in src]
            o = self.be.compute_rhs(st, self.X, self.Y, np.ascontiguousarray(self.Z[s0:s1]), self.eps, co=co,
                                    gauge=self.gauge)
            for v in range(NF):
                dst[v][k0:k1] = o["rhs"][v][3:3 + (k1 - k0)]
            return o["gont"]
        return max(self.pool.map(job, self.owned))

    def _planes(self, fn):
        list(self.pool.map(fn, self.planes))

    def step(self, U):
        """One RK4 step in place on U (frozen-ghost boundary)."""
        shape = U[0].shape
        pre = [np.zeros(shape) for _ in range(NF)]
        cor = [np.zeros(shape) for _ in range(NF)]
        acc = [np.zeros(shape) for _ in range(NF)]
        gw, dt, tinny = self.gw, self.dt, self.tinny
        interior = (slice(gw, shape[0] - gw), slice(gw, shape[1] - gw), slice(gw, shape[2] - gw))

        def enforce(F):
            self._planes(lambda ab: _enforce_planes(self.be, F, ab))

        def rk_bc(out, k, stage):
            def job(ab):
                a, b = ab
                for v in range(NF):
                    f0 = U[v][a:b]
                    f1 = out[v][a:b] if stage == 0 else k[v][a:b]
                    r = acc[v][a:b]
                    if stage == 0:
                        np.copyto(r, k[v][a:b])
                    self.be.rk4(f0, f1, r, dt, stage)
                    if stage > 0:
                        out[v][a:b] = f1
            self._planes(job)
            # frozen ghost zones (every point within gw of a face keeps U0)
            for v in range(NF):
                keep = out[v][interior].copy()
                out[v][...] = U[v]
                out[v][interior] = keep
            self.be.lowerboundset(out[0], tinny)

        enforce(U)
        k = [np.zeros(shape) for _ in range(NF)]
        gont = self._rhs(U, k, 0)
        rk_bc(pre, k, 0)
        for it in (1, 2, 3):
            enforce(pre)
            gont |= self._rhs(pre, cor, 1)
            rk_bc(cor, cor, it)
            if it < 3:
                pre, cor = cor, pre
        for v in range(NF):
            np.copyto(U[v], cor[v])
        return gont


def _enforce_planes(be, F, ab):
    a, b = ab
    views = [f[a:b] for f in F]
    be.enforce_ga(views)


def time_reference_steps(cfg, nsteps=2, nthreads=None, prefer_reference=True):
    """Seconds per RK4 step of the CPU baseline on cfg's first patch, and what ran."""
    if prefer_reference and O.have_ref():
        be, kind, gauge = O.Ref("_o3") if os.path.exists(os.path.join(O.HERE, "_ref", "libamssref_o3.so")) \
            else O.Ref(), "reference", None
    else:
        be, kind, gauge = O.Orc(), "port", 0
    geom = cfg.patches[0]
    stp = SlabStepper(be, geom, cfg.dt, cfg.eps, cfg.chitiny, gauge=gauge, nthreads=nthreads, gw=geom.gw)
    U = cfg.initial_state(0)
    stp.step(U)  # warm-up
    t0 = time.perf_counter()
    for _ in range(nsteps):
        stp.step(U)
    dt = (time.perf_counter() - t0) / nsteps
    stp.close()
    return dt, kind, stp.nthreads
