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

    def _rhs(self, src, dst, co, res=None):
        def job(ab):
            k0, k1 = ab
            s0, s1 = k0 - 3, k1 + 3
            st = [np.ascontiguousarray(f[s0:s1]) for f in src]
            o = self.be.compute_rhs(st, self.X, self.Y, np.ascontiguousarray(self.Z[s0:s1]), self.eps, co=co,
                                    gauge=self.gauge)
            for v in range(NF):
                dst[v][k0:k1] = o["rhs"][v][3:3 + (k1 - k0)]
            if res is not None:
                for r in range(7):
                    res[r][k0:k1] = o["res"][r][3:3 + (k1 - k0)]
            return o["gont"]
        return max(self.pool.map(job, self.owned))

    def residual_norms(self, U):
        """The 7 constraint L2 norms (Ham, Px..Pz, Gx..Gz) of the predictor of
        a step starting from U (residuals of the enforced U, co = 0) over the
        owned points [gw, n - gw)^3 (frozen ghost zones)."""
        shape = U[0].shape
        V = [u.copy() for u in U]
        self._planes(lambda ab: _enforce_planes(self.be, V, ab))
        k = [np.zeros(shape) for _ in range(NF)]
        res = [np.zeros(shape) for _ in range(7)]
        self._rhs(V, k, 0, res)
        gw = self.gw
        dv = (self.X[1] - self.X[0]) * (self.Y[1] - self.Y[0]) * (self.Z[1] - self.Z[0])
        s = np.s_[gw:shape[0] - gw, gw:shape[1] - gw, gw:shape[2] - gw]
        return [float(np.sqrt(np.sum(r[s] ** 2) * dv)) for r in res]

    def _planes(self, fn):
        list(self.pool.map(fn, self.planes))

    def step(self, U):
        """One RK4 step in place on U (frozen-ghost boundary)."""
        shape = U[0].shape
        pre = [np.zeros(shape) for _ in range(NF)]
        cor = [np.zeros(shape) for _ in range(NF)]
        acc = [np.zeros(shape) for _ in range(NF)]
        gw, dt, tinny = self.gw, self.dt, self.tinny

        def enforce(F):
            self._planes(lambda ab: _enforce_planes(self.be, F, ab))

        def rk_bc(out, k, stage):
            # RK4 update, then the frozen-ghost rule (every point within gw of
            # a face keeps U0) and the chi bound, plane range by plane range
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
                    for z in range(a, b):  # the ghost shell of each plane
                        o, u = out[v][z], U[v][z]
                        if z < gw or z >= shape[0] - gw:
                            o[...] = u
                        else:
                            o[:gw] = u[:gw]
                            o[shape[1] - gw:] = u[shape[1] - gw:]
                            o[:, :gw] = u[:, :gw]
                            o[:, shape[2] - gw:] = u[:, shape[2] - gw:]
                self.be.lowerboundset(out[0][a:b], tinny)
            self._planes(job)

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


def time_reference_steps(cfg, nsteps=5, nthreads=None, prefer_reference=True):
    """Median seconds per RK4 step of the CPU baseline on cfg's first patch
    (after one warm-up step), what ran, the threads used and every step's
    time."""
    if prefer_reference and O.have_ref():
        be, kind, gauge = O.Ref("_o3") if os.path.exists(os.path.join(O.HERE, "_ref", "libamssref_o3.so")) \
            else O.Ref(), "reference", None
    else:
        be, kind, gauge = O.Orc(), "port", 0
    geom = cfg.patches[0]
    stp = SlabStepper(be, geom, cfg.dt, cfg.eps, cfg.chitiny, gauge=gauge, nthreads=nthreads, gw=geom.gw)
    U = cfg.initial_state(0)
    stp.step(U)  # warm-up
    times = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        stp.step(U)
        times.append(time.perf_counter() - t0)
    stp.close()
    return float(np.median(times)), kind, stp.nthreads, times
