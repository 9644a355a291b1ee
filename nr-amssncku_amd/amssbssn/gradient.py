"""Gradients of a scalar observable of the evolved BSSN state with respect to
initial-data parameters, through the hand-written step adjoint.

Parameters theta = (m1, m2, x1, x2, c[n^3]): the two puncture masses, their
x positions, and (n > 0) the coefficients of a cubic B-spline field added to
A_xx.  The forward run evolves the level's fused RK4 step (one patch, or a
batched level of uneven patches) and keeps a checkpoint every `checkpoint`
steps; the backward pass rebuilds each segment from its checkpoint and runs
the level's step adjoint (libamssbssn.so) backwards through it, exchange
included.  EvolveObservable is the torch.autograd.Function; the initial-data
map and the observable are small elementwise torch expressions on the
device, so torch's autograd chains dJ/dU0 to dJ/dtheta.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import time
from dataclasses import dataclass

import numpy as np

from .grid import AXX, CHI, LAP, NF, PatchGeom, cube_patch, philox_uniform, uneven_level
from .level import LIST_STATE, Level
from .adjoint import LevelAdjoint, _ptrs
from ._lib import check, lib


@dataclass
class DiffConfig:
    """Patches with frozen outer ghost zones, two punctures, A_ij noise, an
    optional spline perturbation of A_xx; the observable is the
    Gaussian-shell integral of alpha * chi over the owned points."""
    patches: list
    dt: float
    steps: int = 200
    checkpoint: int = 20
    eps: float = 0.02
    chitiny: float = 1e-5
    masses: tuple = (0.5, 0.5)
    xs: tuple = (3.0, -3.0)
    spline_n: int = 16           # 0: no spline parameters
    spline_sigma: float = 1e-6
    spline_seed: int = 3
    a_noise: float = 1e-8
    noise_seed: int = 0
    shift0: tuple = (0.0, 0.0, 0.0)  # constant initial shift (0 in the benchmark configurations)
    shell_r: float = 6.0
    shell_width: float = 1.0

    def theta0(self):
        """Initial parameter vector: masses, x positions, spline coefficients
        N(0, sigma^2) from numpy's default_rng(spline_seed)."""
        rng = np.random.default_rng(self.spline_seed)
        c = self.spline_sigma * rng.standard_normal(self.spline_n ** 3) if self.spline_n else np.zeros(0)
        return np.concatenate([np.array(tuple(self.masses) + tuple(self.xs), dtype=np.float64), c])

    @property
    def interior_pts(self):
        return sum(p.interior_pts for p in self.patches)


def config3(n: int = 64, steps: int = 200, checkpoint: int = 20) -> DiffConfig:
    """One n^3 patch (+3 ghosts) over [-16,16]^3 (h = 32/n), config-1 data,
    16^3 spline coefficients on A_xx."""
    g = cube_patch(n, 16.0)
    return DiffConfig([g], dt=0.25 * g.h[0], steps=steps, checkpoint=checkpoint)


def config4(steps: int = 400, checkpoint: int = 0) -> DiffConfig:
    """16 uneven patches (multiples of 8 in [48,128], seed 2) packed in
    [-24,24]^3; masses 0.55 / 0.45 at x = +-4; gradient w.r.t. masses and
    positions.  checkpoint 0: chosen by plan_checkpoints at run time."""
    patches, _ = uneven_level(16, 48, 128, seed=2, half_width=24.0)
    return DiffConfig(patches, dt=0.25 * patches[0].h[0], steps=steps, checkpoint=checkpoint,
                      masses=(0.55, 0.45), xs=(4.0, -4.0), spline_n=0)


def plan_checkpoints(steps: int, state_bytes: int, free_bytes: int, reserve: float = 0.1,
                     segment_factor: float = 1.0) -> int:
    """Checkpoint spacing k for a run of `steps`: the backward pass holds
    ceil(steps/k) checkpoints (state_bytes each) plus what it keeps for the k
    steps of the segment it rebuilds (segment_factor x state_bytes each).
    Each segment rebuild costs k - 1 forward steps, so the recomputation is
    steps - ceil(steps/k) forward steps in total: the smallest k that fits in
    the free memory (less a reserve) recomputes least."""
    budget = free_bytes * (1.0 - reserve) / max(state_bytes, 1)
    for k in range(1, steps + 1):
        if math.ceil(steps / k) + segment_factor * k <= budget:
            return k
    raise MemoryError(f"{steps} steps of {state_bytes} B states do not fit in {free_bytes} B")


# ---------------------------------------------------------------------------
def bspline3(t):
    """Cardinal cubic B-spline (support [-2, 2])."""
    a = np.abs(t)
    return np.where(a < 1, (4.0 - 6.0 * a ** 2 + 3.0 * a ** 3) / 6.0, np.where(a < 2, (2.0 - a) ** 3 / 6.0, 0.0))


def spline_basis(coord, n, lo=None, hi=None):
    """[len(coord), n] matrix of n uniform cubic B-splines whose centres span
    [lo, hi] (default: the coordinate range) end to end."""
    lo = coord[0] if lo is None else lo
    hi = coord[-1] if hi is None else hi
    d = (hi - lo) / (n - 1)
    centres = lo + d * np.arange(n)
    return bspline3((coord[:, None] - centres[None, :]) / d)


class InitialData:
    """theta -> [U0 of each patch] as torch expressions on one device."""

    def __init__(self, cfg: DiffConfig, device="cuda"):
        import torch
        self.cfg = cfg
        t = lambda a: torch.tensor(a, dtype=torch.float64, device=device)  # noqa: E731
        # the spline spans the level's coordinate range (ghost cells included)
        lo = [min(p.coords()[a][0] for p in cfg.patches) for a in range(3)]
        hi = [max(p.coords()[a][-1] for p in cfg.patches) for a in range(3)]
        self.parts = []
        for g in cfg.patches:
            X, Y, Z = g.coords()
            z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
            part = dict(x=t(x), y=t(y), z=t(z))
            if cfg.spline_n:
                part["B"] = [t(spline_basis(c, cfg.spline_n, lo[a], hi[a])) for a, c in enumerate((X, Y, Z))]
            base = np.zeros((NF,) + x.shape)
            if cfg.a_noise:
                for s in range(6):
                    u = philox_uniform(g.npts, AXX + s, cfg.noise_seed).reshape(x.shape)
                    base[AXX + s] = cfg.a_noise * (2.0 * u - 1.0)
            for a in range(3):
                base[18 + a] += cfg.shift0[a]
            part["base"] = t(base)
            self.parts.append(part)

    def __call__(self, theta):
        import torch
        n = self.cfg.spline_n
        m1, m2, x1, x2 = theta[0], theta[1], theta[2], theta[3]
        out = []
        for P in self.parts:
            r2 = P["y"] ** 2 + P["z"] ** 2 + 0.01
            psi = (1.0 + m1 / (2.0 * torch.sqrt((P["x"] - x1) ** 2 + r2)) +
                   m2 / (2.0 * torch.sqrt((P["x"] - x2) ** 2 + r2)))
            fields = [P["base"][v] for v in range(NF)]
            fields[CHI] = psi ** -4 - 1.0
            fields[LAP] = psi ** -2 - 1.0
            if n:
                Bx, By, Bz = P["B"]
                coef = theta[4:].reshape(n, n, n)  # [c_z, c_y, c_x]
                fields[AXX] = fields[AXX] + torch.einsum("cba,kc,jb,ia->kji", coef, Bz, By, Bx)
            out.append(torch.stack(fields).contiguous())
        return out


class Observable:
    """J = sum over owned points of alpha chi w(r) dV, w = exp(-(r - r0)^2 / (2 sigma^2))."""

    def __init__(self, cfg: DiffConfig, device="cuda"):
        import torch
        self.w = []
        for g in cfg.patches:
            X, Y, Z = g.coords()
            z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
            r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
            w = np.exp(-(r - cfg.shell_r) ** 2 / (2.0 * cfg.shell_width ** 2)) * float(np.prod(g.h))
            mask = np.zeros_like(w)
            gw = g.gw
            mask[gw:-gw, gw:-gw, gw:-gw] = 1.0
            self.w.append(torch.tensor(w * mask, dtype=torch.float64, device=device))

    def value(self, Us):
        return sum(((U[LAP] + 1.0) * (U[CHI] + 1.0) * w).sum() for U, w in zip(Us, self.w))

    def grad(self, Us):
        """dJ/dU of each patch."""
        import torch
        out = []
        for U, w in zip(Us, self.w):
            lam = torch.zeros_like(U)
            lam[CHI] = (U[LAP] + 1.0) * w
            lam[LAP] = (U[CHI] + 1.0) * w
            out.append(lam)
        return out


class Evolution:
    """The forward evolution (the level's fused step) with checkpoints, and
    the backward pass through the level's step adjoint."""

    def __init__(self, cfg: DiffConfig, mode: int = 0, tape: bool = True):
        self.cfg = cfg
        self.level = Level(cfg.patches, gauge=0, eps=cfg.eps, chitiny=cfg.chitiny)
        self.adj = LevelAdjoint(cfg.patches, eps=cfg.eps, chitiny=cfg.chitiny)
        # with a tape, each segment's rebuild keeps every step's stage data and
        # the step adjoints reuse it instead of re-evaluating the step
        self.tape = tape
        if tape:
            self.adj.reserve(cfg.checkpoint)
        self.mode = mode
        if mode > 0:
            self.level.autotune()  # outside the timed forward (the first batched step would run it)
        self.timing = {}
        self.min_free = None

    def close(self):
        self.level.close()
        self.adj.close()

    def _note_memory(self):
        import torch
        free, total = torch.cuda.mem_get_info()
        self.total = total
        self.min_free = free if self.min_free is None else min(self.min_free, free)

    def _copy(self, Us, direction):
        """State of every patch in (0) or out (1), ordered on the level's stream."""
        for p, U in enumerate(Us):
            check(lib().amss_level_copy_list(self.level.h, p, LIST_STATE, _ptrs(U), direction), "amss_level_copy_list")

    def forward(self, U0s):
        """Evolve for cfg.steps steps; returns (final states, checkpoints)."""
        import torch
        cfg = self.cfg
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ckpts = [[U.detach().clone() for U in U0s]]
        torch.cuda.synchronize()  # the level's stream reads what torch's stream wrote
        self._copy(ckpts[0], 0)
        done = 0
        while done < cfg.steps:
            k = min(cfg.checkpoint, cfg.steps - done)
            rc = self.level.step(cfg.dt, k, self.mode)
            if rc:
                raise RuntimeError(f"forward evolution failed (code {rc})")
            done += k
            Us = [torch.empty_like(U) for U in U0s]
            self._copy(Us, 1)
            ckpts.append(Us)
        torch.cuda.synchronize()
        self.timing["forward_s"] = time.perf_counter() - t0
        self._note_memory()
        return ckpts[-1], ckpts[:-1]

    def backward(self, lams, ckpts):
        """dJ/dU(T) of each patch -> dJ/dU0, rebuilding each segment from its checkpoint."""
        import torch
        cfg = self.cfg
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lams = [l.clone().contiguous() for l in lams]
        for s in range(len(ckpts) - 1, -1, -1):
            k = min(cfg.checkpoint, cfg.steps - s * cfg.checkpoint)
            states = [ckpts[s]]
            for n in range(k if self.tape else k - 1):
                nxt = [torch.empty_like(U) for U in states[-1]]
                self.adj.forward(cfg.dt, states[-1], nxt, n if self.tape else -1)
                states.append(nxt)
            if s == len(ckpts) - 1:
                self._note_memory()
            for n in range(k - 1, -1, -1):
                self.adj.step(cfg.dt, states[n], lams, n if self.tape else -1)
        torch.cuda.synchronize()
        self.timing["backward_s"] = time.perf_counter() - t0
        return lams


def _autograd_function():
    import torch

    class EvolveObservable(torch.autograd.Function):
        """J(U0 of every patch) = observable(evolve(U0)); backward through the step adjoint."""

        @staticmethod
        def forward(ctx, evo, obs, *U0s):
            UT, ckpts = evo.forward(list(U0s))
            ctx.evo, ctx.ckpts, ctx.lamT = evo, ckpts, obs.grad(UT)
            return obs.value(UT)

        @staticmethod
        def backward(ctx, gJ):
            lam0 = ctx.evo.backward([l * gJ for l in ctx.lamT], ctx.ckpts)
            ctx.ckpts = None
            return (None, None) + tuple(lam0)

    return EvolveObservable


_EVOLVE = None


def evolve_observable(U0s, evo: Evolution, obs: Observable):
    """J as a differentiable function of the initial states (device tensors)."""
    global _EVOLVE
    if _EVOLVE is None:
        _EVOLVE = _autograd_function()
    return _EVOLVE.apply(evo, obs, *U0s)


TAPE_FIELDS = 192  # stage data of one step (Y0..Y3, z0..z3), grid functions per patch


def _plan(cfg: DiffConfig, device):
    """(config with its checkpoint spacing, whether the segment tape fits)."""
    import torch
    free, _ = torch.cuda.mem_get_info()
    state = sum(NF * 8 * p.npts for p in cfg.patches)
    # the level (127 grid functions) and the adjoint work arrays (~520) sit
    # next to the checkpoints
    work = sum((127 + 520) * 8 * p.npts for p in cfg.patches)
    factor = 1.0 + TAPE_FIELDS / NF  # a segment step keeps its state and its stage data
    if cfg.checkpoint > 0:
        k = cfg.checkpoint
        fits = math.ceil(cfg.steps / k) + factor * k <= 0.9 * (free - work) / state
        return cfg, fits
    try:
        return dataclasses.replace(cfg, checkpoint=plan_checkpoints(cfg.steps, state, free - work,
                                                                    segment_factor=factor)), True
    except MemoryError:
        return dataclasses.replace(cfg, checkpoint=plan_checkpoints(cfg.steps, state, free - work)), False


def value_and_grad(cfg: DiffConfig, theta=None, device="cuda", mode: int = 0):
    """(J, dJ/dtheta, info) for the configuration's parameters; info holds the
    forward / backward wall times, the checkpoint spacing and peak device memory."""
    import torch
    cfg, tape = _plan(cfg, device)
    init = InitialData(cfg, device)
    obs = Observable(cfg, device)
    evo = Evolution(cfg, mode, tape=tape)
    th = torch.tensor(cfg.theta0() if theta is None else theta, dtype=torch.float64, device=device,
                      requires_grad=True)
    try:
        J = evolve_observable(init(th), evo, obs)
        J.backward()
        info = dict(evo.timing, checkpoint=cfg.checkpoint, tape=tape,
                    peak_device_bytes=int(evo.total - evo.min_free) if evo.min_free is not None else None)
        return float(J.detach()), th.grad.detach().cpu().numpy(), info
    finally:
        evo.close()


def value(cfg: DiffConfig, theta, device="cuda", mode: int = 0) -> float:
    """J alone (no tape), for finite differences."""
    import torch
    cfg = dataclasses.replace(cfg, checkpoint=cfg.checkpoint or cfg.steps)
    init = InitialData(cfg, device)
    obs = Observable(cfg, device)
    evo = Evolution(cfg, mode, tape=False)
    try:
        with torch.no_grad():
            UT, _ = evo.forward(init(torch.tensor(theta, dtype=torch.float64, device=device)))
            return float(obs.value(UT))
    finally:
        evo.close()


__all__ = ["DiffConfig", "config3", "config4", "plan_checkpoints", "InitialData", "Observable", "Evolution",
           "evolve_observable", "value_and_grad", "value", "spline_basis", "bspline3", "PatchGeom"]
