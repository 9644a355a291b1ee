"""Gradients of a scalar observable of the evolved BSSN state with respect to
initial-data parameters, through the hand-written step adjoint.

Parameters theta = (m1, m2, x1, x2, c[16^3]): the two puncture masses, their
x positions, and the coefficients of a cubic B-spline field added to A_xx.
The forward run evolves the level's fused RK4 step and keeps a checkpoint
every `checkpoint` steps; the backward pass recomputes each segment from its
checkpoint and runs the step adjoint (libamssbssn.so) backwards through it.
EvolveObservable is the torch.autograd.Function; the initial-data map and the
observable are small elementwise torch expressions on the device, so torch's
autograd chains dJ/dU0 to dJ/dtheta.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from .grid import AXX, CHI, LAP, NF, PatchGeom, cube_patch, philox_uniform
from .level import LIST_STATE, Level
from .adjoint import StepAdjoint
from ._lib import check, lib


@dataclass
class DiffConfig:
    """One patch with frozen ghost zones, two punctures, A_ij noise, and a
    spline perturbation of A_xx; the observable is the Gaussian-shell
    integral of alpha * chi."""
    patch: PatchGeom
    dt: float
    steps: int = 200
    checkpoint: int = 20
    eps: float = 0.02
    chitiny: float = 1e-5
    masses: tuple = (0.5, 0.5)
    xs: tuple = (3.0, -3.0)
    spline_n: int = 16
    spline_sigma: float = 1e-6
    spline_seed: int = 3
    a_noise: float = 1e-8
    noise_seed: int = 0
    shift0: tuple = (0.0, 0.0, 0.0)  # constant initial shift (0 in the benchmark)
    shell_r: float = 6.0
    shell_width: float = 1.0

    def theta0(self):
        """Initial parameter vector: masses, x positions, spline coefficients
        N(0, sigma^2) from numpy's default_rng(spline_seed)."""
        rng = np.random.default_rng(self.spline_seed)
        c = self.spline_sigma * rng.standard_normal(self.spline_n ** 3)
        return np.concatenate([np.array(self.masses + self.xs, dtype=np.float64), c])


def config3(n: int = 64, steps: int = 200, checkpoint: int = 20) -> DiffConfig:
    """One n^3 patch (+3 ghosts) over [-16,16]^3 (h = 32/n), config-1 data."""
    g = cube_patch(n, 16.0)
    return DiffConfig(g, dt=0.25 * g.h[0], steps=steps, checkpoint=checkpoint)


# ---------------------------------------------------------------------------
def bspline3(t):
    """Cardinal cubic B-spline (support [-2, 2])."""
    a = np.abs(t)
    return np.where(a < 1, (4.0 - 6.0 * a ** 2 + 3.0 * a ** 3) / 6.0, np.where(a < 2, (2.0 - a) ** 3 / 6.0, 0.0))


def spline_basis(coord, n):
    """[len(coord), n] matrix of n uniform cubic B-splines whose centres span
    the coordinate range end to end."""
    lo, hi = coord[0], coord[-1]
    d = (hi - lo) / (n - 1)
    centres = lo + d * np.arange(n)
    return bspline3((coord[:, None] - centres[None, :]) / d)


class InitialData:
    """theta -> U0 as torch expressions on one device (differentiable)."""

    def __init__(self, cfg: DiffConfig, device="cuda"):
        import torch
        self.cfg = cfg
        g = cfg.patch
        X, Y, Z = g.coords()
        z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
        t = lambda a: torch.tensor(a, dtype=torch.float64, device=device)  # noqa: E731
        self.x, self.y, self.z = t(x), t(y), t(z)
        self.Bx = t(spline_basis(X, cfg.spline_n))
        self.By = t(spline_basis(Y, cfg.spline_n))
        self.Bz = t(spline_basis(Z, cfg.spline_n))
        base = np.zeros((NF,) + x.shape)
        if cfg.a_noise:
            for s in range(6):
                u = philox_uniform(g.npts, AXX + s, cfg.noise_seed).reshape(x.shape)
                base[AXX + s] = cfg.a_noise * (2.0 * u - 1.0)
        self.base = t(base)
        self.device = device

    def __call__(self, theta):
        import torch
        n = self.cfg.spline_n
        m1, m2, x1, x2 = theta[0], theta[1], theta[2], theta[3]
        r2 = self.y ** 2 + self.z ** 2 + 0.01
        psi = 1.0 + m1 / (2.0 * torch.sqrt((self.x - x1) ** 2 + r2)) + m2 / (2.0 * torch.sqrt((self.x - x2) ** 2 + r2))
        coef = theta[4:].reshape(n, n, n)  # [c_z, c_y, c_x]
        axx = torch.einsum("cba,kc,jb,ia->kji", coef, self.Bz, self.By, self.Bx)
        fields = [self.base[v] for v in range(NF)]
        fields[CHI] = psi ** -4 - 1.0
        fields[LAP] = psi ** -2 - 1.0
        fields[AXX] = fields[AXX] + axx
        for a in range(3):
            if self.cfg.shift0[a]:
                fields[18 + a] = fields[18 + a] + self.cfg.shift0[a]
        return torch.stack(fields).contiguous()


class Observable:
    """J = sum over owned points of (alpha)(chi) w(r) dV, w = exp(-(r - r0)^2 / (2 sigma^2))."""

    def __init__(self, cfg: DiffConfig, device="cuda"):
        import torch
        g = cfg.patch
        X, Y, Z = g.coords()
        z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
        r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
        w = np.exp(-(r - cfg.shell_r) ** 2 / (2.0 * cfg.shell_width ** 2)) * float(np.prod(g.h))
        mask = np.zeros_like(w)
        gw = g.gw
        mask[gw:-gw, gw:-gw, gw:-gw] = 1.0
        self.w = torch.tensor(w * mask, dtype=torch.float64, device=device)

    def value(self, U):
        import torch
        return torch.sum((U[LAP] + 1.0) * (U[CHI] + 1.0) * self.w)

    def grad(self, U):
        """dJ/dU (24 fields)."""
        import torch
        lam = torch.zeros_like(U)
        lam[CHI] = (U[LAP] + 1.0) * self.w
        lam[LAP] = (U[CHI] + 1.0) * self.w
        return lam


class Evolution:
    """The forward evolution (fused level step) with checkpoints, and the
    backward pass through the step adjoint."""

    def __init__(self, cfg: DiffConfig, mode: int = 0):
        self.cfg = cfg
        self.level = Level([cfg.patch], gauge=0, eps=cfg.eps, chitiny=cfg.chitiny)
        self.adj = StepAdjoint(cfg.patch.shape, cfg.patch.h, eps=cfg.eps, chitiny=cfg.chitiny)
        self.mode = mode
        self.timing = {}

    def close(self):
        self.level.close()
        self.adj.close()

    def _put(self, U):
        n = U[0].numel()
        for v in range(NF):
            check(lib().amss_level_copy(self.level.h, 0, LIST_STATE, v, C.c_void_p(U.data_ptr() + 8 * v * n), 0),
                  "amss_level_copy(in)")

    def _get(self, U):
        n = U[0].numel()
        for v in range(NF):
            check(lib().amss_level_copy(self.level.h, 0, LIST_STATE, v, C.c_void_p(U.data_ptr() + 8 * v * n), 1),
                  "amss_level_copy(out)")

    def forward(self, U0):
        """Evolve U0 for cfg.steps steps; returns (final state, checkpoints)."""
        import torch
        cfg = self.cfg
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ckpts = [U0.detach().clone()]
        self._put(ckpts[0])
        done = 0
        while done < cfg.steps:
            k = min(cfg.checkpoint, cfg.steps - done)
            rc = self.level.step(cfg.dt, k, self.mode)
            if rc:
                raise RuntimeError(f"forward evolution failed (code {rc})")
            done += k
            U = torch.empty_like(U0)
            self._get(U)
            ckpts.append(U)
        torch.cuda.synchronize()
        self.timing["forward_s"] = time.perf_counter() - t0
        return ckpts[-1], ckpts[:-1]

    def backward(self, lam, ckpts):
        """lam = dJ/dU(T) -> dJ/dU0, recomputing each segment from its checkpoint."""
        import torch
        cfg = self.cfg
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lam = lam.clone().contiguous()
        nseg = len(ckpts)
        for s in range(nseg - 1, -1, -1):
            k = min(cfg.checkpoint, cfg.steps - s * cfg.checkpoint)
            states = [ckpts[s]]
            for _ in range(k - 1):
                nxt = torch.empty_like(states[-1])
                self.adj.forward(cfg.dt, states[-1], nxt)
                states.append(nxt)
            for n in range(k - 1, -1, -1):
                self.adj.step(cfg.dt, states[n], lam)
        torch.cuda.synchronize()
        self.timing["backward_s"] = time.perf_counter() - t0
        return lam


def _autograd_function():
    import torch

    class EvolveObservable(torch.autograd.Function):
        """J(U0) = observable(evolve(U0)); backward through the step adjoint."""

        @staticmethod
        def forward(ctx, U0, evo, obs):
            UT, ckpts = evo.forward(U0)
            ctx.evo, ctx.ckpts, ctx.lamT = evo, ckpts, obs.grad(UT)
            return obs.value(UT)

        @staticmethod
        def backward(ctx, gJ):
            lam0 = ctx.evo.backward(ctx.lamT * gJ, ctx.ckpts)
            ctx.ckpts = None
            return lam0, None, None

    return EvolveObservable


_EVOLVE = None


def evolve_observable(U0, evo: Evolution, obs: Observable):
    """J as a differentiable function of U0 (a [24, nz, ny, nx] device tensor)."""
    global _EVOLVE
    if _EVOLVE is None:
        _EVOLVE = _autograd_function()
    return _EVOLVE.apply(U0, evo, obs)


def value_and_grad(cfg: DiffConfig, theta=None, device="cuda", mode: int = 0):
    """(J, dJ/dtheta, timings) for the configuration's parameters."""
    import torch
    init = InitialData(cfg, device)
    obs = Observable(cfg, device)
    evo = Evolution(cfg, mode)
    th = torch.tensor(cfg.theta0() if theta is None else theta, dtype=torch.float64, device=device,
                      requires_grad=True)
    try:
        J = evolve_observable(init(th), evo, obs)
        J.backward()
        return float(J.detach()), th.grad.detach().cpu().numpy(), dict(evo.timing)
    finally:
        evo.close()


def value(cfg: DiffConfig, theta, device="cuda", mode: int = 0) -> float:
    """J alone (no tape), for finite differences."""
    import torch
    init = InitialData(cfg, device)
    obs = Observable(cfg, device)
    evo = Evolution(cfg, mode)
    try:
        with torch.no_grad():
            U0 = init(torch.tensor(theta, dtype=torch.float64, device=device))
            UT, _ = evo.forward(U0)
            return float(obs.value(UT))
    finally:
        evo.close()


__all__ = ["DiffConfig", "config3", "InitialData", "Observable", "Evolution", "evolve_observable",
           "value_and_grad", "value", "spline_basis", "bspline3"]
