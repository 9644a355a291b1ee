"""dJ/dtheta through the checkpointed step adjoint (the autograd.Function)
against torch autograd of the PyTorch restatement on the CPU, and against
central finite differences of the GPU forward evolution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small(steps=4, checkpoint=2, n=16):
    """An n^3 copy at the benchmark spacing h = 0.5: [-n/4, n/4]^3, the
    punctures at x = +-3, the shell at r = 3."""
    import amssbssn.gradient as G
    from amssbssn.grid import cube_patch
    g = cube_patch(n, n / 4.0)
    return G.DiffConfig([g], dt=0.25 * g.h[0], steps=steps, checkpoint=checkpoint, shell_r=3.0, spline_sigma=1e-3)


def _torch_cpu_grad(cfg):
    import torch
    import torch_rhs
    import amssbssn.gradient as G
    init = G.InitialData(cfg, device="cpu")
    obs = G.Observable(cfg, device="cpu")
    th = torch.tensor(cfg.theta0(), dtype=torch.float64, requires_grad=True)
    U = list(init(th)[0].unbind(0))
    for _ in range(cfg.steps):
        U = torch_rhs.step(U, cfg.dt, cfg.patches[0].h, cfg.eps, cfg.chitiny)
    J = obs.value([torch.stack(U)])
    J.backward()
    return float(J.detach()), th.grad.numpy()


def test_gradient_matches_torch_cpu_autograd():
    import amssbssn.gradient as G
    cfg = _small()
    J, g, _ = G.value_and_grad(cfg)
    Jr, gr = _torch_cpu_grad(cfg)
    assert abs(J / Jr - 1) < 1e-12
    scale = np.max(np.abs(gr[4:]))
    assert np.max(np.abs(g[4:] - gr[4:])) <= 1e-9 * scale
    for k in range(4):
        assert abs(g[k] - gr[k]) <= 1e-9 * abs(gr[k]), (k, g[k], gr[k])
    print(f"J {J:.12e}; dJ/d(m1,m2,x1,x2) {g[:4]}; spline grad scale {scale:.3e}")


def test_gradient_matches_central_differences():
    """Central differences with h = 1e-6 for the masses and positions; for the
    4096 spline coefficients, the directional derivative along the unit
    gradient direction (a step of 1e-6 per coefficient on average), where a
    single coefficient's change of J would drown in rounding.  (The exact
    derivative of the discrete map, kinks included, is the torch autograd
    comparison above.)"""
    import dataclasses
    import amssbssn.gradient as G
    # a constant shift keeps every beta^a away from zero, where the upwind
    # stencil switches sides: the map is smooth and differences converge
    cfg = dataclasses.replace(_small(steps=10, checkpoint=4, n=24), shift0=(0.02, -0.015, 0.01))
    theta = cfg.theta0()
    _, g, _ = G.value_and_grad(cfg, theta)
    for k in range(4):
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (G.value(cfg, tp) - G.value(cfg, tm)) / (2 * h)
        print(f"theta[{k}]: adjoint {g[k]:.10e} fd {fd:.10e} rel {abs(fd - g[k]) / abs(g[k]):.2e}")
        assert abs(fd - g[k]) <= 1e-6 * abs(g[k]), k
    d = np.zeros_like(theta)
    d[4:] = g[4:] / np.linalg.norm(g[4:])
    h = 1e-6 * np.sqrt(len(theta) - 4)
    fd = (G.value(cfg, theta + h * d) - G.value(cfg, theta - h * d)) / (2 * h)
    lin = float(np.dot(g, d))
    print(f"spline direction: adjoint {lin:.10e} fd {fd:.10e} rel {abs(fd - lin) / abs(lin):.2e}")
    assert abs(fd - lin) <= 1e-6 * abs(lin)


def test_level_gradient_matches_central_differences():
    """A level of 3 uneven patches (exchange on every sub-step): dJ/d(masses,
    positions) through the level adjoint against central differences."""
    import dataclasses
    import amssbssn.gradient as G
    from amssbssn.grid import uneven_level
    patches, _ = uneven_level(3, 16, 24, seed=4, half_width=5.0)
    cfg = G.DiffConfig(patches, dt=0.25 * patches[0].h[0], steps=6, checkpoint=4, spline_n=0,
                       shell_r=3.0, shift0=(0.02, -0.015, 0.01))
    theta = cfg.theta0()
    _, g, info = G.value_and_grad(cfg, theta, mode=2)
    for k in range(4):
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (G.value(cfg, tp, mode=2) - G.value(cfg, tm, mode=2)) / (2 * h)
        print(f"level theta[{k}]: adjoint {g[k]:.10e} fd {fd:.10e} rel {abs(fd - g[k]) / abs(g[k]):.2e}")
        assert abs(fd - g[k]) <= 1e-6 * abs(g[k]), k
