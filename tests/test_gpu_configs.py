"""The benchmark configurations as GPU tests at their own sizes.

* 37 uneven patches (multiples of 8 in [24,96], seed 1, [-24,24]^3): the
  batched tile-list step and its hipGraph replay against the per-patch step,
  residuals in every predictor, 3 steps: states to 1e-12, the seven
  constraint norms of every step to 1e-12.
* 64^3 patch, 200 RK4 steps, checkpoint every 20 (the differentiation
  workload): dJ/dm1 and dJ/dx1 from the adjoint against central differences
  (dJ/dm1 to 1e-6; dJ/dx1, 2e4 times smaller, to 1e-5); and on a 24^3 copy
  (4 steps) dJ/dtheta against torch
  autograd of the PyTorch FP64 restatement on the CPU to 1e-9.
* 16 uneven patches (multiples of 8 in [48,128], seed 2), 400 RK4 steps,
  batched hipGraph forward, checkpoint spacing planned from the free HBM:
  dJ/dm1 against central differences to 1e-6, with peak memory and the
  adjoint/forward ratio printed.
* 128^3 two-puncture run at GAUGE 0, 100 steps, against the fixture the C
  restatement wrote (512 samples, max |f| and sum f^2 per field, the seven
  constraint norms of steps 1, 50 and 100) to 1e-9.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    """|a - b| relative to |b|, or to the largest entry of b where b is zero
    (the Gamma-constraint norms of conformally flat data vanish)."""
    b = np.asarray(b)
    floor = 1e-2 * np.max(np.abs(b)) if b.size else 1.0
    return np.abs(a - b) / np.maximum(np.abs(b), max(floor, 1e-300))


@pytest.mark.parametrize("mode", [1, 2])
def test_37_patch_level_batched_and_graph_equal_per_patch(mode):
    import amssbssn as A
    cfg = A.config2()
    out = []
    for m in (0, mode):
        lv = A.Level.from_config(cfg, gauge=0)
        try:
            assert lv.step(cfg.dt, 3, m, want_ham=A.RES_EVERY_STEP) == 0
            out.append(([lv.get_state(p) for p in range(len(cfg.patches))], lv.constraint_series(),
                        [lv.constraint_series(p) for p in range(len(cfg.patches))]))
        finally:
            lv.close()
    (sa, na, pa), (sb, nb, pb) = out
    # per field, relative to the field's largest value on the level (a patch
    # far from the punctures holds a nearly vanishing trK, say)
    worst = 0.0
    for v in range(24):
        m = max(float(np.max(np.abs(sa[p][v]))) for p in range(len(cfg.patches)))
        d = max(float(np.max(np.abs(sb[p][v] - sa[p][v]))) for p in range(len(cfg.patches)))
        e = d / m if m > 0 else d
        worst = max(worst, e)
        assert e <= 1e-12, (v, e)
    assert np.max(_rel(nb, na)) <= 1e-12
    # per patch, relative to the largest patch's norm (a far patch's residual
    # is rounding-sized, and the tile kernel rounds differently)
    A_, B_ = np.stack(pa), np.stack(pb)
    scale = np.max(np.abs(A_), axis=0)
    assert np.all(np.abs(B_ - A_) <= 1e-12 * scale + 1e-300)
    print(f"37 patches, mode {mode} vs per patch: worst state {worst:.2e}, Ham norms {nb[:, 0]}")


def _central(G, cfg, theta, k, h=1e-6, mode=0):
    tp, tm = theta.copy(), theta.copy()
    tp[k] += h
    tm[k] -= h
    return (G.value(cfg, tp, mode=mode) - G.value(cfg, tm, mode=mode)) / (2 * h)


def test_64cubed_200_steps_gradient_matches_central_differences():
    """dJ/dx1 (about 0.022) is 2e4 times smaller than dJ/dm1 on J of about
    757, whose rounding over 200 steps is about 1e-13 absolute: at h = 1e-6
    that alone is 4e-6 of dJ/dx1, so x1 takes h = 1e-5, where the truncation
    error is 4e-6 (it falls as h^2 from 3.8e-4 at h = 1e-4), checked to 1e-5."""
    import amssbssn.gradient as G
    cfg = G.config3(64, steps=200, checkpoint=20)
    theta = cfg.theta0()
    J, g, info = G.value_and_grad(cfg, theta)
    for k, name, h, tol in ((0, "m1", 1e-6, 1e-6), (2, "x1", 1e-5, 1e-5)):
        fd = _central(G, cfg, theta, k, h=h)
        rel = abs(fd - g[k]) / abs(g[k])
        print(f"64^3 200 steps dJ/d{name}: adjoint {g[k]:.12e} fd (h {h:.0e}) {fd:.12e} rel {rel:.2e}")
        assert rel <= tol, (name, rel)
    print(f"J {J:.12e}; forward {info['forward_s']:.3f} s, backward {info['backward_s']:.3f} s")


def test_24cubed_copy_gradient_matches_torch_cpu_autograd():
    """The differentiation workload's data on a 24^3 copy at its spacing
    (h = 0.5 over [-6,6]^3, punctures at x = +-3, the shell at r = 3)."""
    import torch
    import torch_rhs
    import amssbssn.gradient as G
    from amssbssn.grid import cube_patch
    g = cube_patch(24, 6.0)
    cfg = G.DiffConfig([g], dt=0.25 * g.h[0], steps=4, checkpoint=2, shell_r=3.0, spline_sigma=1e-3)
    J, grad, _ = G.value_and_grad(cfg)
    init = G.InitialData(cfg, device="cpu")
    obs = G.Observable(cfg, device="cpu")
    th = torch.tensor(cfg.theta0(), dtype=torch.float64, requires_grad=True)
    U = list(init(th)[0].unbind(0))
    for _ in range(cfg.steps):
        U = torch_rhs.step(U, cfg.dt, g.h, cfg.eps, cfg.chitiny)
    Jr = obs.value([torch.stack(U)])
    Jr.backward()
    gr = th.grad.numpy()
    assert abs(J / float(Jr.detach()) - 1) < 1e-12
    scale = np.max(np.abs(gr[4:]))
    assert np.max(np.abs(grad[4:] - gr[4:])) <= 1e-9 * scale
    for k in range(4):
        assert abs(grad[k] - gr[k]) <= 1e-9 * abs(gr[k]), (k, grad[k], gr[k])


@pytest.mark.timeout(900)
def test_16_patch_level_400_steps_gradient_matches_central_differences():
    import amssbssn.gradient as G
    cfg = G.config4(steps=400)
    theta = cfg.theta0()
    J, g, info = G.value_and_grad(cfg, theta, mode=2)
    fd = _central(G, cfg, theta, 0, mode=2)
    rel = abs(fd - g[0]) / abs(g[0])
    ratio = info["backward_s"] / info["forward_s"]
    peak = info["peak_device_bytes"] / 1e9 if info["peak_device_bytes"] else float("nan")
    print(f"16 patches, 400 steps: checkpoint every {info['checkpoint']} (tape {info['tape']}), "
          f"peak device memory {peak:.1f} GB, forward {info['forward_s']:.2f} s, backward "
          f"{info['backward_s']:.2f} s (x{ratio:.2f}); dJ/dm1 adjoint {g[0]:.12e} fd {fd:.12e} rel {rel:.2e}")
    assert rel <= 1e-6
    assert info["peak_device_bytes"] is None or info["peak_device_bytes"] < 288e9


def _fixture(name):
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    if not os.path.exists(p):
        pytest.skip(f"fixture {name} not generated")
    return np.load(p)


@pytest.mark.parametrize("gauge,name", [(0, "orc_c1_128cubed_step100_gauge0.npz"),
                                        (2, "ref_c1_128cubed_step100_gauge2.npz")])
def test_128cubed_100_steps_with_constraint_norms_vs_fixture(gauge, name):
    import amssbssn as A
    g = _fixture(name)
    if "con_l2" not in g.files:
        pytest.skip(f"{name} holds no constraint norms")
    cfg = A.config1(n=128)
    lv = A.Level.from_config(cfg, gauge=gauge)
    try:
        assert lv.step(cfg.dt, 100, want_ham=A.RES_EVERY_STEP) == 0
        G = lv.get_state(0)
        series = lv.constraint_series()
    finally:
        lv.close()
    idx = g["idx"]
    worst = 0.0
    for v, u in enumerate(G):
        m = float(g["maxabs"][v])
        d = float(np.max(np.abs(u.ravel()[idx] - g["samples"][v])))
        e = d / m if m > 0 else d
        worst = max(worst, e)
        assert e <= 1e-9, (v, e)
        if m > 0:
            assert abs(float(np.max(np.abs(u))) / m - 1) <= 1e-9, v
            assert abs(float(np.sum(u * u)) / float(g["sumsq"][v]) - 1) <= 2e-9, v
    rows = series[np.array(g["con_steps"]) - 1]
    want = g["con_l2"]
    # relative per norm; the Gamma-constraint norms of the conformally flat
    # initial data are exactly zero, where rounding is measured against the
    # largest norm
    tol = 1e-9 * np.abs(want) + 1e-14 * np.max(np.abs(want))
    assert np.all(np.abs(rows - want) <= tol), (rows, want)
    print(f"128^3 GAUGE {gauge}, 100 steps: worst sample {worst:.2e}; Ham norms {rows[:, 0]}")
