"""The hand-written adjoint kernels: the RHS vector-Jacobian product against
torch autograd of the PyTorch restatement, the adjoint's forward step
against the level's fused step, and the step adjoint against central
finite differences of the step itself."""
import numpy as np
import pytest

from helpers import grid, random_state, rel_err

pytestmark = pytest.mark.gpu


def _dev(st):
    import torch
    return torch.tensor(np.stack(st), dtype=torch.float64, device="cuda").contiguous()


def test_rhs_vjp_matches_torch_autograd():
    import torch
    import torch_rhs
    from amssbssn.adjoint import StepAdjoint
    nx, ny, nz = 14, 15, 16
    X, Y, Z, _ = grid(nx, ny, nz)
    h = (X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0])
    st = random_state((nz, ny, nx), seed=31, amp=0.05)
    rng = np.random.default_rng(3)
    lam_i = [rng.standard_normal((nz - 6, ny - 6, nx - 6)) for _ in range(24)]
    _, ref = torch_rhs.interior_vjp(st, lam_i, h, 0.02)
    lam = [np.zeros((nz, ny, nx)) for _ in range(24)]
    for v in range(24):
        lam[v][3:-3, 3:-3, 3:-3] = lam_i[v]
    A = StepAdjoint((nx, ny, nz), h, eps=0.02)
    U, L = _dev(st), _dev(lam)
    out = torch.zeros_like(U)
    A.rhs_vjp(U, L, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    A.close()
    worst = 0.0
    for v in range(24):
        e = rel_err(got[v], ref[v])
        worst = max(worst, e)
        assert e <= 1e-12, (v, e)
    print(f"RHS VJP vs torch autograd: worst {worst:.2e}")


def _perturbed_puncture(n=16, seed=41, amp=1e-3, shift=(0.0, 0.0, 0.0)):
    import amssbssn as A
    cfg = A.config1(n=n)
    st = cfg.initial_state(0)
    rng = np.random.default_rng(seed)
    for v in range(24):
        st[v] = st[v] + amp * rng.standard_normal(st[v].shape)
    for a in range(3):
        st[18 + a] = st[18 + a] + shift[a]
    p = cfg.patches[0]
    return cfg, p, st


def test_adjoint_forward_matches_level_step():
    import torch
    import amssbssn as A
    from amssbssn.adjoint import StepAdjoint
    cfg, p, st = _perturbed_puncture()
    lv = A.Level(cfg.patches, eps=cfg.eps)
    lv.set_state(0, st)
    assert lv.step(cfg.dt, 1) == 0
    ref = lv.get_state(0)
    lv.close()
    adj = StepAdjoint(p.shape, p.h, eps=cfg.eps)
    U = _dev(st)
    Un = torch.empty_like(U)
    adj.forward(cfg.dt, U, Un)
    torch.cuda.synchronize()
    got = Un.cpu().numpy()
    adj.close()
    for v in range(24):
        assert rel_err(got[v], ref[v]) <= 1e-13, v


@pytest.mark.parametrize("nsteps", [1, 3])
def test_step_adjoint_matches_central_differences(nsteps):
    """J = <w, U^{n+nsteps}>: <dJ/dU^n, e> from the adjoint against the
    central difference of J along a random direction e."""
    import torch
    from amssbssn.adjoint import StepAdjoint
    # a shift well away from zero: the upwind stencil's switch at beta = 0 is
    # a kink that a finite difference of 1e-6 would straddle at some points
    cfg, p, st = _perturbed_puncture(shift=(0.03, -0.02, 0.025))
    rng = np.random.default_rng(7)
    w = _dev([rng.standard_normal(st[0].shape) for _ in range(24)])
    e = _dev([rng.standard_normal(st[0].shape) for _ in range(24)])
    adj = StepAdjoint(p.shape, p.h, eps=cfg.eps)
    U0 = _dev(st)

    def run(U):
        traj = [U]
        for _ in range(nsteps):
            nxt = torch.empty_like(U)
            adj.forward(cfg.dt, traj[-1], nxt)
            traj.append(nxt)
        return traj

    traj = run(U0)
    lam = w.clone()
    for n in range(nsteps - 1, -1, -1):
        adj.step(cfg.dt, traj[n], lam)
    torch.cuda.synchronize()
    lin = float(torch.sum(lam * e))
    eps = 1e-6

    def J(s):
        return float(torch.sum(w * run(U0 + s * e)[-1]))
    fd = (J(eps) - J(-eps)) / (2 * eps)
    adj.close()
    print(f"{nsteps} step(s): adjoint {lin:.12e} fd {fd:.12e} rel {abs(fd - lin) / abs(lin):.2e}")
    assert abs(fd - lin) <= 1e-6 * abs(lin)


def test_step_adjoint_matches_torch_autograd_of_the_step():
    """One full step (enforce, four RHS, RK4 combination, chi bound, ghost
    fill) differentiated by torch autograd on the CPU against the adjoint
    kernels, on an 18 x 19 x 20 block with generic shift signs."""
    import torch
    import torch_rhs
    from amssbssn.adjoint import StepAdjoint
    nx, ny, nz = 18, 19, 20
    X, Y, Z, _ = grid(nx, ny, nz)
    h = (X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0])
    st = random_state((nz, ny, nx), seed=33, amp=0.03)
    rng = np.random.default_rng(9)
    w = [rng.standard_normal((nz, ny, nx)) for _ in range(24)]
    dt = 0.25 * min(h)
    U = [torch.tensor(u, requires_grad=True) for u in st]
    out = torch_rhs.step(U, dt, h, 0.02)
    ref = torch.autograd.grad(out, U, grad_outputs=[torch.tensor(x) for x in w])
    adj = StepAdjoint((nx, ny, nz), h, eps=0.02)
    Ud, lam = _dev(st), _dev(w)
    nxt = torch.empty_like(Ud)
    adj.forward(dt, Ud, nxt)
    adj.step(dt, Ud, lam)
    torch.cuda.synchronize()
    fwd, got = nxt.cpu().numpy(), lam.cpu().numpy()
    adj.close()
    worst = 0.0
    for v in range(24):
        assert rel_err(fwd[v], out[v].detach().numpy()) <= 1e-12, ("forward", v)
        e = rel_err(got[v], ref[v].numpy())
        worst = max(worst, e)
        assert e <= 1e-11, (v, e)
    print(f"step adjoint vs torch autograd: worst {worst:.2e}")


def test_level_adjoint_split_equals_whole():
    """A 20^3 cube and the same cube as two 10 x 20 x 20 patches: the split
    forward step reproduces the whole interior, and for J = <w, U^{n+2}>
    the split adjoint (exchange reversed) gives the same directional
    derivative sum_p <dJ/dU0_p, e|p> as the whole cube's <dJ/dU0, e>."""
    import torch
    import amssbssn as A
    from amssbssn.adjoint import LevelAdjoint
    n, gw = 20, 3
    whole = A.cube_patch(n, 8.0)
    h = whole.h[0]
    lo = -8.0 - gw * h
    left = A.PatchGeom((n // 2 + 2 * gw, n + 2 * gw, n + 2 * gw),
                       (lo, lo, lo, 0.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    right = A.PatchGeom((n // 2 + 2 * gw, n + 2 * gw, n + 2 * gw),
                        (0.0 - gw * h, lo, lo, 8.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    cfg = A.Config("whole", [whole], dt=0.25 * h, steps=2, punctures=[(0.5, 1.0, 0.5, 0.0), (0.5, -1.5, 0.0, 0.3)],
                   a_noise=1e-6)
    rng = np.random.default_rng(12)
    U = _dev(cfg.initial_state(0))
    U = U + 1e-3 * torch.tensor(rng.standard_normal(U.shape), device="cuda")
    w = torch.zeros_like(U)
    w[:, gw:-gw, gw:-gw, gw:-gw] = torch.tensor(rng.standard_normal((24, n, n, n)), device="cuda")
    e = torch.tensor(rng.standard_normal(U.shape), device="cuda")
    cut = [lambda t: t[..., : n // 2 + 2 * gw].contiguous(), lambda t: t[..., n // 2:].contiguous()]
    Lw = LevelAdjoint([whole], eps=cfg.eps)
    Ls = LevelAdjoint([left, right], eps=cfg.eps)
    # forward: two steps each
    tw = [[U]]
    ts = [[c(U) for c in cut]]
    for _ in range(2):
        nw = [torch.empty_like(U)]
        Lw.forward(cfg.dt, tw[-1], nw)
        tw.append(nw)
        ns = [torch.empty_like(x) for x in ts[-1]]
        Ls.forward(cfg.dt, ts[-1], ns)
        ts.append(ns)
    W = tw[-1][0][:, gw:-gw, gw:-gw, gw:-gw]
    a = ts[-1][0][:, gw:-gw, gw:-gw, gw:-gw]
    b = ts[-1][1][:, gw:-gw, gw:-gw, gw:-gw]
    m = float(W.abs().max())
    assert float((W[..., : n // 2] - a).abs().max()) <= 1e-13 * m
    assert float((W[..., n // 2:] - b).abs().max()) <= 1e-13 * m
    # adjoint
    lw = [w.clone()]
    wl = torch.zeros_like(ts[0][0])
    wr = torch.zeros_like(ts[0][1])
    wl[:, gw:-gw, gw:-gw, gw:-gw] = w[:, gw:-gw, gw:-gw, gw:gw + n // 2]
    wr[:, gw:-gw, gw:-gw, gw:-gw] = w[:, gw:-gw, gw:-gw, gw + n // 2:-gw]
    ls = [wl, wr]
    for k in (1, 0):
        Lw.step(cfg.dt, tw[k], lw)
        Ls.step(cfg.dt, ts[k], ls)
    torch.cuda.synchronize()
    dw = float(torch.sum(lw[0] * e))
    ds = sum(float(torch.sum(l * c(e))) for l, c in zip(ls, cut))
    Lw.close()
    Ls.close()
    print(f"whole {dw:.12e} split {ds:.12e} rel {abs(ds - dw) / abs(dw):.2e}")
    assert abs(ds - dw) <= 1e-11 * abs(dw)


def test_tape_slots_give_the_same_adjoint():
    """Reversing steps from tape slots written by the forward pass equals
    reversing them with the step re-evaluated (same kernels, same inputs)."""
    import torch
    from amssbssn.adjoint import LevelAdjoint
    cfg, p, st = _perturbed_puncture(n=12, shift=(0.01, 0.0, -0.01))
    rng = np.random.default_rng(8)
    w = _dev([rng.standard_normal(st[0].shape) for _ in range(24)])
    L = LevelAdjoint([p], eps=cfg.eps)
    L.reserve(3)
    traj = [_dev(st)]
    for n in range(3):
        nxt = torch.empty_like(traj[-1])
        L.forward(cfg.dt, [traj[-1]], [nxt], n)
        traj.append(nxt)
    a, b = w.clone(), w.clone()
    for n in (2, 1, 0):
        L.step(cfg.dt, [traj[n]], [a], n)
    for n in (2, 1, 0):
        L.step(cfg.dt, [traj[n]], [b], -1)
    torch.cuda.synchronize()
    L.close()
    assert torch.equal(a, b)
