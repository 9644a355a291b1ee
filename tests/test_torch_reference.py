"""The PyTorch FP64 restatement of the RHS (the autograd reference of the
adjoint) against the C oracle at interior points, and its autograd
vector-Jacobian product against central finite differences."""
import numpy as np
import pytest

from helpers import grid, random_state


@pytest.mark.parametrize("gauge", [0, 2])
def test_torch_rhs_matches_oracle_interior(orc, gauge):
    import torch_rhs
    import torch
    nx, ny, nz = 12, 13, 14
    X, Y, Z, _ = grid(nx, ny, nz)
    st = random_state((nz, ny, nx), seed=21, amp=0.05)
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=1, gauge=gauge)
    h = (X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0])
    t = torch_rhs.rhs([torch.tensor(u) for u in st], h, 0.02, gauge)
    s = np.s_[3:-3, 3:-3, 3:-3]
    for v in range(24):
        ref = o["rhs"][v][s]
        m = float(np.max(np.abs(ref)))
        assert float(np.max(np.abs(t[v].numpy() - ref))) <= 1e-12 * m, v


def test_torch_vjp_matches_central_differences():
    """<lam, J e> from autograd equals the central difference of <lam, rhs>
    along a random direction e (O(h^2) error at h = 1e-5)."""
    import torch_rhs
    nx, ny, nz = 11, 11, 12
    X, Y, Z, _ = grid(nx, ny, nz)
    h = (X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0])
    st = random_state((nz, ny, nx), seed=22, amp=0.05)
    rng = np.random.default_rng(5)
    lam = [rng.standard_normal((nz - 6, ny - 6, nx - 6)) for _ in range(24)]
    e = [rng.standard_normal((nz, ny, nx)) for _ in range(24)]
    _, g = torch_rhs.interior_vjp(st, lam, h, 0.02)
    lin = sum(float(np.sum(gv * ev)) for gv, ev in zip(g, e))

    import torch

    def J(eps):
        out = torch_rhs.rhs([torch.tensor(u + eps * d) for u, d in zip(st, e)], h, 0.02)
        return sum(float(np.sum(l * o.numpy())) for l, o in zip(lam, out))
    fd = (J(1e-5) - J(-1e-5)) / 2e-5
    assert abs(fd - lin) <= 1e-6 * abs(lin)


def test_torch_step_matches_oracle_step(orc):
    """The PyTorch restatement of one RK4 step (frozen ghosts, width 3)
    against the C oracle's step."""
    import torch
    import torch_rhs
    nx, ny, nz = 18, 19, 20
    X, Y, Z, bbox = grid(nx, ny, nz)
    h = (X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0])
    st = random_state((nz, ny, nx), seed=33, amp=0.03)
    dt = 0.25 * min(h)
    U = [u.copy() for u in st]
    orc.step(U, X, Y, Z, bbox, dt, 0.02, 1e-5, gauge=0, bc=0, gw=3)
    T = torch_rhs.step([torch.tensor(u) for u in st], dt, h, 0.02)
    for v in range(24):
        m = max(float(np.max(np.abs(U[v]))), 1e-300)
        assert float(np.max(np.abs(T[v].numpy() - U[v]))) <= 1e-13 * m, v
