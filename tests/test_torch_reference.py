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
