"""Psi4 on the GPU (getnp4_ and amss_psi4_dev) against the oracle's getnp4
and the closed form of linear data: random states with and without symmetry
planes, a vertex-centred grid through the z axis and the origin, and a
two-puncture state evolved 10 steps on a 64^3 patch; relative to the largest
|Psi4|, 1e-12 (the oracle rounds in another order)."""
import numpy as np
import pytest

from helpers import random_state, rel_err
from test_oracle_psi4 import _axis_grid, _expected_linear, _linear_case, _sym_grids

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sym", [0, 1, 2, "axis"])
def test_getnp4_matches_oracle(orc, sym):
    import amssbssn.fortran_abi as F
    if sym == "axis":
        X, Y, Z = _axis_grid()
        sym = 0
    else:
        X, Y, Z = _sym_grids(sym, (13, 12, 11))
    st = random_state((len(Z), len(Y), len(X)), seed=80 + sym)
    aux = orc.compute_rhs(st, X, Y, Z, 0.02, symmetry=sym, gauge=0)
    want = orc.getnp4(st, aux["gam"], aux["ric"], X, Y, Z, symmetry=sym)
    got = F.getnp4(st, aux["gam"], aux["ric"], X, Y, Z, symmetry=sym)
    for g, w in zip(got, want):
        assert np.all(np.isfinite(g))
        e = rel_err(g, w)
        print(f"Symmetry {sym}: rel err {e:.2e}")
        assert e <= 1e-12


@pytest.mark.parametrize("seed", [1, 2])
def test_getnp4_linear_data_closed_form(seed):
    import amssbssn.fortran_abi as F
    X, Y, Z, st, gam, ric, g, c, kk, b = _linear_case(seed, n=(12, 11, 10))
    got = F.getnp4(st, gam, ric, X, Y, Z)
    want = _expected_linear(X, Y, Z, st, gam, ric, g, c, kk, b)
    for a, w in zip(got, want):
        assert rel_err(a, w) <= 1e-12


def test_getnp4_rejects_bad_arguments():
    import amssbssn as A
    import amssbssn.fortran_abi as F
    X, Y, Z = _sym_grids(0, (6, 6, 6))
    st = random_state((6, 6, 6), seed=3)
    z = [np.zeros((6, 6, 6)) for _ in range(24)]
    with pytest.raises(A.AmssError, match="Symmetry"):
        F.getnp4(st, z[:18], z[:6], X, Y, Z, symmetry=3)


def test_psi4_dev_of_an_evolved_two_puncture_state(orc):
    """64^3 + ghosts, two punctures, 10 RK4 steps on the GPU; the Christoffels
    and Ricci tensor from the device RHS of that state feed both sides."""
    import torch
    import amssbssn as A
    import amssbssn.fortran_abi as F
    cfg = A.config1(n=64)
    lv = A.Level.from_config(cfg, gauge=0)
    try:
        assert lv.step(cfg.dt, 10) == 0
        U = lv.get_state(0)
    finally:
        lv.close()
    geom = cfg.patches[0]
    X, Y, Z = geom.coords()
    dev = torch.device("cuda:0")
    st = torch.tensor(np.stack(U), dtype=torch.float64, device=dev)
    aux = F.compute_rhs_dev(st, geom.h, cfg.eps, co=1, gauge=0)
    assert aux["flag"] == 0
    xs = [torch.tensor(a, dtype=torch.float64, device=dev) for a in (X, Y, Z)]
    rp, ip = F.psi4_dev(st, aux["gam"], aux["ric"], *xs, geom.h)
    gam = [a for a in aux["gam"].cpu().numpy()]
    ric = [a for a in aux["ric"].cpu().numpy()]
    wr, wi = orc.getnp4(U, gam, ric, X, Y, Z)
    # the Fortran-ABI entry point runs the same kernel
    fr, fi = F.getnp4(U, gam, ric, X, Y, Z)
    assert np.array_equal(fr, rp.cpu().numpy()) and np.array_equal(fi, ip.cpu().numpy())
    for g, w in ((rp, wr), (ip, wi)):
        g = g.cpu().numpy()
        assert float(np.max(np.abs(w))) > 1e-6
        e = rel_err(g, w)
        print(f"64^3 two punctures, 10 steps: max |Psi4| {np.max(np.abs(w)):.3e}, rel err {e:.2e}")
        assert e <= 1e-12
