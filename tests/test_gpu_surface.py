"""Psi4's spin-weighted modes on the GPU (amss_surf_wave_dev) against the
oracle's surf_Wave: the polynomial fields of the closed-form test (Symmetry
0, 1, 2), a random field, and the Psi4 of a two-puncture state evolved 10
steps on a 64^3 patch (amss_psi4_dev -> modes at r = 6); relative to the
largest mode, 1e-12."""
import numpy as np
import pytest

from test_oracle_surface import SOA_I, SOA_R, _expected, _fields, _grid, _mesh

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")


def _rel(got, want):
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


@pytest.mark.parametrize("sym", [0, 1, 2])
def test_modes_of_polynomial_fields_match_oracle_and_closed_form(orc, sym):
    import amssbssn.fortran_abi as F
    X, Y, Z = _grid(sym)
    R, I = _fields(*_mesh(X, Y, Z), sym)
    rex, maxl = 2.5, 5
    rp, ip = F.surf_wave_dev(_dev(R), _dev(I), X, Y, Z, rex, 12, maxl, soa_r=SOA_R, soa_i=SOA_I, symmetry=sym)
    rc, orp, oip = orc.surf_wave(R, I, X, Y, Z, rex, 12, maxl, soa_r=SOA_R, soa_i=SOA_I, symmetry=sym)
    assert rc == 0
    e = _rel(rp + 1j * ip, orp + 1j * oip)
    print(f"Symmetry {sym}: GPU vs oracle {e:.2e}")
    assert e <= 1e-12
    assert _rel(rp + 1j * ip, _expected(sym, rex, maxl)) <= 1e-12


def test_modes_of_a_random_field_match_oracle(orc):
    import amssbssn.fortran_abi as F
    X, Y, Z = _grid(0, n_int=24, h=0.3)
    rng = np.random.default_rng(21)
    R = rng.standard_normal((len(Z), len(Y), len(X)))
    I = rng.standard_normal(R.shape)
    rp, ip = F.surf_wave_dev(_dev(R), _dev(I), X, Y, Z, 2.9, 10, 6)
    rc, orp, oip = orc.surf_wave(R, I, X, Y, Z, 2.9, 10, 6)
    assert rc == 0
    assert _rel(rp + 1j * ip, orp + 1j * oip) <= 1e-12


def test_psi4_modes_of_an_evolved_two_puncture_state(orc):
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
    g = cfg.patches[0]
    X, Y, Z = g.coords()
    st = _dev(np.stack(U))
    aux = F.compute_rhs_dev(st, g.h, cfg.eps, co=1, gauge=0)
    rpsi, ipsi = F.psi4_dev(st, aux["gam"], aux["ric"], _dev(X), _dev(Y), _dev(Z), g.h)
    rp, ip = F.surf_wave_dev(rpsi, ipsi, X, Y, Z, 6.0, 20, 4)
    rc, orp, oip = orc.surf_wave(rpsi.cpu().numpy(), ipsi.cpu().numpy(), X, Y, Z, 6.0, 20, 4)
    assert rc == 0
    e = _rel(rp + 1j * ip, orp + 1j * oip)
    print(f"two punctures, r = 6: |Psi4_22| {abs(rp[4] + 1j * ip[4]):.3e}, GPU vs oracle {e:.2e}")
    assert e <= 1e-12
    # the equal-mass pair on the x axis radiates mostly in l = 2, m = +-2 and m = 0
    amp = np.abs(rp + 1j * ip)
    assert np.argmax(amp[:5]) in (0, 2, 4)


def test_sphere_outside_the_block_is_refused():
    import amssbssn as A
    import amssbssn.fortran_abi as F
    X, Y, Z = _grid(0)
    R, I = _fields(*_mesh(X, Y, Z), 0)
    with pytest.raises(A.AmssError, match="leaves the block"):
        F.surf_wave_dev(_dev(R), _dev(I), X, Y, Z, 4.5, 8, 2)
