"""The extraction of Psi4's spin-weighted modes (surface_integral::surf_Wave
with the Gauss-Legendre rule of surface_integral.C:39-181, and global_interp,
fmisc.f90:712-794) in the oracle: global_interp bit for bit against the
reference build, and the modes against what the mathematics fixes.

Closed form: a Psi4 field that is a polynomial of degree <= 5 per axis is
interpolated exactly by the 6-point rule, and the modes of its restriction
to the sphere of radius R,
    R int Psi4(R n) sqrt((2l+1)/4pi) d^l_{m,2}(theta) e^{-i m phi} dOmega,
are computed here by an independent quadrature (96 Gauss-Legendre nodes in
cos(theta), 192 midpoint nodes in phi, both exact for these integrands) with
d^l_{m,s}(theta) = <l m| exp(-i theta J_y) |l s> from the matrix exponential
of the angular-momentum operator -- not from the factorial sum the reference
and the oracle use.  Symmetry 1 and 2 unfold a hemisphere / octant through
the fields' parities; fields with those parities on the full sphere give the
same modes."""
import math

import numpy as np
import pytest

SOA_R = (1.0, 1.0, 1.0)
SOA_I = (-1.0, -1.0, 1.0)


def _axis(n_int, h, plane):
    """Cell-centred coordinates: from the plane (no cells below it) or
    symmetric about 0 with 3 ghost cells at each end."""
    if plane:
        return (np.arange(n_int + 3) + 0.5) * h
    return (np.arange(n_int + 6) - (n_int + 6) / 2 + 0.5) * h


def _grid(sym, n_int=32, h=0.25):
    X = _axis(n_int if sym < 2 else n_int // 2, h, sym == 2)
    Y = _axis(n_int if sym < 2 else n_int // 2, h, sym == 2)
    Z = _axis(n_int if sym < 1 else n_int // 2, h, sym >= 1)
    return X, Y, Z


def _fields(x, y, z, sym):
    if sym == 0:
        R = 0.3 + 0.7 * x - 0.2 * y * z + 0.4 * x * x * z - 0.05 * x * y ** 3 + 0.01 * z ** 5
        I = -0.1 + 0.5 * y + 0.3 * x * z - 0.2 * y * y * x + 0.02 * x ** 4 * z
    else:  # even in x, y, z (SOA_R); odd in x and y, even in z (SOA_I)
        R = 0.3 + x * x - 0.5 * y * y * z * z + 0.2 * x ** 4 * y * y + 0.1 * z ** 4
        I = x * y * (1.0 + z * z) + 0.1 * x ** 3 * y - 0.05 * x * y ** 3 * z ** 4
    return R, I


def _mesh(X, Y, Z):
    z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
    return x, y, z


def _wigner_d(l, m, s, theta):
    """<l m| exp(-i theta J_y) |l s> from the eigenvectors of J_y (states
    ordered m = l .. -l, J_+ |l m> = sqrt(l(l+1) - m(m+1)) |l m+1>)."""
    ms = list(range(l, -l - 1, -1))
    Jp = np.zeros((2 * l + 1, 2 * l + 1))
    for a, mm in enumerate(ms):
        if mm < l:
            Jp[a - 1, a] = math.sqrt(l * (l + 1) - mm * (mm + 1))
    lam, V = np.linalg.eigh((Jp - Jp.T) / 2j)
    a, b = ms.index(m), ms.index(s)
    theta = np.asarray(theta, dtype=float)
    return np.real(np.exp(-1j * theta[..., None] * lam) @ (V[a, :] * np.conj(V[b, :])))


def _expected(sym, rex, maxl, spinw=2):
    ct, wt = np.polynomial.legendre.leggauss(96)
    th = np.arccos(ct)
    nph = 192
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    x, y, z = rex * np.sin(T) * np.cos(P), rex * np.sin(T) * np.sin(P), rex * np.cos(T)
    R, I = _fields(x, y, z, sym)
    psi = R + 1j * I
    out = []
    for l in range(spinw, maxl + 1):
        for m in range(-l, l + 1):
            d = _wigner_d(l, m, spinw, th)
            Y = math.sqrt((2 * l + 1) / (4 * np.pi)) * d[:, None] * np.exp(1j * m * P)
            out.append(rex * np.sum(psi * np.conj(Y) * wt[:, None]) * (2 * np.pi / nph))
    return np.array(out)


@pytest.mark.parametrize("sym", [0, 1, 2])
def test_surf_wave_closed_form(orc, sym):
    X, Y, Z = _grid(sym)
    R, I = _fields(*_mesh(X, Y, Z), sym)
    rex, maxl = 2.5, 5
    rc, rp, ip = orc.surf_wave(R, I, X, Y, Z, rex, 12, maxl, soa_r=SOA_R, soa_i=SOA_I, symmetry=sym)
    assert rc == 0
    want = _expected(sym, rex, maxl)
    got = rp + 1j * ip
    scale = float(np.max(np.abs(want)))
    assert scale > 0.1
    assert np.max(np.abs(got - want)) <= 1e-12 * scale, np.max(np.abs(got - want)) / scale
    # the fields excite several modes, not just one
    assert np.sum(np.abs(want) > 1e-3 * scale) >= 4


def test_surf_wave_orthonormal_modes_pick_out_one_coefficient(orc):
    """A field whose restriction to the sphere is a combination of two of the
    basis functions, 2Y_{2,2} and 2Y_{3,-1} times 1/r: on the sphere the
    modes are those two coefficients and zeros (orthonormality of the basis,
    exact in the quadrature); the field is not a polynomial and not smooth on
    the z axis, so the grid interpolation limits the agreement to ~4e-4."""
    X, Y, Z = _grid(0, n_int=64, h=0.125)
    x, y, z = _mesh(X, Y, Z)
    r = np.sqrt(x * x + y * y + z * z) + 1e-300
    th = np.arccos(np.clip(z / r, -1, 1))
    ph = np.arctan2(y, x)
    # 2Y_{2,2} and 2Y_{3,-1} (the reference's basis), times a radial factor 1/r
    basis = []
    for l, m in ((2, 2), (3, -1)):
        d = _wigner_d(l, m, 2, th.ravel()).reshape(th.shape)
        basis.append(math.sqrt((2 * l + 1) / (4 * np.pi)) * d * np.exp(1j * m * ph))
    psi = (0.7 * basis[0] + (0.2 - 0.3j) * basis[1]) * 3.0 / r
    rc, rp, ip = orc.surf_wave(psi.real.copy(), psi.imag.copy(), X, Y, Z, 3.0, 20, 4)
    assert rc == 0
    got = rp + 1j * ip
    idx = {lm: k for k, lm in enumerate((l, m) for l in range(2, 5) for m in range(-l, l + 1))}
    assert abs(got[idx[(2, 2)]] - 0.7 * 3.0) < 2e-3
    assert abs(got[idx[(3, -1)]] - (0.2 - 0.3j) * 3.0) < 2e-3
    others = [abs(got[k]) for lm, k in idx.items() if lm not in ((2, 2), (3, -1))]
    assert max(others) < 2e-3


def test_surf_wave_refuses_a_sphere_outside_the_block(orc):
    X, Y, Z = _grid(0)
    R, I = _fields(*_mesh(X, Y, Z), 0)
    rc, _, _ = orc.surf_wave(R, I, X, Y, Z, 4.5, 8, 2)
    assert rc == 2


def test_global_interp_exact_on_quintics(orc):
    X, Y, Z = _grid(0, n_int=16, h=0.3)
    x, y, z = _mesh(X, Y, Z)
    f = (1 + x - 2 * x ** 5) * (0.5 - y ** 3 + 0.1 * y ** 4) * (2 + z ** 2 - z ** 5 / 7)
    rng = np.random.default_rng(3)
    for _ in range(50):
        p = rng.uniform(X[3], X[-4], 3)
        rc, v = orc.global_interp(f, X, Y, Z, p)
        want = (1 + p[0] - 2 * p[0] ** 5) * (0.5 - p[1] ** 3 + 0.1 * p[1] ** 4) * (2 + p[2] ** 2 - p[2] ** 5 / 7)
        assert rc == 0 and abs(v - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("sym", [0, 1, 2])
def test_global_interp_bitwise(orc, ref, sym):
    X, Y, Z = _grid(sym, n_int=16, h=0.3)
    rng = np.random.default_rng(10 + sym)
    f = rng.standard_normal((len(Z), len(Y), len(X)))
    soa = (-1.0, 1.0, -1.0)
    lo = [A[0] if (sym == 2 or (a == 2 and sym == 1)) else A[3] for a, A in enumerate((X, Y, Z))]
    hi = [A[-4] for A in (X, Y, Z)]
    for _ in range(60):
        p = [rng.uniform(lo[a], hi[a]) for a in range(3)]
        a = orc.global_interp(f, X, Y, Z, p, soa, sym)
        b = ref.global_interp(f, X, Y, Z, p, soa, sym)
        assert a[0] == 0 and a[1] == b[1], (p, a, b)
