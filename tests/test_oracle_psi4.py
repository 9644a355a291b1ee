"""Psi4 (getnp4, tetradtype 2) of the oracle: bit for bit against the
reference build, and against closed forms the mathematics fixes.

Closed form: with a constant conformal metric g, a state whose chi, trK and
A_ij are linear in the coordinates, and arbitrary Christoffels and Ricci
inputs, every first derivative the routine takes is exact (centred
differences of either order are exact on linear data), so

    chi D_k K_ij = b_ijk - G^m_ki A_mj - G^m_kj A_im - (c_k / chi) A_ij + g_ij k_k / 3,
    E_ij = R_ij - (K_ik g^kl K_lj - K_ij K) / chi - v^k (T_ijk - (T_ikj + T_jki) / 2) / chi,
    Psi4 = -E_ij (u^i u^j - w^i w^j) / 2 - i E_ij (u^i w^j + w^i u^j) / 2,

with chi K_ij = A_ij + g_ij K / 3 and (v, u, w) the Gram-Schmidt frame of
(g^-1 x, (-y, x, 0), (xz, yz, -x^2 - y^2)) in the metric g / chi.  Points
without +-1 room on every axis (and the last index of each axis) take zero
derivatives.  The expected values are computed here with tensor contractions
and an explicit Gram-Schmidt, independently of the oracle's component code.
Conformally flat, time-symmetric data reduce to Psi4 = -chi (R_pp - R_tt) / 2
- i chi R_pt in the spherical unit vectors p = phi-hat, t = theta-hat.
"""
import numpy as np
import pytest

from helpers import grid, random_state

SIDX = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def _full(sym6):
    """(6, ...) symmetric components -> (3, 3, ...)."""
    out = np.empty((3, 3) + sym6[0].shape)
    for s, (a, b) in enumerate(SIDX):
        out[a, b] = out[b, a] = sym6[s]
    return out


def _gram_schmidt(g, chi, X, Y, Z):
    """The frame (v, u, w) of getnp4's tetradtype 2 in the metric g / chi
    (g: (3, 3, ...)), points off the z axis."""
    Zg, Yg, Xg = np.meshgrid(Z, Y, X, indexing="ij")
    r = np.stack([Xg, Yg, Zg])
    gi = np.moveaxis(np.linalg.inv(np.moveaxis(g, (0, 1), (-2, -1))), (-2, -1), (0, 1))
    v = np.einsum("ab...,b...->a...", gi, r)
    u = np.stack([-Yg, Xg, np.zeros_like(Xg)])
    w = np.stack([Xg * Zg, Yg * Zg, -(Xg ** 2 + Yg ** 2)])

    def dot(a, b):
        return np.einsum("ab...,a...,b...->...", g, a, b) / chi

    v = v / np.sqrt(dot(v, v))
    u = u - dot(v, u) * v
    u = u / np.sqrt(dot(u, u))
    w = w - dot(v, w) * v
    w = w - dot(u, w) * u
    w = w / np.sqrt(dot(w, w))
    return v, u, w


def _linear_case(seed, n=(9, 10, 11)):
    rng = np.random.default_rng(seed)
    X, Y, Z, _ = grid(*n)
    shape = (len(Z), len(Y), len(X))
    Zg, Yg, Xg = np.meshgrid(Z, Y, X, indexing="ij")
    xs = np.stack([Xg, Yg, Zg])
    st = [np.zeros(shape) for _ in range(24)]
    g = np.eye(3) + 0.1 * rng.standard_normal((3, 3))
    g = 0.5 * (g + g.T)
    for s, (a, b) in enumerate(SIDX):
        st[2 + s][:] = g[a, b] - (1.0 if a == b else 0.0)
    c0, c = 0.7, 0.05 * rng.standard_normal(3)
    st[0][:] = c0 - 1.0 + np.einsum("k,k...->...", c, xs)
    k0, kk = 0.1, 0.05 * rng.standard_normal(3)
    st[1][:] = k0 + np.einsum("k,k...->...", kk, xs)
    a0 = 0.1 * rng.standard_normal((3, 3))
    a0 = a0 + a0.T
    b = 0.1 * rng.standard_normal((3, 3, 3))
    b = b + b.transpose(1, 0, 2)
    for s, (i, j) in enumerate(SIDX):
        st[8 + s][:] = a0[i, j] + np.einsum("k,k...->...", b[i, j], xs)
    gam = [0.3 * rng.standard_normal(shape) for _ in range(18)]
    ric = [0.3 * rng.standard_normal(shape) for _ in range(6)]
    return X, Y, Z, st, gam, ric, g, c, kk, b


def _expected_linear(X, Y, Z, st, gam, ric, g, c, kk, b):
    shape = st[0].shape
    chi = st[0] + 1.0
    K = st[1]
    A = _full(st[8:14])
    G = np.stack([_full(gam[6 * m:6 * m + 6]) for m in range(3)])  # G[m, k, i] = Gamma^m_ki
    R = _full(ric)
    gg = np.broadcast_to(g[:, :, None, None, None], (3, 3) + shape)
    gi = np.linalg.inv(g)
    # room: +-1 on every axis, not the last index
    nz, ny, nx = shape
    room = np.zeros(shape, bool)
    room[1:-1, 1:-1, 1:-1] = True
    T = (np.broadcast_to(b[:, :, :, None, None, None], (3, 3, 3) + shape)
         - np.einsum("mki...,mj...->ijk...", G, A) - np.einsum("mkj...,im...->ijk...", G, A)
         - np.einsum("k,ij...->ijk...", c, A) / chi + np.einsum("ij...,k->ijk...", gg, kk) / 3.0)
    # where the derivatives are zero, T keeps only the Christoffel terms
    T0 = -np.einsum("mki...,mj...->ijk...", G, A) - np.einsum("mkj...,im...->ijk...", G, A)
    T = np.where(room, T, T0)
    v, u, w = _gram_schmidt(gg, chi, X, Y, Z)
    Bm = np.einsum("k...,ijk...->ij...", v, T - 0.5 * (T.transpose(0, 2, 1, 3, 4, 5) + T.transpose(2, 1, 0, 3, 4, 5)))
    Kh = A + gg * K / 3.0
    KK = np.einsum("ik...,kl,lj...->ij...", Kh, gi, Kh)
    E = R - (KK - Kh * K) / chi - Bm / chi
    re = -0.5 * np.einsum("ij...,ij...->...", E, np.einsum("i...,j...->ij...", u, u) - np.einsum("i...,j...->ij...", w, w))
    im = -0.5 * np.einsum("ij...,ij...->...", E, np.einsum("i...,j...->ij...", u, w) + np.einsum("i...,j...->ij...", w, u))
    return re, im


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_linear_data_closed_form(orc, seed):
    X, Y, Z, st, gam, ric, g, c, kk, b = _linear_case(seed)
    rp, ip = orc.getnp4(st, gam, ric, X, Y, Z)
    er, ei = _expected_linear(X, Y, Z, st, gam, ric, g, c, kk, b)
    for got, want in ((rp, er), (ip, ei)):
        scale = float(np.max(np.abs(want)))
        assert scale > 1e-2
        assert np.max(np.abs(got - want)) <= 1e-12 * scale


def test_time_symmetric_conformally_flat_closed_form(orc):
    X, Y, Z, _ = grid(8, 9, 10)
    shape = (len(Z), len(Y), len(X))
    rng = np.random.default_rng(5)
    st = [np.zeros(shape) for _ in range(24)]
    st[0][:] = -0.4 + 0.1 * rng.random(shape)  # chi - 1, point by point
    gam = [rng.standard_normal(shape) for _ in range(18)]  # multiplies A = 0
    ric = [rng.standard_normal(shape) for _ in range(6)]
    rp, ip = orc.getnp4(st, gam, ric, X, Y, Z)
    Zg, Yg, Xg = np.meshgrid(Z, Y, X, indexing="ij")
    rho = np.hypot(Xg, Yg)
    r = np.sqrt(rho ** 2 + Zg ** 2)
    ph = np.stack([-Yg / rho, Xg / rho, np.zeros_like(rho)])
    th = np.stack([Xg * Zg / (r * rho), Yg * Zg / (r * rho), -rho / r])
    R = _full(ric)
    Rpp = np.einsum("i...,ij...,j...->...", ph, R, ph)
    Rtt = np.einsum("i...,ij...,j...->...", th, R, th)
    Rpt = np.einsum("i...,ij...,j...->...", ph, R, th)
    chi = st[0] + 1.0
    assert np.max(np.abs(rp - (-0.5 * chi * (Rpp - Rtt)))) <= 1e-13 * np.max(np.abs(rp))
    assert np.max(np.abs(ip - (-chi * Rpt))) <= 1e-13 * np.max(np.abs(ip))


def _sym_grids(sym, n):
    nx, ny, nz = n
    if sym == 0:
        X, Y, Z, _ = grid(nx, ny, nz)
    else:
        X = (np.arange(nx) + 0.5) * 0.3
        Y = (np.arange(ny) + 0.5) * 0.27
        Z = (np.arange(nz) + 0.5) * 0.33
    return X, Y, Z


def _axis_grid(n=(9, 9, 9), h=0.25):
    """Vertex-centred grid through the origin: the z axis and the origin are
    grid points (the tetrad's stand-in branches)."""
    c = (np.arange(n[0]) - n[0] // 2) * h
    return c.copy(), c.copy(), c.copy()


@pytest.mark.parametrize("sym", [0, 1, 2, "axis"])
def test_getnp4_bitwise(orc, ref, sym):
    if sym == "axis":
        X, Y, Z = _axis_grid()
        sym = 0
    else:
        X, Y, Z = _sym_grids(sym, (10, 11, 12))
    st = random_state((len(Z), len(Y), len(X)), seed=70 + (sym if isinstance(sym, int) else 9))
    aux = ref.compute_rhs(st, X, Y, Z, 0.02, symmetry=sym)
    a = orc.getnp4(st, aux["gam"], aux["ric"], X, Y, Z, symmetry=sym)
    b = ref.getnp4(st, aux["gam"], aux["ric"], X, Y, Z, symmetry=sym)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
        assert np.all(np.isfinite(x)) and np.max(np.abs(x)) > 0
