"""Closed-form checks of the oracle, chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails: polynomial exactness of every
stencil, the KO operator on x^6, the RK4 amplification factor, the algebraic
constraints after enforce_ga, the RHS of flat space and of a pure B^i
field, reflection symmetry of the puncture RHS, and Philox known answers."""
import numpy as np
import pytest

from helpers import grid, random_state


def mesh(X, Y, Z):
    z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
    return x, y, z


def test_fderivs_exact_on_quartics(orc):
    X, Y, Z, _ = grid(11, 12, 13)
    x, y, z = mesh(X, Y, Z)
    f = x ** 3 - 2 * x ** 2 * y + y ** 4 + z ** 4 + x * y * z
    fx, fy, fz = orc.fderivs(f, X, Y, Z, (1, 1, 1))
    ex = (3 * x ** 2 - 4 * x * y + y * z, -2 * x ** 2 + 4 * y ** 3 + x * z, 4 * z ** 3 + x * y)
    s = np.s_[2:-3, 2:-3, 2:-3]  # 4th-order points (+-2 room, last index excluded)
    for a, b in zip((fx, fy, fz), ex):
        assert np.max(np.abs(a[s] - b[s])) < 1e-11 * np.max(np.abs(b[s]))
    # 2nd-order points (exactly +-1 room on some axis) are exact on quadratics
    q = 2 * x ** 2 - x * y + 3 * z ** 2
    qx, qy, qz = orc.fderivs(q, X, Y, Z, (1, 1, 1))
    edge = np.s_[1, 1:-1, 1:-1]
    assert np.allclose(qx[edge], (4 * x - y)[edge], rtol=0, atol=1e-12)
    assert np.allclose(qz[edge], (6 * z)[edge], rtol=0, atol=1e-12)
    # zero where an axis lacks +-1 room, and on the last index
    assert np.all(fx[0] == 0) and np.all(fx[:, :, -1] == 0)


def test_fdderivs_exact(orc):
    X, Y, Z, _ = grid(11, 12, 13)
    x, y, z = mesh(X, Y, Z)
    f = x ** 5 + y ** 5 + z ** 4 + x ** 3 * y ** 2 + x ** 2 * z ** 3 + y ** 4 * z ** 4
    d = orc.fdderivs(f, X, Y, Z, (1, 1, 1))
    ex = (20 * x ** 3 + 6 * x * y ** 2 + 2 * z ** 3,          # xx
          6 * x ** 2 * y,                                     # xy
          6 * x * z ** 2,                                     # xz
          20 * y ** 3 + 2 * x ** 3 + 12 * y ** 2 * z ** 4,    # yy
          16 * y ** 3 * z ** 3,                               # yz
          12 * z ** 2 + 6 * x ** 2 * z + 12 * y ** 4 * z ** 2)  # zz
    s = np.s_[2:-3, 2:-3, 2:-3]
    for a, b in zip(d, ex):
        assert np.max(np.abs(a[s] - b[s])) < 1e-10 * np.max(np.abs(b[s]))


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_lopsided_exact_on_quartics(orc, sign):
    X, Y, Z, _ = grid(12, 12, 12)
    x, y, z = mesh(X, Y, Z)
    f = x ** 4 + 2 * y ** 3 * x - z ** 4 + x * y * z
    bx, by, bz = (sign * c * np.ones_like(x) for c in (0.7, -1.3, 0.4))
    r = orc.lopsided(f, np.zeros_like(f), bx, by, bz, X, Y, Z, (1, 1, 1))
    grad = (4 * x ** 3 + 2 * y ** 3 + y * z, 6 * y ** 2 * x + x * z, -4 * z ** 3 + x * y)
    ex = bx * grad[0] + by * grad[1] + bz * grad[2]
    s = np.s_[3:-3, 3:-3, 3:-3]  # full upwind room along every axis
    assert np.max(np.abs(r[s] - ex[s])) < 1e-10 * np.max(np.abs(ex[s]))


def test_kodis_annihilates_quintics_and_gives_720_h5_on_x6(orc):
    h = 0.3
    X, Y, Z, _ = grid(12, 12, 12, h=(h, h, h))
    x, y, z = mesh(X, Y, Z)
    eps = 0.3
    p5 = x ** 5 - 3 * y ** 4 * z + z ** 5 + x * y * z
    r = orc.kodis(p5, np.zeros_like(p5), X, Y, Z, (1, 1, 1), eps)
    s = np.s_[3:-3, 3:-3, 3:-3]
    assert np.max(np.abs(r[s])) < 1e-9
    r6 = orc.kodis(x ** 6, np.zeros_like(x), X, Y, Z, (1, 1, 1), eps)
    assert np.allclose(r6[s], eps / 64 * 720 * h ** 5, rtol=1e-9, atol=0)
    # only where every axis has +-3 room
    assert np.all(r6[:3] == 0) and np.all(r6[:, :, -3:] == 0)


def test_rk4_amplification_factor(orc):
    """The reference's four-call RK4 sequence (bssn_class.C:3119-3123, 3487-3491)
    on f' = lambda f reproduces 1 + z + z^2/2 + z^3/6 + z^4/24."""
    lam, dt = -0.7, 0.3
    sh = (2, 2, 2)
    U0 = np.ones(sh)
    pre, cor, acc = np.zeros(sh), np.zeros(sh), np.zeros(sh)
    acc[:] = lam * U0
    orc.rk4(U0, pre, acc, dt, 0)
    for it in (1, 2, 3):
        cor[:] = lam * pre
        orc.rk4(U0, cor, acc, dt, it)
        if it < 3:
            pre, cor = cor, pre
    zz = lam * dt
    assert np.allclose(cor, 1 + zz + zz ** 2 / 2 + zz ** 3 / 6 + zz ** 4 / 24, rtol=1e-15, atol=0)


def test_enforce_ga_restores_algebraic_constraints(orc):
    st = random_state((6, 7, 8), seed=3, amp=0.1)
    orc.enforce_ga(st)
    g = [st[2] + 1, st[3], st[4], st[5] + 1, st[6], st[7] + 1]
    det = g[0] * g[3] * g[5] + 2 * g[1] * g[4] * g[2] - g[2] ** 2 * g[3] - g[1] ** 2 * g[5] - g[0] * g[4] ** 2
    assert np.max(np.abs(det - 1)) < 1e-14
    cof = [g[3] * g[5] - g[4] ** 2, g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
           g[0] * g[5] - g[2] ** 2, g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] ** 2]
    A = st[8:14]
    tr = cof[0] * A[0] + cof[3] * A[3] + cof[5] * A[5] + 2 * (cof[1] * A[1] + cof[2] * A[2] + cof[4] * A[4])
    assert np.max(np.abs(tr)) < 1e-15


@pytest.mark.parametrize("gauge", [0, 2])
def test_flat_space_rhs_is_zero(orc, gauge):
    X, Y, Z, _ = grid(10, 10, 10)
    st = [np.zeros((10, 10, 10)) for _ in range(24)]
    if gauge == 2:  # GAUGE 2's damping is 0/0 at chi = 1; perturb chi slightly
        st[0][:] = -1e-3
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=0, gauge=gauge)
    tol = 0.0 if gauge == 0 else 1e-18  # constant chi - 1 != 0: stencil sums cancel to rounding
    for v in range(24):
        assert np.max(np.abs(o["rhs"][v])) <= tol, v
    assert np.max(np.abs(o["res"][0])) <= tol


def test_gauge0_B_driver_closed_form(orc):
    """Uniform B^i on flat space: beta^i_rhs = 3/4 B^i, B^i_rhs = -eta B^i with eta = 2."""
    X, Y, Z, _ = grid(10, 10, 10)
    st = [np.zeros((10, 10, 10)) for _ in range(24)]
    c = (0.3, -0.2, 0.05)
    for a in range(3):
        st[21 + a][:] = c[a]
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=1, gauge=0)
    s = np.s_[3:-3, 3:-3, 3:-3]
    for a in range(3):
        assert np.all(o["rhs"][18 + a][s] == 0.75 * c[a])
        assert np.all(o["rhs"][21 + a][s] == -2.0 * c[a])


def test_puncture_rhs_reflection_symmetry(orc):
    """Single puncture at the centre of a symmetric grid: each RHS field has
    the reflection parity of its variable (bssn_class.C:183-361)."""
    import oracle
    import amssbssn as A
    g = A.cube_patch(14, 4.0)
    cfg = A.Config("sym", [g], dt=0.1, steps=1, punctures=[(1.0, 0.0, 0.0, 0.0)])
    st = cfg.initial_state(0)
    rng = np.random.default_rng(0)
    # add a reflection-symmetric shift so advection and the Lie terms are exercised
    X, Y, Z = g.coords()
    x, y, z = mesh(X, Y, Z)
    bump = np.exp(-(x ** 2 + y ** 2 + z ** 2) / 4)
    st[18], st[19], st[20] = 0.1 * x * bump, 0.1 * y * bump, 0.1 * z * bump
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=1, gauge=0)
    s = np.s_[3:-3, 3:-3, 3:-3]  # [3, n-4]: full stencils, symmetric under i -> n-1-i
    for v in range(24):
        r = o["rhs"][v][s]
        m = max(float(np.max(np.abs(r))), 1e-300)
        for axis, flip in ((2, np.s_[:, :, ::-1]), (1, np.s_[:, ::-1, :]), (0, np.s_[::-1, :, :])):
            par = oracle.PARITY[v][2 - axis]
            assert np.max(np.abs(r[flip] - par * r)) <= 1e-12 * m, (v, axis)


@pytest.mark.parametrize("sym", [0, 2])
def test_sommerfeld_bam_exact_on_quadratics(orc, sym):
    """Every difference of sommerfeld_routbam (backward, centred, forward;
    second order) is exact on quadratics, so on each face layer the result
    is -v (x.grad f + f) / r of the polynomial; interior points keep their
    value.  Octant symmetry with an even polynomial: the mirrored reads at
    the planes see the same polynomial."""
    if sym == 0:
        X, Y, Z, bbox = grid(9, 10, 11)
        f_of = lambda x, y, z: 0.3 + 2 * x ** 2 - x * y + 3 * z ** 2 - 0.5 * y + z  # noqa: E731
        grad = lambda x, y, z: (4 * x - y, -x - 0.5, 6 * z + 1)  # noqa: E731
        par = (1, 1, 1)
    else:
        X, Y, Z = ((np.arange(n) + 0.5) * h for n, h in ((9, 0.3), (10, 0.27), (11, 0.33)))
        bbox = (0.0, 0.0, 0.0, 9 * 0.3, 10 * 0.27, 11 * 0.33)
        f_of = lambda x, y, z: 0.3 + 2 * x ** 2 + y ** 2 - 3 * z ** 2  # noqa: E731
        grad = lambda x, y, z: (4 * x, 2 * y, -6 * z)  # noqa: E731
        par = (1, 1, 1)
    x, y, z = mesh(X, Y, Z)
    f = f_of(x, y, z)
    r = np.sqrt(x * x + y * y + z * z)
    gx, gy, gz = grad(x, y, z)
    want = -1.3 * (gx * x + gy * y + gz * z + f) / r
    out = np.full(f.shape, 7.0)
    orc.sommerfeld_bam(out, f, X, Y, Z, bbox, 1.3, par, sym)
    layer = np.zeros(f.shape, bool)
    layer[-1], layer[:, -1], layer[:, :, -1] = True, True, True
    if sym == 0:
        layer[0], layer[:, 0], layer[:, :, 0] = True, True, True
    assert np.allclose(out[layer], want[layer], rtol=0, atol=1e-12 * np.max(np.abs(want)))
    assert np.all(out[~layer] == 7.0)


def _poly5(x, y, z):
    return (0.3 + x - 0.7 * x ** 5 + 0.2 * x ** 3 * y ** 2 + y ** 4 - 0.5 * y ** 5 * z + 1.1 * z ** 5
            - x * y * z + 0.05 * (x * y * z) ** 2)


def test_prolong3_restrict3_exact_on_degree5(orc):
    """6-point Lagrange interpolation (prolong: fine cell centres at 1/4 and
    3/4 of a coarse cell; restrict: coarse centres midway between fine ones)
    is exact for polynomials of degree <= 5 in each variable; regions kept
    off the arrays' lower edges (no mirrored reads)."""
    CD, FD = 0.5, 0.25
    extc, extf = np.array([14, 12, 13]), np.array([12, 14, 10])
    llbc = np.array([-3.5, -3.0, -3.25])
    llbf = llbc + FD * np.array([6, 7, 8])  # even and odd offsets
    cen = lambda lo, n, d: [lo[a] + (np.arange(n[a]) + 0.5) * d for a in range(3)]  # noqa: E731
    Xc, Xf = cen(llbc, extc, CD), cen(llbf, extf, FD)
    zc, yc, xc = np.meshgrid(Xc[2], Xc[1], Xc[0], indexing="ij")
    zf, yf, xf = np.meshgrid(Xf[2], Xf[1], Xf[0], indexing="ij")
    func, funf = _poly5(xc, yc, zc), _poly5(xf, yf, zf)
    # prolong onto the fine interior
    out = np.full(funf.shape, 9.0)
    assert orc.prolong3(func, llbc, llbc + CD * extc, out, llbf, llbf + FD * extf, llbf + FD * 2,
                        llbf + FD * (extf - 2), (1, 1, 1)) == 0
    s = np.s_[2:-2, 2:-2, 2:-2]
    assert np.max(np.abs(out[s] - funf[s])) < 1e-11 * np.max(np.abs(funf[s]))
    assert np.all(out[0] == 9.0)
    # restrict onto coarse cells covered by the fine data with a 3-cell margin
    out = np.full(func.shape, 9.0)
    lo, hi = llbf + 3 * FD + np.array([0.0, 0.25, 0.0]), llbf + FD * (extf - 3) - np.array([0.0, 0.25, 0.0])
    assert orc.restrict3(out, llbc, llbc + CD * extc, funf, llbf, llbf + FD * extf, lo, hi, (1, 1, 1)) == 0
    m = out != 9.0
    assert m.sum() >= 12
    assert np.max(np.abs(out[m] - func[m])) < 1e-11 * np.max(np.abs(func))


def test_philox_known_answers(orc):
    import amssbssn.grid as G
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        assert tuple(orc.philox(ctr, key)) == want
        got = G.philox4x32_10(*[np.array([c], np.uint64) for c in ctr], *key)
        assert tuple(int(w[0]) for w in got) == want


@pytest.mark.parametrize("sym,soa", [(0, (1, 1, 1)), (1, (1, 1, 1)), (1, (1, 1, -1)), (2, (-1, 1, -1))])
def test_sommerfeld_interpolating_predictor_exact_on_polynomials(orc, sym, soa):
    """sommerfeld_rout (precor 0) sets each face point to the outgoing wave
    (1 - r) f0(x - r x), r = alpha sqrt(chi) dT / |x|; its 6-point
    interpolation is exact for polynomials of degree <= 5 per axis, also when
    the block is mirrored across a symmetry plane (f0 with the parity SoA)."""
    h = (0.3, 0.27, 0.33)
    n = (11, 12, 13)
    if sym == 0:
        X, Y, Z, bbox = grid(*n, h=h, origin=(-1.6, -1.5, -2.0))
    else:  # the lower faces on the planes (z for sym 1; x, y, z for sym 2)
        org = (0.0 if sym == 2 else -1.6, 0.0 if sym == 2 else -1.5, 0.0)
        X, Y, Z, bbox = grid(*n, h=h, origin=org)
    x, y, z = mesh(X, Y, Z)
    sx, sy, sz = soa

    def par(t, s):  # degree <= 5, parity s under t -> -t
        return t ** 3 - t if s == -1 else t ** 4 - 2 * t ** 2 + 0.7

    def P(x, y, z):  # degree <= 5 per axis, parity SoA on the mirrored axes
        if sym == 0:
            return x ** 5 - 2 * x ** 3 * y ** 2 + 0.5 * y ** 4 - 0.3 * y * z + z ** 5 - x * y * z + 1.0
        if sym == 1:
            return (x ** 5 - x * y ** 3 + 0.4 * y ** 2 + 1.0) * par(z, sz)
        return par(x, sx) * par(y, sy) * par(z, sz)

    f0 = P(x, y, z)
    rng = np.random.default_rng(31)
    chi0 = -0.3 + 0.05 * rng.standard_normal(f0.shape)
    lap0 = -0.2 + 0.05 * rng.standard_normal(f0.shape)
    dT = 0.4
    f = np.full(f0.shape, 1e30)
    assert orc.sommerfeld_interp(f0, f, chi0, lap0, X, Y, Z, bbox, dT, soa, sym) == 0
    r = (lap0 + 1) * np.sqrt(1 + chi0) * dT / np.sqrt(x * x + y * y + z * z)
    want = (1 - r) * P((1 - r) * x, (1 - r) * y, (1 - r) * z)
    written = f != 1e30
    face = np.zeros(f0.shape, bool)
    face[-1], face[:, -1], face[:, :, -1] = True, True, True
    if sym < 2:
        face[:, :, 0] = face[:, 0] = True
    if sym == 0:
        face[0] = True
    assert np.array_equal(written, face)
    scale = np.max(np.abs(want[face]))
    assert np.max(np.abs(f[face] - want[face])) < 1e-11 * scale
    assert np.max(r[face]) * np.max(np.abs(np.stack([x, y, z])[:, face]) / 0.27) > 1.0  # moves > 1 cell
