"""The oracle (C restatement) against the reference Fortran build, bit for bit.

Runs only where oracle/_ref was built from the reference sources.  The
reference build is configured with GAUGE 2 (its checked-in macrodef.fh);
the oracle's GAUGE 0 is pinned separately (test_oracle_golden.py and the
closed-form eta test in test_oracle_kat.py).
"""
import numpy as np
import pytest

from helpers import grid, random_state

PARS = [(1, 1, 1), (-1, -1, 1), (-1, 1, -1), (1, -1, -1), (-1, 1, 1), (1, 1, -1)]


def _grids(sym, n=(9, 10, 11)):
    nx, ny, nz = n
    if sym == 0:
        X, Y, Z, _ = grid(nx, ny, nz)
    else:  # grid starting at the symmetry plane(s)
        X = (np.arange(nx) + 0.5) * 0.3
        Y = (np.arange(ny) + 0.5) * 0.27
        Z = (np.arange(nz) + 0.5) * 0.33
    return X, Y, Z


@pytest.mark.parametrize("sym", [0, 1, 2])
@pytest.mark.parametrize("par", PARS)
def test_stencils_bitwise(orc, ref, sym, par):
    X, Y, Z = _grids(sym)
    rng = np.random.default_rng(hash((sym, par)) % 2**32)
    shape = (len(Z), len(Y), len(X))
    f = rng.standard_normal(shape)
    r0 = rng.standard_normal(shape)
    b = [rng.standard_normal(shape) for _ in range(3)]
    for a, c in zip(orc.fderivs(f, X, Y, Z, par, sym), ref.fderivs(f, X, Y, Z, par, sym)):
        assert np.array_equal(a, c)
    for a, c in zip(orc.fdderivs(f, X, Y, Z, par, sym), ref.fdderivs(f, X, Y, Z, par, sym)):
        assert np.array_equal(a, c)
    assert np.array_equal(orc.lopsided(f, r0, *b, X, Y, Z, par, sym), ref.lopsided(f, r0, *b, X, Y, Z, par, sym))
    assert np.array_equal(orc.kodis(f, r0, X, Y, Z, par, 0.25, sym), ref.kodis(f, r0, X, Y, Z, par, 0.25, sym))


def test_symmetry_bd_matches_reference_halo(ref):
    import oracle
    orc_lib = oracle.Orc()
    rng = np.random.default_rng(3)
    f = rng.standard_normal((6, 7, 8))
    import ctypes as C
    for ordn in (2, 3):
        for par in PARS[:3]:
            a = ref.symmetry_bd(ordn, f, par)
            b = np.zeros_like(a)
            soa = (C.c_double * 3)(*par)
            orc_lib.lib.orc_symmetry_bd(C.c_int(ordn), oracle._ex(f.shape), oracle._ptr(f), oracle._ptr(b), soa)
            assert np.array_equal(a, b)


@pytest.mark.parametrize("sym", [0, 1, 2])
@pytest.mark.parametrize("co", [0, 1])
def test_rhs_bitwise(orc, ref, sym, co):
    X, Y, Z = _grids(sym, (10, 11, 12))
    st = random_state((len(Z), len(Y), len(X)), seed=40 + sym * 2 + co)
    rng = np.random.default_rng(9)
    mat = [0.01 * rng.standard_normal(st[0].shape) for _ in range(10)]
    a = orc.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=sym, gauge=2, matter=mat)
    b = ref.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=sym, matter=mat)
    assert a["gont"] == b["gont"] == 0
    for key in ("rhs", "gam", "ric", "res"):
        for i, (x, y) in enumerate(zip(a[key], b[key])):
            assert np.array_equal(x, y), (key, i)


def test_rhs_nan_flag(orc, ref):
    X, Y, Z = _grids(0, (8, 8, 8))
    st = random_state((8, 8, 8), seed=1)
    st[12][3, 3, 3] = np.nan
    assert orc.compute_rhs(st, X, Y, Z, 0.02, gauge=2)["gont"] == 1
    assert ref.compute_rhs(st, X, Y, Z, 0.02)["gont"] == 1
    st = random_state((8, 8, 8), seed=1)
    st[22][3, 3, 3] = np.nan  # dtSf is outside the reference's NaN test
    assert orc.compute_rhs(st, X, Y, Z, 0.02, gauge=2)["gont"] == 0
    assert ref.compute_rhs(st, X, Y, Z, 0.02)["gont"] == 0


def test_pointwise_routines_bitwise(orc, ref):
    X, Y, Z, bbox = grid(9, 10, 11)
    shape = (11, 10, 9)
    st = random_state(shape, seed=5)
    a, b = [x.copy() for x in st], [x.copy() for x in st]
    orc.enforce_ga(a)
    ref.enforce_ga(b)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    rng = np.random.default_rng(6)
    for stage in range(4):
        f0, f1, r = (rng.standard_normal(shape) for _ in range(3))
        a1, ar, b1, br = f1.copy(), r.copy(), f1.copy(), r.copy()
        orc.rk4(f0, a1, ar, 0.07, stage)
        ref.rk4(f0, b1, br, 0.07, stage)
        assert np.array_equal(a1, b1) and np.array_equal(ar, br)
    c = rng.standard_normal(shape)
    c1, c2 = c.copy(), c.copy()
    orc.lowerboundset(c1, -0.2)
    ref.lowerboundset(c2, -0.2)
    assert np.array_equal(c1, c2)
    for bb, gw in ((bbox, 3), ((-5, -5, -5, 5, 5, 5), 3), (bbox, 1)):
        assert orc.l2norm(c, X, Y, Z, bb, gw) == pytest.approx(ref.l2norm(c, X, Y, Z, bb, gw), rel=1e-14)
    f0 = rng.standard_normal(shape)
    for sym in (0, 1, 2):
        g1, g2 = c.copy(), c.copy()
        orc.sommerfeld_copy(f0, g1, X, Y, Z, bbox, sym)
        ref.sommerfeld_copy(f0, g2, X, Y, Z, bbox, sym)
        assert np.array_equal(g1, g2)


@pytest.mark.parametrize("sym", [1, 2])
@pytest.mark.parametrize("bc", [1, 2, 3])
def test_step_with_symmetry_planes_bitwise(orc, ref, sym, bc):
    """Step on a block starting at the symmetry planes: mirrored stencils,
    boundary layers skipping the plane faces."""
    import oracle
    X, Y, Z = ((np.arange(n) + 0.5) * 0.3 for n in (12, 13, 11))
    bbox = (0.0, 0.0, 0.0, 12 * 0.3, 13 * 0.3, 11 * 0.3)
    st = random_state((11, 13, 12), seed=90 + sym)
    U1 = [x.copy() for x in st]
    U2 = [x.copy() for x in st]
    for _ in range(2):
        g1, _ = orc.step(U1, X, Y, Z, bbox, 0.075, 0.02, symmetry=sym, gauge=2, bc=bc)
        g2, _ = oracle.step_driver(ref, U2, X, Y, Z, bbox, 0.075, 0.02, 1e-5, sym, bc, 3, False)
        assert g1 == g2 == 0
    for a, b in zip(U1, U2):
        assert np.array_equal(a, b)


def test_prolong3_restrict3_bitwise(orc, ref):
    """prolongrestrict_cell.f90 (Cell, ghost width 3) on random 2:1 pairs and
    regions; the reference is called only on regions the oracle accepts
    (restrict3 stops the process on a bad one)."""
    from helpers import transfer_cases
    np_, nr = 0, 0
    for c in transfer_cases(5, 150):
        a = c["funf"].copy()
        if orc.prolong3(c["func"], c["llbc"], c["uubc"], a, c["llbf"], c["uubf"], c["llbp"], c["uubp"],
                        c["soa"]) == 0:
            b = c["funf"].copy()
            ref.prolong3(c["func"], c["llbc"], c["uubc"], b, c["llbf"], c["uubf"], c["llbp"], c["uubp"], c["soa"])
            assert np.array_equal(a, b) and not np.array_equal(a, c["funf"])
            np_ += 1
        if c["llbr"] is not None:
            a = c["func"].copy()
            if orc.restrict3(a, c["llbc"], c["uubc"], c["funf"], c["llbf"], c["uubf"], c["llbr"], c["uubr"],
                             c["soa"]) == 0:
                b = c["func"].copy()
                ref.restrict3(b, c["llbc"], c["uubc"], c["funf"], c["llbf"], c["uubf"], c["llbr"], c["uubr"],
                              c["soa"])
                assert np.array_equal(a, b) and not np.array_equal(a, c["func"])
                nr += 1
    assert np_ > 80 and nr > 80, (np_, nr)


def _bam_cases():
    """(X, Y, Z, bbox, symmetry): a box whose faces all lie on bbox and
    straddle the origin, boxes starting at the symmetry planes, and a box
    with some faces inside bbox."""
    X, Y, Z, bbox = grid(9, 10, 11)
    yield X, Y, Z, bbox, 0
    P = [(np.arange(n) + 0.5) * h for n, h in ((9, 0.3), (10, 0.27), (11, 0.33))]
    pb = (0.0, 0.0, 0.0, 9 * 0.3, 10 * 0.27, 11 * 0.33)
    for sym in (0, 1, 2):
        yield P[0], P[1], P[2], pb, sym
    yield X, Y, Z, (bbox[0] - 1.0, bbox[1], bbox[2] - 1.0, bbox[3], bbox[4] + 1.0, bbox[5]), 0


def test_sommerfeld_routbam_bitwise(orc, ref):
    """sommerfeld_routbam: every face layer, both wave directions per axis,
    every one-sided / centred fallback, mirrored reads at symmetry planes."""
    import oracle
    rng = np.random.default_rng(12)
    for X, Y, Z, bbox, sym in _bam_cases():
        shape = (len(Z), len(Y), len(X))
        f0 = rng.standard_normal(shape)
        r0 = rng.standard_normal(shape)
        for vel in (1.0, 2 ** 0.5):
            for par in PARS:
                a, b = r0.copy(), r0.copy()
                orc.sommerfeld_bam(a, f0, X, Y, Z, bbox, vel, par, sym)
                ref.sommerfeld_bam(b, f0, X, Y, Z, bbox, vel, par, sym)
                assert np.array_equal(a, b), (sym, vel, par)
                assert not np.array_equal(a, r0)
    assert oracle.PROPSPEED[0] == oracle.PROPSPEED[1] == oracle.PROPSPEED[17] == np.sqrt(2.0)


@pytest.mark.parametrize("bc", [0, 1, 2, 3])
def test_step_bitwise(orc, ref, bc):
    """The oracle's Step (C) against bssn_class::Step driven over the
    reference routines, three RK4 steps."""
    import oracle
    X, Y, Z, bbox = grid(13, 12, 14, h=(0.3, 0.3, 0.3), origin=(-2.0, -1.8, -2.1))
    st = random_state((14, 12, 13), seed=77)
    U1 = [x.copy() for x in st]
    U2 = [x.copy() for x in st]
    for _ in range(3):
        g1, h1 = orc.step(U1, X, Y, Z, bbox, 0.075, 0.02, gauge=2, bc=bc, want_ham=True)
        g2, h2 = oracle.step_driver(ref, U2, X, Y, Z, bbox, 0.075, 0.02, 1e-5, 0, bc, 3, True)
        assert g1 == g2 == 0
        assert np.array_equal(h1, h2)
    for a, b in zip(U1, U2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dT", [0.05, 0.6])
def test_sommerfeld_interpolating_predictor_bitwise(orc, ref, dT):
    """sommerfeld_rout with precor == 0: the outgoing-wave value interpolated
    from f0 at the point moved inwards by alpha sqrt(chi) dT along the ray
    (6^3 Neville block, clamped at the array's edges, mirrored below a
    symmetry plane), on every face layer; small and large displacements."""
    rng = np.random.default_rng(21)
    for X, Y, Z, bbox, sym in _bam_cases():
        shape = (len(Z), len(Y), len(X))
        f0 = rng.standard_normal(shape)
        chi0 = -0.3 + 0.05 * rng.standard_normal(shape)
        lap0 = -0.2 + 0.05 * rng.standard_normal(shape)
        start = rng.standard_normal(shape)
        for par in PARS[:3]:
            a, b = start.copy(), start.copy()
            assert orc.sommerfeld_interp(f0, a, chi0, lap0, X, Y, Z, bbox, dT, par, sym) == 0
            ref.sommerfeld_interp(f0, b, chi0, lap0, X, Y, Z, bbox, dT, par, sym)
            assert np.array_equal(a, b), (sym, par)
            assert not np.array_equal(a, start)
