"""GPU parity of the HIP path against the oracle (C restatement of the
reference Fortran, itself pinned to the reference build bit for bit).

All calls go through the C ABI of libamssbssn.so.  Metric: per field,
max |gpu - oracle| / max |oracle| (absolute when the oracle field is zero).
Tolerances: 1e-12 per field for one RHS call and one RK4 step, 1e-9 after
many steps.  The GPU evaluates the same algebra in
a different operation order (tensor-index form, FMA contraction), so
agreement is to rounding, not bitwise.
"""
import numpy as np
import pytest

from helpers import grid, random_state, rel_err

pytestmark = pytest.mark.gpu

TOL_RHS = 1e-12


@pytest.fixture(scope="module")
def fa():
    import amssbssn.fortran_abi as fa
    import amssbssn
    amssbssn.lib()
    return fa


def _cmp_rhs(a, b, co, label, tol=TOL_RHS):
    worst = {}
    for key in ("rhs", "gam", "ric") + (("res",) if co == 0 else ()):
        for i, (x, y) in enumerate(zip(a[key], b[key])):
            e = rel_err(x, y)
            worst[f"{key}[{i}]"] = e
    bad = {k: v for k, v in worst.items() if v > tol}
    assert not bad, f"{label}: fields over {tol}: {bad}"
    return max(worst.values())


@pytest.mark.parametrize("gauge", [0, 2])
@pytest.mark.parametrize("co", [0, 1])
def test_rhs_random_state_matches_oracle(fa, orc, gauge, co):
    """Whole array, including the reference's edge semantics (degraded
    stencils, lopsided fallbacks for both shift signs, last index)."""
    nx, ny, nz = 13, 14, 15
    X, Y, Z, _ = grid(nx, ny, nz)
    st = random_state((nz, ny, nx), seed=10 + gauge + co)
    fa.set_gauge(gauge)
    g = fa.compute_rhs(st, X, Y, Z, 0.02, co=co)
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=co, gauge=gauge)
    assert g["gont"] == 0 and o["gont"] == 0
    worst = _cmp_rhs(g, o, co, f"gauge {gauge} co {co}")
    print(f"gauge {gauge} co {co}: worst rel err {worst:.2e}")


def test_rhs_with_matter_matches_oracle(fa, orc):
    nx, ny, nz = 11, 12, 10
    X, Y, Z, _ = grid(nx, ny, nz)
    st = random_state((nz, ny, nx), seed=3)
    rng = np.random.default_rng(4)
    mat = [0.01 * rng.standard_normal((nz, ny, nx)) for _ in range(10)]
    fa.set_gauge(0)
    g = fa.compute_rhs(st, X, Y, Z, 0.02, co=0, matter=mat)
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=0, gauge=0, matter=mat)
    _cmp_rhs(g, o, 0, "matter")


def test_rhs_nan_input_returns_gont_and_writes_nothing(fa):
    nx, ny, nz = 9, 9, 9
    X, Y, Z, _ = grid(nx, ny, nz)
    st = random_state((nz, ny, nx), seed=5)
    st[5][4, 4, 4] = np.nan
    fa.set_gauge(0)
    g = fa.compute_rhs(st, X, Y, Z, 0.02, co=0)
    assert g["gont"] == 1
    assert all(np.all(a == 0) for a in g["rhs"])


def test_rhs_flat_space_is_zero(fa):
    nx, ny, nz = 12, 12, 12
    X, Y, Z, _ = grid(nx, ny, nz)
    st = [np.zeros((nz, ny, nx)) for _ in range(24)]
    fa.set_gauge(0)
    g = fa.compute_rhs(st, X, Y, Z, 0.02, co=0)
    assert max(float(np.max(np.abs(a))) for a in g["rhs"]) == 0.0


def _plane_grid(n, h, at_plane):
    """Cell-centred coordinates starting half a spacing above a symmetry
    plane, or well away from it."""
    return (np.arange(n) + 0.5) * h if at_plane else -1.7 + np.arange(n) * h


@pytest.mark.parametrize("sym", [1, 2])
@pytest.mark.parametrize("gauge", [0, 2])
@pytest.mark.parametrize("co", [0, 1])
def test_rhs_symmetry_planes_match_oracle(fa, orc, sym, gauge, co):
    """Equatorial (z) and octant (x, y, z) symmetry: stencils at the lower
    faces read mirror points with each field's parity, every degraded-order
    and lopsided branch moving with the lowered bound."""
    nx, ny, nz = 11, 12, 13
    X, Y, Z = _plane_grid(nx, 0.3, True), _plane_grid(ny, 0.27, True), _plane_grid(nz, 0.33, True)
    st = random_state((nz, ny, nx), seed=60 + 4 * sym + 2 * gauge + co)
    fa.set_gauge(gauge)
    g = fa.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=sym)
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=sym, gauge=gauge)
    assert g["gont"] == 0 and o["gont"] == 0
    worst = _cmp_rhs(g, o, co, f"symmetry {sym} gauge {gauge} co {co}")
    print(f"symmetry {sym} gauge {gauge} co {co}: worst rel err {worst:.2e}")
    # the planes change the lower-face values only
    z = orc.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=0, gauge=gauge)
    assert not np.array_equal(z["rhs"][0][0], o["rhs"][0][0])


@pytest.mark.parametrize("shape,sym", [((5, 6, 23), 0), ((9, 5, 7), 0), ((5, 5, 5), 0), ((6, 7, 6), 0),
                                       ((3, 7, 5), 2), ((4, 4, 6), 2), ((8, 6, 3), 1)])
def test_rhs_ragged_and_minimal_shapes_match_oracle(fa, orc, shape, sym):
    """Thin or empty interiors (the boundary kernels' stencil-order shells
    then cover most or all points), the smallest axes the entry point accepts
    (5 points, 3 on a symmetry plane), odd row pitches."""
    nx, ny, nz = shape
    X = _plane_grid(nx, 0.3, sym > 1)
    Y = _plane_grid(ny, 0.27, sym > 1)
    Z = _plane_grid(nz, 0.33, sym > 0)
    st = random_state((nz, ny, nx), seed=200 + nx + 7 * ny + 31 * nz + sym)
    fa.set_gauge(0)
    for co in (0, 1):
        g = fa.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=sym)
        o = orc.compute_rhs(st, X, Y, Z, 0.02, co=co, symmetry=sym, gauge=0)
        assert g["gont"] == 0 and o["gont"] == 0
        _cmp_rhs(g, o, co, f"shape {shape} symmetry {sym} co {co}")


@pytest.mark.parametrize("faces", [(False, False, True), (True, False, True), (False, True, False)])
def test_rhs_octant_with_some_faces_off_the_plane(fa, orc, faces):
    """Symmetry 2 mirrors only the axes whose first coordinate is within one
    spacing of 0."""
    nx, ny, nz = 10, 9, 11
    X, Y, Z = (_plane_grid(n, hh, on) for n, hh, on in zip((nx, ny, nz), (0.3, 0.27, 0.33), faces))
    st = random_state((nz, ny, nx), seed=71)
    fa.set_gauge(0)
    g = fa.compute_rhs(st, X, Y, Z, 0.02, co=0, symmetry=2)
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=0, symmetry=2, gauge=0)
    _cmp_rhs(g, o, 0, f"octant faces {faces}")


@pytest.mark.parametrize("sym", [0, 2])
def test_device_rhs_entry_point_matches_oracle(fa, orc, sym):
    """amss_rhs_bssn_dev on device arrays, with matter and the residuals."""
    import torch
    nx, ny, nz = 12, 10, 11
    X, Y, Z = _plane_grid(nx, 0.3, sym > 0), _plane_grid(ny, 0.27, sym > 0), _plane_grid(nz, 0.33, sym > 0)
    st = random_state((nz, ny, nx), seed=80 + sym)
    rng = np.random.default_rng(81)
    mat = [0.01 * rng.standard_normal((nz, ny, nx)) for _ in range(10)]
    dev = torch.device("cuda:0")
    S = torch.from_numpy(np.stack(st)).to(dev)
    M = torch.from_numpy(np.stack(mat)).to(dev)
    g = fa.compute_rhs_dev(S, (X[1] - X[0], Y[1] - Y[0], Z[1] - Z[0]), 0.02, co=0, symmetry=sym,
                           origin=(X[0], Y[0], Z[0]), gauge=0, matter=M)
    o = orc.compute_rhs(st, X, Y, Z, 0.02, co=0, symmetry=sym, gauge=0, matter=mat)
    assert g["flag"] == 0
    gh = {k: [a.cpu().numpy() for a in g[k]] for k in ("rhs", "gam", "ric", "res")}
    _cmp_rhs(gh, o, 0, f"device symmetry {sym}")


def test_pointwise_abi_entry_points_match_oracle(fa, orc):
    nx, ny, nz = 10, 11, 12
    X, Y, Z, bbox = grid(nx, ny, nz)
    rng = np.random.default_rng(7)
    st = random_state((nz, ny, nx), seed=8)
    a = [x.copy() for x in st]
    b = [x.copy() for x in st]
    fa.enforce_ga(a)
    orc.enforce_ga(b)
    for x, y in zip(a, b):
        assert rel_err(x, y) < 1e-14
    for stage in range(4):
        f0, f1, r = (rng.standard_normal((nz, ny, nx)) for _ in range(3))
        g1, gr, o1, orr = f1.copy(), r.copy(), f1.copy(), r.copy()
        assert fa.rk4(f0, g1, gr, 0.1, stage) == 0
        orc.rk4(f0, o1, orr, 0.1, stage)
        assert rel_err(g1, o1) < 1e-15 and rel_err(gr, orr) < 1e-15
    assert fa.rk4(f0, g1, gr, 0.1, 7) != 0
    chi = rng.standard_normal((nz, ny, nx))
    c1, c2 = chi.copy(), chi.copy()
    fa.lowerboundset(c1, 0.1)
    orc.lowerboundset(c2, 0.1)
    assert np.array_equal(c1, c2)
    f = rng.standard_normal((nz, ny, nx))
    for bb, gw in ((bbox, 3), ((-9, -9, -9, 9, 9, 9), 2)):
        assert abs(fa.l2norm(f, X, Y, Z, bb, gw) / orc.l2norm(f, X, Y, Z, bb, gw) - 1) < 1e-13
    f0 = rng.standard_normal((nz, ny, nx))
    g, o = f.copy(), f.copy()
    fa.sommerfeld_copy(f0, g, X, Y, Z, bbox)
    orc.sommerfeld_copy(f0, o, X, Y, Z, bbox)
    assert np.array_equal(g, o)


def _level_vs_oracle(orc, cfg, gauge, bc, nsteps, tol, variant=0):
    import amssbssn as A
    lv = A.Level.from_config(cfg, gauge=gauge, bc=bc)
    lv.set_variant(variant)
    p = cfg.patches[0]
    X, Y, Z = p.coords()
    U = cfg.initial_state(0)
    for _ in range(nsteps):
        orc.step(U, X, Y, Z, p.bbox, cfg.dt, cfg.eps, cfg.chitiny, gauge=gauge, bc=bc, gw=p.gw)
    assert lv.step(cfg.dt, nsteps) == 0
    G = lv.get_state(0)
    errs = [rel_err(g, u) for g, u in zip(G, U)]
    lv.close()
    assert max(errs) <= tol, f"per-field rel errs {errs}"
    return max(errs)


# variant 0: one thread per point, stencils straight from HBM; variant 1: LDS
# tiles of 8x8x4 points on the tile-aligned part of the interior, variant 0 on
# the remainder slabs (n = 18 leaves a 2-point remainder on every axis).
@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("bc", [0, 1, 2, 3])
def test_one_rk4_step_matches_oracle(orc, bc, variant):
    import amssbssn as A
    cfg = A.config1(n=18)
    worst = _level_vs_oracle(orc, cfg, 0, bc, 1, 1e-12, variant)
    print(f"bc {bc} variant {variant}: worst {worst:.2e}")


def test_one_rk4_step_single_puncture_small(orc):
    import amssbssn as A
    g = A.cube_patch(20, 8.0)
    cfg = A.Config("single puncture 20^3", [g], dt=0.25 * g.h[0], steps=1, punctures=[(1.0, 0, 0, 0)])
    _level_vs_oracle(orc, cfg, 0, 0, 1, 1e-12)


@pytest.mark.parametrize("variant", [0, 1])
def test_twenty_steps_match_oracle(orc, variant):
    import amssbssn as A
    cfg = A.config1(n=16)
    worst = _level_vs_oracle(orc, cfg, 0, 0, 20, 1e-10, variant)
    print(f"20 steps variant {variant}: worst {worst:.2e}")


@pytest.mark.parametrize("variant", [0, 1])
def test_gauge2_steps_match_oracle(orc, variant):
    import amssbssn as A
    cfg = A.config1(n=14)
    _level_vs_oracle(orc, cfg, 2, 0, 5, 1e-10, variant)


@pytest.mark.parametrize("gauge", [0, 2])
def test_sommerfeld_boundary_steps_match_oracle(orc, gauge):
    """The level-0 BAM Sommerfeld boundary over several steps, where the
    boundary layers feed back into the interior."""
    import amssbssn as A
    cfg = A.config1(n=16)
    worst = _level_vs_oracle(orc, cfg, gauge, A.BC_SOMMERFELD, 6, 1e-10, 1)
    print(f"Sommerfeld gauge {gauge}, 6 steps: worst {worst:.2e}")
    # the boundary is visible at that tolerance: without it the state differs
    p = cfg.patches[0]
    X, Y, Z = p.coords()
    U3, U2 = cfg.initial_state(0), cfg.initial_state(0)
    for _ in range(6):
        orc.step(U3, X, Y, Z, p.bbox, cfg.dt, cfg.eps, cfg.chitiny, gauge=gauge, bc=A.BC_SOMMERFELD, gw=p.gw)
        orc.step(U2, X, Y, Z, p.bbox, cfg.dt, cfg.eps, cfg.chitiny, gauge=gauge, bc=A.BC_NONE, gw=p.gw)
    assert max(rel_err(a, b) for a, b in zip(U3, U2)) > 1e-6


@pytest.mark.parametrize("sym", [1, 2])
@pytest.mark.parametrize("bc", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_level_with_symmetry_planes_matches_oracle(orc, sym, bc, mode):
    """A level whose patch starts at the symmetry planes (equatorial: z;
    octant: x, y, z), two steps, per-patch and batched: mirrored boundary
    stencils, one-layer boundaries skipping the plane faces."""
    import amssbssn as A
    n = (20, 18, 22)
    h = 0.3
    p = A.PatchGeom(n, (0.0, 0.0, 0.0, n[0] * h, n[1] * h, n[2] * h), 3)
    X, Y, Z = p.coords()
    st = random_state((n[2], n[1], n[0]), seed=100 + 10 * sym + bc)
    lv = A.Level([p], gauge=0, eps=0.02, bc=bc, symmetry=sym)
    lv.set_state(0, st)
    U = [x.copy() for x in st]
    for _ in range(2):
        orc.step(U, X, Y, Z, p.bbox, 0.075, 0.02, 1e-5, symmetry=sym, gauge=0, bc=bc, gw=3)
    assert lv.step(0.075, 2, mode=mode) == 0
    G = lv.get_state(0)
    lv.close()
    errs = [rel_err(g, u) for g, u in zip(G, U)]
    assert max(errs) <= 1e-12, f"per-field rel errs {errs}"
    # the planes matter at that tolerance
    V = [x.copy() for x in st]
    for _ in range(2):
        orc.step(V, X, Y, Z, p.bbox, 0.075, 0.02, 1e-5, symmetry=0, gauge=0, bc=bc, gw=3)
    assert max(rel_err(a, b) for a, b in zip(U, V)) > 1e-8


def test_sommerfeld_routbam_abi_matches_oracle(fa, orc):
    """sommerfeld_routbam_ on host arrays: faces on and off the bbox,
    symmetry planes (one mirrored layer with the parity), both speeds."""
    rng = np.random.default_rng(21)
    X, Y, Z, bbox = grid(9, 10, 11)
    P = [(np.arange(n) + 0.5) * h for n, h in ((9, 0.3), (10, 0.27), (11, 0.33))]
    pb = (0.0, 0.0, 0.0, 9 * 0.3, 10 * 0.27, 11 * 0.33)
    cases = [(X, Y, Z, bbox, 0), (P[0], P[1], P[2], pb, 1), (P[0], P[1], P[2], pb, 2),
             (X, Y, Z, (bbox[0] - 1.0, bbox[1], bbox[2], bbox[3], bbox[4] + 1.0, bbox[5]), 0)]
    for X_, Y_, Z_, bb, sym in cases:
        shape = (len(Z_), len(Y_), len(X_))
        f0 = rng.standard_normal(shape)
        r0 = rng.standard_normal(shape)
        for vel, par in ((1.0, (1, 1, 1)), (2 ** 0.5, (-1, 1, -1)), (1.0, (1, -1, -1))):
            g, o = r0.copy(), r0.copy()
            fa.sommerfeld_bam(g, f0, X_, Y_, Z_, bb, vel, par, sym)
            orc.sommerfeld_bam(o, f0, X_, Y_, Z_, bb, vel, par, sym)
            assert rel_err(g, o) <= 1e-14, (sym, vel, par)
            assert np.array_equal(g == r0, o == r0)  # the same points written


@pytest.mark.parametrize("n", [24, 44])
def test_lds_variant_matches_direct_kernel(n):
    """Both interior kernels evaluate the same expression tree per point (the
    compiler may contract different products into FMAs), so a 24^3 patch
    (three whole tiles on x and y, six on z) and a 44^3 one (5 x 5 x 11
    tiles, with remainder slabs for the direct kernel) agree to rounding,
    the seven constraint norms of both steps included."""
    import amssbssn as A
    cfg = A.config1(n=n)
    out, res = [], []
    for variant in (0, 1):
        lv = A.Level.from_config(cfg, gauge=0)
        lv.set_variant(variant)
        assert lv.variant() == variant
        assert lv.step(cfg.dt, 2, want_ham=A.RES_EVERY_STEP) == 0
        out.append(lv.get_state(0))
        res.append(lv.constraint_series())
        lv.close()
    for v in range(24):
        assert rel_err(out[1][v], out[0][v]) <= 1e-13, v
    assert np.all(np.abs(res[1] - res[0]) <= 1e-12 * np.max(np.abs(res[0]), axis=0))


def test_two_patches_equal_one_patch():
    """A cube split into two touching patches evolves identically to the
    whole cube: the ghost exchange hands each patch its neighbour's values."""
    import amssbssn as A
    n, gw = 16, 3
    whole = A.cube_patch(n, 8.0)
    h = whole.h[0]
    lo = -8.0 - gw * h
    left = A.PatchGeom((n // 2 + 2 * gw, n + 2 * gw, n + 2 * gw),
                       (lo, lo, lo, 0.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    right = A.PatchGeom((n // 2 + 2 * gw, n + 2 * gw, n + 2 * gw),
                        (0.0 - gw * h, lo, lo, 8.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    pun = [(0.5, 1.0, 0.5, 0.0), (0.5, -1.5, 0.0, 0.3)]
    c1 = A.Config("whole", [whole], dt=0.25 * h, steps=3, punctures=pun, a_noise=1e-6)
    c2 = A.Config("split", [left, right], dt=0.25 * h, steps=3, punctures=pun, a_noise=1e-6)
    L1 = A.Level.from_config(c1)
    L2 = A.Level(c2.patches, eps=c2.eps)
    U = c1.initial_state(0)
    # same initial data on the split patches (cut from the whole array)
    L2.set_state(0, [u[:, :, : n // 2 + 2 * gw].copy() for u in U])
    L2.set_state(1, [u[:, :, n // 2:].copy() for u in U])
    L1.step(c1.dt, 3)
    L2.step(c2.dt, 3)
    W = L1.get_state(0)
    A0, A1 = L2.get_state(0), L2.get_state(1)
    for v in range(24):
        w = W[v][gw:-gw, gw:-gw, gw:-gw]
        a = A0[v][gw:-gw, gw:-gw, gw:-gw]
        b = A1[v][gw:-gw, gw:-gw, gw:-gw]
        assert np.array_equal(w[:, :, : n // 2], a), v
        assert np.array_equal(w[:, :, n // 2:], b), v


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_split_patches_match_whole_cube_oracle(orc, mode):
    """Four patches tiling a cube (split along x and z, uneven), evolved with
    the on-device exchange, against the oracle's whole-cube evolution: the
    owned points of every patch to 1e-12 after 3 steps."""
    import amssbssn as A
    n, gw = 16, 3
    whole = A.cube_patch(n, 8.0)
    h = whole.h[0]
    cuts = ((0, 6), (6, n))  # owned index ranges along x and along z
    patches, ranges = [], []
    for kz in cuts:
        for ix in cuts:
            shape = (ix[1] - ix[0] + 2 * gw, n + 2 * gw, kz[1] - kz[0] + 2 * gw)
            lo = (-8.0 + (ix[0] - gw) * h, -8.0 - gw * h, -8.0 + (kz[0] - gw) * h)
            bb = lo + (lo[0] + shape[0] * h, lo[1] + shape[1] * h, lo[2] + shape[2] * h)
            patches.append(A.PatchGeom(shape, bb, gw))
            ranges.append((ix, kz))
    pun = [(0.5, 1.0, 0.5, 0.0), (0.5, -1.5, 0.0, 0.3)]
    c1 = A.Config("whole", [whole], dt=0.25 * h, steps=3, punctures=pun, a_noise=1e-6)
    X, Y, Z = whole.coords()
    U = c1.initial_state(0)
    lv = A.Level(patches, eps=c1.eps)
    try:
        for p, (ix, kz) in enumerate(ranges):
            lv.set_state(p, [u[kz[0]:kz[1] + 2 * gw, :, ix[0]:ix[1] + 2 * gw].copy() for u in U])
        assert lv.step(c1.dt, 3, mode) == 0
        got = [lv.get_state(p) for p in range(len(patches))]
    finally:
        lv.close()
    for _ in range(3):
        orc.step(U, X, Y, Z, whole.bbox, c1.dt, c1.eps, c1.chitiny, gauge=0, bc=0, gw=gw)
    for v in range(24):
        ref = U[v][gw:-gw, gw:-gw, gw:-gw]
        m = float(np.max(np.abs(ref)))
        for (ix, kz), G in zip(ranges, got):
            a = G[v][gw:-gw, gw:-gw, gw:-gw]
            b = ref[kz[0]:kz[1], :, ix[0]:ix[1]]
            e = float(np.max(np.abs(a - b)))
            assert e <= 1e-12 * (m if m > 0 else 1.0), (v, ix, kz, e)


# ---------------------------------------------------------------------------
# Against the reference Fortran's golden vectors (tests/golden, GAUGE 2: the
# reference build's configured gauge).  Per field: relative difference at the
# sampled points (scaled by the field's max |f|), max |f| and sum f^2.
def _gold(name):
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    if not os.path.exists(p):
        pytest.skip(f"fixture {name} not generated")
    return np.load(p)


def _vs_gold(G, g, tol):
    idx = g["idx"]
    worst = 0.0
    for v, u in enumerate(G):
        m = float(g["maxabs"][v])
        d = float(np.max(np.abs(u.ravel()[idx] - g["samples"][v])))
        e = d / m if m > 0 else d
        worst = max(worst, e)
        assert e <= tol, (v, e)
        if m > 0:
            assert abs(float(np.max(np.abs(u))) / m - 1) <= tol, v
            assert abs(float(np.sum(u * u)) / float(g["sumsq"][v]) - 1) <= 2 * tol, v
    return worst


def test_c0_one_step_vs_reference_vectors():
    import amssbssn as A
    g = _gold("ref_c0_step1_gauge2.npz")
    cfg = A.config0()
    lv = A.Level.from_config(cfg, gauge=2)
    assert lv.step(cfg.dt, 1, want_ham=True) == 0
    worst = _vs_gold(lv.get_state(0), g, 1e-12)
    assert abs(lv.ham_l2() / float(g["ham_l2"]) - 1) < 1e-12
    lv.close()
    print(f"single puncture 48^3, one step vs reference: worst {worst:.2e}")


@pytest.mark.parametrize("variant", [0, 1])
def test_c1_32cubed_100_steps_vs_reference_vectors(variant):
    import amssbssn as A
    g = _gold("ref_c1_32cubed_step100_gauge2.npz")
    cfg = A.config1(n=32)
    lv = A.Level.from_config(cfg, gauge=2)
    lv.set_variant(variant)
    assert lv.step(cfg.dt, 100, want_ham=True) == 0
    worst = _vs_gold(lv.get_state(0), g, 1e-9)
    assert abs(lv.ham_l2() / float(g["ham_l2"][-1]) - 1) < 1e-9
    lv.close()
    print(f"two punctures 32^3, 100 steps vs reference: worst {worst:.2e}")


@pytest.mark.parametrize("variant", [0, 1])
def test_c1_128cubed_100_steps_vs_reference_vectors(variant):
    import amssbssn as A
    g = _gold("ref_c1_128cubed_step100_gauge2.npz")
    cfg = A.config1(n=128)
    lv = A.Level.from_config(cfg, gauge=2)
    lv.set_variant(variant)
    assert lv.step(cfg.dt, 100) == 0
    worst = _vs_gold(lv.get_state(0), g, 1e-9)
    lv.close()
    print(f"two punctures 128^3, 100 steps vs reference: worst {worst:.2e}")


@pytest.mark.parametrize("variant", [0, 1])
def test_odd_row_pitch_matches_oracle(orc, variant):
    """A 17^3 interior (23 points per axis with ghosts): an odd row pitch, which
    the LDS path hands to the direct kernel (its 16-byte loads need an even one)."""
    import amssbssn as A
    cfg = A.config1(n=17)
    _level_vs_oracle(orc, cfg, 0, 0, 2, 1e-12, variant)


@pytest.mark.parametrize("dT", [0.05, 0.6])
def test_sommerfeld_interpolating_predictor_abi_matches_oracle(fa, orc, dT):
    """sommerfeld_rout_ with precor 0 (SommerType 1's predictor) against the
    oracle: the same face points written, values to 1e-13 relative (same
    Neville order; the GPU contracts multiply-adds)."""
    from test_oracle_reference import _bam_cases, PARS
    rng = np.random.default_rng(41)
    for X, Y, Z, bbox, sym in _bam_cases():
        shape = (len(Z), len(Y), len(X))
        f0 = rng.standard_normal(shape)
        chi0 = -0.3 + 0.05 * rng.standard_normal(shape)
        lap0 = -0.2 + 0.05 * rng.standard_normal(shape)
        start = rng.standard_normal(shape)
        for par in PARS[:3]:
            g, o = start.copy(), start.copy()
            fa.sommerfeld_interp(f0, g, chi0, lap0, X, Y, Z, bbox, dT, par, sym)
            assert orc.sommerfeld_interp(f0, o, chi0, lap0, X, Y, Z, bbox, dT, par, sym) == 0
            assert np.array_equal(g != start, o != start), (sym, par)
            assert np.max(np.abs(g - o)) <= 1e-13 * np.max(np.abs(o)), (sym, par)
