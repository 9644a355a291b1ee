"""Constraint residuals in every predictor and their per-step L2 norms
(bssn_class::Step with co = 0 in each predictor, bssn_class.C:3086-3092;
Constraint_Out, bssn_class.C:7475-7481) against the oracle.

The oracle's residuals of a step are those of compute_rhs_bssn (co = 0) on
the step's enforced base state; the norms follow l2normhelper
(fmisc.f90:1263-1309) with the patch bbox: under the reference's boundary
modes every point of the Block counts, with frozen ghost zones the points
[gw, n - gw).  Tolerance: 1e-12 relative per norm (the GPU evaluates the
residual algebra in a different order; the sums differ in order too)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_series(orc, cfg, gauge, bc, nsteps, split=None):
    """Norms of the 7 residuals per step, oracle side (whole cube)."""
    p = cfg.patches[0]
    X, Y, Z = p.coords()
    U = cfg.initial_state(0)
    gw = p.gw
    dv = p.h[0] * p.h[1] * p.h[2]
    rows = []
    for _ in range(nsteps):
        V = [u.copy() for u in U]
        orc.enforce_ga(V)
        res = orc.compute_rhs(V, X, Y, Z, cfg.eps, co=0, gauge=gauge)["res"]
        if bc == 0:
            s = np.s_[gw:-gw, gw:-gw, gw:-gw]
            rows.append([np.sqrt(np.sum(r[s] ** 2) * dv) for r in res])
        else:
            rows.append([np.sqrt(orc.l2norm(r, X, Y, Z, p.bbox, gw)) for r in res])
        orc.step(U, X, Y, Z, p.bbox, cfg.dt, cfg.eps, cfg.chitiny, gauge=gauge, bc=bc, gw=gw)
    return np.array(rows), U


def _rel(a, b):
    """|a - b| relative to |b|, or to the largest entry of b where b is zero
    (the Gamma-constraint norms of conformally flat data vanish)."""
    b = np.asarray(b)
    floor = 1e-2 * np.max(np.abs(b)) if b.size else 1.0
    return np.abs(a - b) / np.maximum(np.abs(b), max(floor, 1e-300))


@pytest.mark.parametrize("bc", [0, 1, 3])
@pytest.mark.parametrize("gauge", [0, 2])
def test_constraint_series_matches_oracle(orc, bc, gauge):
    import amssbssn as A
    cfg = A.config1(n=20)
    nsteps = 3
    want, U = _oracle_series(orc, cfg, gauge, bc, nsteps)
    lv = A.Level.from_config(cfg, gauge=gauge, bc=bc)
    try:
        assert lv.step(cfg.dt, nsteps, A.MODE_PLAIN, want_ham=A.RES_EVERY_STEP) == 0
        got = lv.constraint_series()
        assert got.shape == (nsteps, 7)
        assert np.all(want[:, 0] > 0) and np.all(want[1:] > 0)
        assert np.max(_rel(got, want)) <= 1e-12, (got, want)
        assert np.array_equal(lv.constraint_series(0), got)  # one patch: patch 0 is the level
        # the last step's norms, read again, and the Hamiltonian one
        last = lv.constraint_norms()
        assert np.max(_rel(last, want[-1])) <= 1e-12
        assert abs(lv.ham_l2() - last[0]) == 0.0
        # the evolution itself is unchanged by the residuals
        G = lv.get_state(0)
        for g, u in zip(G, U):
            m = float(np.max(np.abs(u)))
            assert float(np.max(np.abs(g - u))) <= 1e-12 * (m if m > 0 else 1.0)
    finally:
        lv.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_constraint_series_batched_and_graph_equal_per_patch(mode):
    """Residuals in every predictor through the batched tile kernels and the
    captured graphs give the per-patch path's norms (uneven 3-patch level)."""
    import amssbssn as A
    patches, _ = A.uneven_level(3, 16, 32, seed=5, half_width=6.0)
    h = patches[0].h[0]
    cfg = A.Config("3 uneven patches", patches, dt=0.25 * h, steps=4,
                   punctures=[(0.5, 1.5, 0.0, 0.0), (0.5, -1.5, 0.0, 0.0)], a_noise=1e-6, noise_seed=0)
    out = []
    for m in (0, mode):
        lv = A.Level.from_config(cfg, gauge=0)
        try:
            assert lv.step(cfg.dt, 4, m, want_ham=A.RES_EVERY_STEP) == 0
            out.append((lv.constraint_series(), [lv.constraint_series(p) for p in range(len(cfg.patches))]))
        finally:
            lv.close()
    (a, ap), (b, bp) = out
    assert np.max(_rel(b, a)) <= 1e-12
    A_, B_ = np.stack(ap), np.stack(bp)
    assert np.all(np.abs(B_ - A_) <= 1e-12 * np.max(np.abs(A_), axis=0) + 1e-300)
    # the level's norm is the root of the patches' summed squares
    assert np.allclose(a ** 2, sum(x ** 2 for x in ap), rtol=1e-13, atol=0)


def test_split_level_norms_match_whole_cube_oracle(orc):
    """Two patches tiling a cube (frozen ghosts): the level's norm over both
    equals the oracle's whole-cube norm, and each patch's equals the oracle's
    over that patch's owned half."""
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
    c1 = A.Config("whole", [whole], dt=0.25 * h, steps=2, punctures=pun, a_noise=1e-6)
    X, Y, Z = whole.coords()
    U = c1.initial_state(0)
    L2 = A.Level([left, right], eps=c1.eps)
    try:
        L2.set_state(0, [u[:, :, : n // 2 + 2 * gw].copy() for u in U])
        L2.set_state(1, [u[:, :, n // 2:].copy() for u in U])
        assert L2.step(c1.dt, 2, A.MODE_PLAIN, want_ham=A.RES_EVERY_STEP) == 0
        got = [L2.constraint_series(p) for p in (0, 1)]
        tot = L2.constraint_series()
    finally:
        L2.close()
    dv = h ** 3
    for s in range(2):
        V = [u.copy() for u in U]
        orc.enforce_ga(V)
        res = orc.compute_rhs(V, X, Y, Z, c1.eps, co=0, gauge=0)["res"]
        core = [r[gw:-gw, gw:-gw, gw:-gw] for r in res]
        want_l = np.array([np.sqrt(np.sum(c[:, :, : n // 2] ** 2) * dv) for c in core])
        want_r = np.array([np.sqrt(np.sum(c[:, :, n // 2:] ** 2) * dv) for c in core])
        want = np.array([np.sqrt(np.sum(c ** 2) * dv) for c in core])
        assert np.max(_rel(got[0][s], want_l)) <= 1e-12
        assert np.max(_rel(got[1][s], want_r)) <= 1e-12
        assert np.max(_rel(tot[s], want)) <= 1e-12
        orc.step(U, X, Y, Z, whole.bbox, c1.dt, c1.eps, c1.chitiny, gauge=0, bc=0, gw=gw)
