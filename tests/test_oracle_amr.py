"""The oracle's statement of Berger-Oliger sub-cycling (RecursiveStep /
RestrictProlong, bssn_class.C:2323-2403, 5724-5904) checked by what the
mathematics fixes:

* the time interpolation of prepare_inter_time_level (average, average2 of
  fmisc.f90:1532-1592) is exact at t + dt/2 for fields linear (two levels)
  and quadratic (three levels) in time;
* restriction and prolongation over the reference's regions leave a field
  that is a polynomial of degree <= 5 per axis on both levels unchanged, and
  write exactly the coarse cells under the fine patch less its buffer zone
  and exactly the fine buffer zone;
* flat space stays exactly flat over a two-level step;
* a uniform state whose shift grows linearly in time (GAUGE 2: beta' = 3/4 B,
  B' = 0, every other right-hand side zero) is integrated exactly by RK4, so
  after one coarse step (two fine steps of half the size) every point that
  the edges cannot reach (the fine level's one-layer boundary, the coarse
  level's mirrored edge stencils) holds beta0 + 3/4 B dT.
"""
import numpy as np
import pytest

NF = 24


def _two_levels(n_coarse, half, lo_cell, hi_cell):
    import amssbssn as A
    coarse = A.cube_patch(n_coarse, half)
    fine = A.nested_box(coarse, lo_cell, hi_cell)
    return coarse, fine


def _levels(oracle, coarse, fine, Uc, Uf):
    C = oracle.AmrLevel(Uc, coarse.bbox)
    F = oracle.AmrLevel(Uf, fine.bbox, lli=(6, 6, 6), uui=(6, 6, 6))
    return C, F


def _mesh(p):
    X, Y, Z = p.coords()
    z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
    return x, y, z


def test_time_interpolation_is_exact_on_polynomials_in_time():
    rng = np.random.default_rng(3)
    a, b, c = rng.standard_normal(3)
    dt, t = 0.3, 1.7
    f = lambda s: a + b * s + c * s * s  # noqa: E731
    # average2: 3/8 f(t + dt) + 3/4 f(t) - 1/8 f(t - dt) = f(t + dt/2) for quadratics
    assert abs(3.0 / 8.0 * f(t + dt) + 3.0 / 4.0 * f(t) + -1.0 / 8.0 * f(t - dt) - f(t + dt / 2)) < 1e-14
    g = lambda s: a + b * s  # noqa: E731
    assert abs(0.5 * (g(t + dt) + g(t)) - g(t + dt / 2)) < 1e-14
    # and the oracle's statement uses exactly these on whole lists
    import oracle
    coarse, fine = _two_levels(16, 4.0, (4, 4, 4), (12, 12, 12))
    shape_c = coarse.shape[::-1]
    C, F = _levels(oracle, coarse, fine, [np.zeros(shape_c)] * NF, [np.zeros(fine.shape[::-1])] * NF)
    C.U, C.old, C.cor = ([np.full(shape_c, f(s))] * NF for s in (t + dt, t, t - dt))
    # the fine level at the intermediate time, the coarse one interpolated to it
    F.U = [np.full(fine.shape[::-1], f(t + dt / 2)) for _ in range(NF)]
    oracle.amr_restrict_prolong(oracle.Orc(), C, F, same_time=False, three=True)
    assert np.max(np.abs(F.U[0] - f(t + dt / 2))) < 1e-13
    # a wrong time level (the coarse state at t + dt) would show in the buffer
    assert abs(f(t + dt) - f(t + dt / 2)) > 1e-3


def _quintic(x, y, z):
    return 1.0 + x ** 5 - 0.3 * x ** 2 * y ** 3 + 0.2 * y ** 4 * z - z ** 5 + x * y * z


@pytest.mark.parametrize("lo,hi", [((4, 4, 4), (12, 12, 12)), ((3, 5, 4), (11, 13, 10))])
def test_restrict_prolong_keep_quintics_and_write_their_regions(orc, lo, hi):
    import oracle
    coarse, fine = _two_levels(16, 4.0, lo, hi)
    Uc = [_quintic(*_mesh(coarse)) for _ in range(NF)]
    Uf = [_quintic(*_mesh(fine)) for _ in range(NF)]
    C, F = _levels(oracle, coarse, fine, Uc, Uf)
    oracle.amr_restrict_prolong(orc, C, F, same_time=True, three=False)
    for a, b in zip(C.U, Uc):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    for a, b in zip(F.U, Uf):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    # which points are written: an infinite coarse level takes finite values
    # where it is restricted, and hands non-finite ones to the fine buffer
    C, F = _levels(oracle, coarse, fine, [np.full(u.shape, np.inf) for u in Uc], [u.copy() for u in Uf])
    oracle.amr_restrict_prolong(orc, C, F, same_time=True, three=False)
    wc = np.isfinite(C.U[0])
    # the coarse cells under the fine patch less its buffer zone (3 coarse cells)
    gw = 3
    exp = np.zeros(wc.shape, bool)
    exp[gw + lo[2]:gw + hi[2], gw + lo[1]:gw + hi[1], gw + lo[0]:gw + hi[0]] = True
    assert np.array_equal(wc, exp)
    # the fine buffer zone is overwritten (with the inf-valued coarse data)
    wf = ~np.isfinite(F.U[0])
    expf = np.ones(wf.shape, bool)
    expf[6:-6, 6:-6, 6:-6] = False
    assert np.array_equal(wf, expf)


def test_flat_space_stays_flat_over_a_two_level_step(orc):
    import oracle
    coarse, fine = _two_levels(16, 4.0, (4, 4, 4), (12, 12, 12))
    C, F = _levels(oracle, coarse, fine, [np.zeros(coarse.shape[::-1])] * NF, [np.zeros(fine.shape[::-1])] * NF)
    t = oracle.amr_step_driver(orc, [C, F], 0.125, 0.02, gauge=0, bc0=0)
    assert t == pytest.approx(0.125)
    for u in C.U + F.U:
        assert not np.any(u)


def test_uniform_linear_shift_is_exact_after_a_coarse_step(orc):
    """GAUGE 2, no outer boundary (bc 2): the coarse level away from the
    fine patch and the fine level away from its one-layer boundary's reach
    (12 points per fine step, restricted back and prolonged again) hold the
    closed form beta0 + 3/4 B dT."""
    import oracle
    n, half = 48, 12.0
    lo, hi = (12, 12, 12), (36, 36, 36)
    coarse, fine = _two_levels(n, half, lo, hi)
    vals = np.zeros(NF)
    vals[0], vals[17] = -0.3, -0.2
    vals[18:21] = (0.1, -0.05, 0.02)
    vals[21:24] = (0.3, -0.2, 0.5)
    C, F = _levels(oracle, coarse, fine, [np.full(coarse.shape[::-1], v) for v in vals],
                   [np.full(fine.shape[::-1], v) for v in vals])
    dT = 0.125
    oracle.amr_step_driver(orc, [C, F], dT, 0.02, gauge=2, bc0=2)
    want = vals.copy()
    want[18:21] += 0.75 * vals[21:24] * dT
    gw = 3
    # coarse points outside the restricted cells and 12 points or more inside
    # the array (the edge stencils read the parity-mirrored layer below the
    # first point, so odd fields are not uniform there)
    far = np.zeros(C.U[0].shape, bool)
    far[12:-12, 12:-12, 12:-12] = True
    far[gw + lo[2]:gw + hi[2], gw + lo[1]:gw + hi[1], gw + lo[0]:gw + hi[0]] = False
    assert far.sum() > 1000
    for v in range(NF):
        assert np.max(np.abs(C.U[v][far] - want[v])) <= 1e-14, v
        deep = F.U[v][24:-24, 24:-24, 24:-24]
        assert np.max(np.abs(deep - want[v])) <= 1e-14, v
    # the excluded points do move (the selections are not vacuous)
    assert np.max(np.abs(C.U[18][:, :, :3] - want[18])) > 1e-6
    assert np.max(np.abs(F.U[18][30, 30, 6:12] - want[18])) > 1e-12
