"""A two-level hierarchy (Berger-Oliger sub-cycling, RestrictProlong) on the
GPU against the oracle's statement of it (oracle.amr_step_driver over the C
restatement): every point of both levels after two coarse steps (the second
one interpolating the coarse level over three time levels), to 1e-12
relative per field; the batched modes equal the per-patch one; and the
closed form of a uniform state whose shift grows linearly in time."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu

NF = 24


def _setup(lo=(6, 6, 6), hi=(18, 18, 18), n=24, half=6.0):
    import amssbssn as A
    coarse = A.cube_patch(n, half)
    fine = A.nested_box(coarse, lo, hi)
    pun = [(0.5, 1.5, 0.3, 0.0), (0.5, -1.5, -0.2, 0.1)]
    cc = A.Config("coarse", [coarse], dt=0.25 * coarse.h[0], steps=2, punctures=pun, a_noise=1e-6)
    cf = A.Config("fine", [fine], dt=0.25 * fine.h[0], steps=2, punctures=pun, a_noise=1e-6)
    return coarse, fine, cc, cf


def _oracle(orc, coarse, fine, Uc, Uf, dT, nsteps, bc0, gauge=0):
    import oracle
    C = oracle.AmrLevel(Uc, coarse.bbox)
    F = oracle.AmrLevel(Uf, fine.bbox, lli=(6, 6, 6), uui=(6, 6, 6))
    t = 0.0
    for _ in range(nsteps):
        t = oracle.amr_step_driver(orc, [C, F], dT, 0.02, gauge=gauge, bc0=bc0, t=(t, 0.0))
    return C.U, F.U, t


def _regions(coarse, fine):
    gw = 3
    rc = np.zeros(coarse.shape[::-1], bool)
    rc[gw + 6:gw + 18, gw + 6:gw + 18, gw + 6:gw + 18] = True   # restricted coarse cells
    bf = np.ones(fine.shape[::-1], bool)
    bf[6:-6, 6:-6, 6:-6] = False                              # fine buffer zone
    return {"coarse restricted": (0, rc), "coarse other": (0, ~rc), "fine buffer": (1, bf),
            "fine owned": (1, ~bf)}


@pytest.mark.parametrize("nsteps", [1, 2])
@pytest.mark.parametrize("bc0", [0, 3])
def test_two_levels_match_oracle(orc, bc0, nsteps):
    import amssbssn as A
    coarse, fine, cc, cf = _setup()
    Uc, Uf = cc.initial_state(0), cf.initial_state(0)
    dT = cc.dt
    H = A.Hierarchy([[coarse], [fine]], gauge=0, eps=0.02, bc0=bc0)
    try:
        H.level(0).set_state(0, Uc)
        H.level(1).set_state(0, Uf)
        assert H.step(dT, nsteps) == 0
        assert H.time == pytest.approx(nsteps * dT)
        Gc, Gf = H.level(0).get_state(0), H.level(1).get_state(0)
    finally:
        H.close()
    Oc, Of, t = _oracle(orc, coarse, fine, Uc, Uf, dT, nsteps, bc0)
    assert t == pytest.approx(nsteps * dT)
    for rname, (lev, m) in _regions(coarse, fine).items():
        G, O = (Gc, Oc) if lev == 0 else (Gf, Of)
        errs = [float(np.max(np.abs(G[v][m] - O[v][m])) / max(float(np.max(np.abs(O[v]))), 1e-300))
                for v in range(NF)]
        print(f"bc0 {bc0}, {nsteps} step(s), {rname}: worst field {int(np.argmax(errs))} {max(errs):.2e}")
    worst = 0.0
    for name, G, O in (("coarse", Gc, Oc), ("fine", Gf, Of)):
        for v in range(NF):
            e = rel_err(G[v], O[v])
            worst = max(worst, e)
            assert e <= 1e-12, (name, v, e)
    # the fine level changed the coarse one (restriction) and moved off its start
    assert rel_err(Of[0], Uf[0]) > 1e-6
    print(f"two levels, bc0 {bc0}, {nsteps} coarse steps: worst {worst:.2e}")


@pytest.mark.parametrize("mode", [1, 2])
def test_two_levels_batched_modes_equal_per_patch(mode):
    import amssbssn as A
    coarse, fine, cc, cf = _setup()
    out = []
    for m in (0, mode):
        H = A.Hierarchy([[coarse], [fine]], gauge=0, eps=0.02)
        try:
            H.level(0).set_state(0, cc.initial_state(0))
            H.level(1).set_state(0, cf.initial_state(0))
            assert H.step(cc.dt, 2, m) == 0
            out.append((H.level(0).get_state(0), H.level(1).get_state(0)))
        finally:
            H.close()
    for a, b in zip(out[0], out[1]):
        for v in range(NF):
            assert rel_err(b[v], a[v]) <= 1e-12, v


def test_two_levels_uniform_linear_shift_closed_form():
    """GAUGE 2, no outer boundary: points the edges cannot reach hold
    beta0 + 3/4 B t after two coarse steps (RK4 is exact for a shift linear
    in time, and so are the restriction, prolongation and time averages)."""
    import amssbssn as A
    n = 48
    coarse = A.cube_patch(n, 12.0)
    lo, hi = (12, 12, 12), (36, 36, 36)
    fine = A.nested_box(coarse, lo, hi)
    vals = np.zeros(NF)
    vals[0], vals[17] = -0.3, -0.2
    vals[18:21] = (0.1, -0.05, 0.02)
    vals[21:24] = (0.3, -0.2, 0.5)
    H = A.Hierarchy([[coarse], [fine]], gauge=2, eps=0.02, bc0=A.BC_NONE)
    dT = 0.125
    try:
        H.level(0).set_state(0, [np.full(coarse.shape[::-1], v) for v in vals])
        H.level(1).set_state(0, [np.full(fine.shape[::-1], v) for v in vals])
        assert H.step(dT, 1) == 0
        Gc, Gf = H.level(0).get_state(0), H.level(1).get_state(0)
    finally:
        H.close()
    want = vals.copy()
    want[18:21] += 0.75 * vals[21:24] * dT
    gw = 3
    far = np.zeros(Gc[0].shape, bool)
    far[12:-12, 12:-12, 12:-12] = True
    far[gw + lo[2]:gw + hi[2], gw + lo[1]:gw + hi[1], gw + lo[0]:gw + hi[0]] = False
    for v in range(NF):
        assert np.max(np.abs(Gc[v][far] - want[v])) <= 1e-14, v
        assert np.max(np.abs(Gf[v][24:-24, 24:-24, 24:-24] - want[v])) <= 1e-14, v


def test_hierarchy_refuses_a_fine_patch_too_close_to_the_coarse_edge():
    import amssbssn as A
    coarse = A.cube_patch(16, 4.0)
    fine = A.nested_box(coarse, (0, 4, 4), (8, 12, 12))  # its buffer reaches the coarse ghosts
    with pytest.raises(A.AmssError, match="buffer zone"):
        A.Hierarchy([[coarse], [fine]])


def test_two_level_checkpoint_restart_continues_bitwise(tmp_path):
    """checkpoint::writecheck_cgh's layout over both levels: a hierarchy read
    back from it continues exactly as the one that wrote it (two coarse steps
    in, so the next steps interpolate the coarse level over three time
    levels from the restored OldStateList)."""
    import struct
    import amssbssn as A
    coarse, fine, cc, cf = _setup()
    dT = cc.dt
    porg = [[(1.5, 0.3, 0.0), (-1.5, -0.2, 0.1)], [(1.5, 0.3, 0.0), (-1.5, -0.2, 0.1)]]
    path = str(tmp_path / "amr_cgh.CHK")

    def make():
        return A.Hierarchy([[coarse], [fine]], gauge=0, eps=0.02)

    H = make()
    try:
        H.level(0).set_state(0, cc.initial_state(0))
        H.level(1).set_state(0, cf.initial_state(0))
        assert H.step(dT, 2) == 0
        H.checkpoint_write(path, movls=1, porg=porg)
        assert H.step(dT, 2) == 0
        want = [[H.level(l).get_field(0, lst, v) for lst in (0, 1) for v in range(NF)] for l in (0, 1)]
        t_want = H.time
    finally:
        H.close()
    with open(path, "rb") as f:
        raw = f.read()
    t, levels, movls, nbh = struct.unpack_from("<diii", raw, 0)
    grids = struct.unpack_from("<2i", raw, 20)
    lt = struct.unpack_from("<2d", raw, 28)
    assert (levels, movls, nbh, grids) == (2, 1, 2, (1, 1))
    assert t == pytest.approx(2 * dT) and lt == pytest.approx((2 * dT, 2 * dT))
    table = 2 * (6 * 8 + 3 * 4 + 3 * 8 + 2 * 3 * 8)
    data = 2 * NF * 8 * (int(np.prod(coarse.shape)) + int(np.prod(fine.shape)))
    assert len(raw) == 44 + table + data
    H2 = make()
    try:
        t2, mv, pg = H2.checkpoint_read(path)
        assert t2 == pytest.approx(2 * dT) and mv == 1
        assert np.allclose(pg, porg, rtol=0, atol=0)
        assert H2.step(dT, 2) == 0
        got = [[H2.level(l).get_field(0, lst, v) for lst in (0, 1) for v in range(NF)] for l in (0, 1)]
        assert H2.time == t_want
    finally:
        H2.close()
    for l in (0, 1):
        for i, (a, b) in enumerate(zip(got[l], want[l])):
            assert np.array_equal(a, b), (l, i)


def test_checkpoint_of_another_hierarchy_is_refused(tmp_path):
    import amssbssn as A
    coarse, fine, cc, cf = _setup()
    path = str(tmp_path / "amr_cgh.CHK")
    H = A.Hierarchy([[coarse], [fine]], gauge=0, eps=0.02)
    try:
        H.checkpoint_write(path)
    finally:
        H.close()
    other = A.nested_box(coarse, (4, 6, 6), (16, 18, 18))
    H2 = A.Hierarchy([[coarse], [other]], gauge=0, eps=0.02)
    try:
        with pytest.raises(A.AmssError, match="bbox differs"):
            H2.checkpoint_read(path)
    finally:
        H2.close()
