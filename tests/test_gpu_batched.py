"""Batched level step (one tile-list launch per sub-step over all patches,
box-table ghost fill and exchange) and its hipGraph replay, against the
per-patch step -- which the single-patch tests pin to the oracle and the
split-patch test pins to the whole-cube evolution."""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _uneven(npatch=5, lo=16, hi=32, seed=3, half_width=6.0, steps=4):
    import amssbssn as A
    patches, _ = A.uneven_level(npatch, lo, hi, seed=seed, half_width=half_width)
    h = patches[0].h[0]
    return A.Config(f"{npatch} uneven patches", patches, dt=0.25 * h, steps=steps,
                    punctures=[(0.5, 1.5, 0.0, 0.0), (0.5, -1.5, 0.0, 0.0)], a_noise=1e-6, noise_seed=0)


def _run(cfg, mode, nsteps, bc=0, tune=None, want_ham=False):
    import amssbssn as A
    lv = A.Level.from_config(cfg, bc=bc)
    if tune is not None:
        lv.set_tune(*tune)
    assert lv.step(cfg.dt, nsteps, mode, want_ham=want_ham) == 0
    out = [lv.get_state(i) for i in range(len(cfg.patches))]
    ham = lv.ham_l2() if want_ham else None
    lv.close()
    return out, ham


def _compare(a, b, tol):
    worst = 0.0
    for pa, pb in zip(a, b):
        for v in range(24):
            e = rel_err(pa[v], pb[v])
            worst = max(worst, e)
            assert e <= tol, (v, e)
    return worst


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_matches_per_patch_uneven_level(mode):
    import amssbssn as A
    cfg = _uneven()
    ref, h0 = _run(cfg, A.MODE_PLAIN, 5, want_ham=True)
    got, h1 = _run(cfg, mode, 5, want_ham=True)
    worst = _compare(got, ref, 1e-12)
    assert abs(h1 / h0 - 1) < 1e-12
    print(f"mode {mode}: worst {worst:.2e}, ham {h1:.6e}")


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("grid,remap", [(0, 0), (0, 1), (24, 1), (24, 0)])
def test_every_tuning_candidate_matches_per_patch(shape, grid, remap):
    import amssbssn as A
    cfg = _uneven(npatch=4, seed=5)
    ref, _ = _run(cfg, A.MODE_PLAIN, 2)
    got, _ = _run(cfg, A.MODE_BATCHED, 2, tune=(shape, grid, remap))
    _compare(got, ref, 1e-12)


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_split_with_remainders_equals_whole(mode):
    """A 20^3 cube cut into two 10 x 20 x 20 patches: every tile shape leaves
    remainders, which the direct kernel covers."""
    import amssbssn as A
    n, gw = 20, 3
    whole = A.cube_patch(n, 8.0)
    h = whole.h[0]
    lo = -8.0 - gw * h
    left = A.PatchGeom((n // 2 + 2 * gw, n + 2 * gw, n + 2 * gw),
                       (lo, lo, lo, 0.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    right = A.PatchGeom((n // 2 + 2 * gw, n + 2 * gw, n + 2 * gw),
                        (0.0 - gw * h, lo, lo, 8.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    pun = [(0.5, 1.0, 0.5, 0.0), (0.5, -1.5, 0.0, 0.3)]
    c1 = A.Config("whole", [whole], dt=0.25 * h, steps=3, punctures=pun, a_noise=1e-6)
    L1 = A.Level.from_config(c1)
    L2 = A.Level([left, right], eps=c1.eps)
    U = c1.initial_state(0)
    L2.set_state(0, [u[:, :, : n // 2 + 2 * gw].copy() for u in U])
    L2.set_state(1, [u[:, :, n // 2:].copy() for u in U])
    assert L1.step(c1.dt, 4) == 0
    assert L2.step(c1.dt, 4, mode) == 0
    W = L1.get_state(0)
    A0, A1 = L2.get_state(0), L2.get_state(1)
    L1.close()
    L2.close()
    for v in range(24):
        w = W[v][gw:-gw, gw:-gw, gw:-gw]
        m = max(float(np.max(np.abs(w))), 1e-300)
        a = A0[v][gw:-gw, gw:-gw, gw:-gw]
        b = A1[v][gw:-gw, gw:-gw, gw:-gw]
        assert np.max(np.abs(w[:, :, : n // 2] - a)) <= 1e-12 * m, v
        assert np.max(np.abs(w[:, :, n // 2:] - b)) <= 1e-12 * m, v


@pytest.mark.parametrize("bc", [1, 2, 3])
def test_batched_other_boundary_modes_match_per_patch(bc):
    import amssbssn as A
    cfg = _uneven(npatch=3, seed=7)
    ref, _ = _run(cfg, A.MODE_PLAIN, 2, bc=bc)
    for mode in (1, 2):
        got, _ = _run(cfg, mode, 2, bc=bc)
        _compare(got, ref, 1e-12)


def test_graph_replays_match_eager_step_by_step():
    """Seven single-step calls in graph mode (the 3 list-slot phases, each
    graph replayed again after eager host work) against eager batched steps:
    the same kernels with the same arguments, so bit for bit."""
    import amssbssn as A
    cfg = _uneven(npatch=4, seed=11)
    la = A.Level.from_config(cfg)
    lb = A.Level.from_config(cfg)
    la.set_tune(0, 0, 1)
    lb.set_tune(0, 0, 1)
    for s in range(7):
        assert la.step(cfg.dt, 1, A.MODE_BATCHED) == 0
        assert lb.step(cfg.dt, 1, A.MODE_GRAPH) == 0
        for p in range(len(cfg.patches)):
            a, b = la.get_state(p), lb.get_state(p)
            for v in range(24):
                assert np.array_equal(a[v], b[v]), (s, p, v)
    la.close()
    lb.close()


def test_autotune_picks_a_valid_candidate():
    import amssbssn as A
    cfg = _uneven(npatch=4, seed=13)
    lv = A.Level.from_config(cfg)
    choice, table = lv.autotune()
    assert len(table) == 24 and all(t[3] > 0 for t in table)
    assert choice["shape"] in (0, 1, 2, 3) and choice["grid"] >= 0
    best = min(t[3] for t in table)
    assert any(t[3] == best and t[0] == choice["shape"] and t[2] == int(choice["xcd_order"]) for t in table)
    assert lv.get_tune() == choice
    assert lv.step(cfg.dt, 2, A.MODE_GRAPH) == 0
    lv.close()
    print(choice)


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_nan_input_is_reported(mode):
    import amssbssn as A
    cfg = _uneven(npatch=3, seed=17)
    lv = A.Level.from_config(cfg)
    st = cfg.initial_state(1)
    st[5][10, 10, 10] = np.nan
    lv.set_state(1, st)
    assert lv.step(cfg.dt, 1, mode) == 1
    lv.close()


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("x", [1, 12])
def test_nan_in_a_ghost_point_is_reported(mode, x):
    """A NaN in the state at a ghost point: x = 1 is a boundary ghost of the
    left patch (filled), x = 12 one the exchange overwrites with the right
    patch's values.  The reference tests every point of an RHS input."""
    import amssbssn as A
    n, gw = 16, 3
    h = 16.0 / n
    lo = -8.0 - gw * h
    left = A.PatchGeom((8 + 2 * gw, n + 2 * gw, n + 2 * gw), (lo, lo, lo, gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    right = A.PatchGeom((8 + 2 * gw, n + 2 * gw, n + 2 * gw), (-gw * h, lo, lo, 8.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    cfg = A.Config("split", [left, right], dt=0.25 * h, steps=1, punctures=[(0.5, 1.0, 0.5, 0.0)], a_noise=1e-6)
    lv = A.Level.from_config(cfg)
    st = cfg.initial_state(0)
    st[5][10, 10, x] = np.nan
    lv.set_state(0, st)
    assert lv.step(cfg.dt, 1, mode) == 1
    lv.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_odd_row_pitch_matches_per_patch(mode):
    """Patches whose row pitch is odd (15 and 23 points): all remainder, direct kernel."""
    import amssbssn as A
    n, gw = 17, 3
    h = 16.0 / n
    lo = -8.0 - gw * h
    left = A.PatchGeom((9 + 2 * gw, n + 2 * gw, n + 2 * gw), (lo, lo, lo, -8.0 + (9 + gw) * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    right = A.PatchGeom((8 + 2 * gw, n + 2 * gw, n + 2 * gw),
                        (-8.0 + (9 - gw) * h, lo, lo, 8.0 + gw * h, 8.0 + gw * h, 8.0 + gw * h), gw)
    cfg = A.Config("odd split", [left, right], dt=0.25 * h, steps=2,
                   punctures=[(0.5, 1.0, 0.5, 0.0), (0.5, -1.5, 0.0, 0.3)], a_noise=1e-6)
    ref, _ = _run(cfg, A.MODE_PLAIN, 2)
    got, _ = _run(cfg, mode, 2)
    _compare(got, ref, 1e-12)


@pytest.mark.parametrize("mode", [1, 2])
def test_batched_wider_frozen_zone_matches_per_patch(mode):
    """Frozen ghost zones 4 deep: the fill overlaps the first layer of tiles
    and must land after them in every sub-step, as in the per-patch step."""
    import amssbssn as A
    cfg = _uneven(npatch=4, seed=5)
    out = []
    for m in (0, mode):
        lv = A.Level(cfg.patches, gauge=0, eps=cfg.eps, chitiny=cfg.chitiny, gw=4)
        try:
            for i in range(len(cfg.patches)):
                lv.set_state(i, cfg.initial_state(i))
            assert lv.step(cfg.dt, 3, m) == 0
            out.append([lv.get_state(i) for i in range(len(cfg.patches))])
        finally:
            lv.close()
    _compare(out[1], out[0], 1e-12)
