"""CPU-side checks of the product: the C-ABI library loads and exports every
function include/amss_bssn.h declares (no compute call without a GPU), the
Python binding's signature table matches the header, and the host-side
geometry / initial-data logic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "amss_bssn.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(\w+)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ("compute_rhs_bssn_", "rungekutta4_rout_", "enforce_ga_", "lowerboundset_", "l2normhelper_",
                 "sommerfeld_rout_", "amss_rhs_bssn_dev", "amss_level_create", "amss_level_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import amssbssn
    lib = C.CDLL(amssbssn.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_the_header():
    from amssbssn._lib import SIGNATURES
    assert set(declared_functions()) == set(SIGNATURES)


def test_version_and_error_text_without_gpu():
    import amssbssn
    L = amssbssn.lib()
    assert L.amss_version() >= 100
    assert L.amss_set_gauge(5) == 4
    assert "gauge" in amssbssn.last_error()
    assert L.amss_set_gauge(0) == 0


def test_rhs_symmetry_arguments_rejected_before_any_gpu_work():
    """Symmetry outside 0..2, and a symmetry axis of fewer than 3 points,
    are argument errors (code 2), raised before the library touches a GPU."""
    import numpy as np
    import amssbssn
    from amssbssn import fortran_abi as fa
    X = (np.arange(6) + 0.5) * 0.3
    st = [np.zeros((6, 6, 6)) for _ in range(24)]
    with pytest.raises(RuntimeError, match="Symmetry must be 0, 1 or 2"):
        fa.compute_rhs(st, X, X, X, 0.02, symmetry=3)
    Z = (np.arange(2) + 0.5) * 0.3
    st2 = [np.zeros((2, 6, 6)) for _ in range(24)]
    with pytest.raises(RuntimeError, match="symmetry axis needs 3 points"):
        fa.compute_rhs(st2, X, X, Z, 0.02, symmetry=1)
    assert "3 points" in amssbssn.last_error()
    # without a plane an axis needs 5 points: the upwind stencil reaches 2 past a shorter one
    Z4 = -1.0 + np.arange(4) * 0.3
    st4 = [np.zeros((4, 6, 6)) for _ in range(24)]
    with pytest.raises(RuntimeError, match="needs 5 points"):
        fa.compute_rhs(st4, X, X, Z4, 0.02)


def test_level_symmetry_needs_a_reference_boundary_mode():
    """Frozen ghost zones have no meaning on a symmetry plane; the level
    refuses the pair before it allocates anything."""
    import amssbssn as A
    p = A.PatchGeom((10, 10, 10), (0.0, 0.0, 0.0, 3.0, 3.0, 3.0), 3)
    with pytest.raises(A.AmssError, match="symmetry plane"):
        A.Level([p], bc=A.BC_FROZEN_GHOSTS, symmetry=1)
    with pytest.raises(A.AmssError, match="Symmetry must be"):
        A.Level([p], bc=A.BC_NONE, symmetry=3)


def test_transfer_arguments_checked_before_any_gpu_work():
    """prolong3_ / restrict3_: grids that are not 2:1 return 2, a region the
    reference rejects returns 1, both before the library touches a GPU."""
    import numpy as np
    import amssbssn
    from amssbssn import fortran_abi as fa
    func, funf = np.zeros((12, 12, 12)), np.ones((10, 10, 10))
    llbc, uubc = (-2.4,) * 3, (2.4,) * 3           # h = 0.4
    llbf, uubf = (-1.0,) * 3, (1.5,) * 3           # h = 0.25: not 2:1
    assert fa.prolong3(func, llbc, uubc, funf, llbf, uubf, llbf, uubf, (1, 1, 1)) == 2
    assert "2:1" in amssbssn.last_error()
    llbf, uubf = (-1.0,) * 3, (1.0,) * 3           # h = 0.2
    assert fa.prolong3(func, llbc, uubc, funf, llbf, uubf, (-3.0,) * 3, uubf, (1, 1, 1)) == 1
    fa.restrict3(func, llbc, uubc, funf, llbf, uubf, llbc, uubc, (1, 1, 1))
    assert "does not fit" in amssbssn.last_error()
    assert np.all(func == 0) and np.all(funf == 1)


def test_config_geometry():
    import amssbssn as A
    c0 = A.config0()
    p = c0.patches[0]
    assert p.shape == (54, 54, 54) and p.h[0] == pytest.approx(1 / 3)
    X, _, _ = p.coords()
    assert X[0] == pytest.approx(-9 + 0.5 / 3) and X[3] == pytest.approx(-8 + 0.5 / 3)
    assert c0.dt == pytest.approx(0.25 / 3)
    c1 = A.config1()
    assert c1.patches[0].shape == (134, 134, 134) and c1.patches[0].h[0] == pytest.approx(0.25)
    assert c1.interior_pts == 128 ** 3


def test_uneven_level_packs_without_overlap():
    import amssbssn as A
    patches, info = A.uneven_level(37, 24, 96, seed=1, half_width=24.0)
    assert len(patches) == 37
    ext, org, w = info["extents"], info["origins"], info["width_cells"]
    for e in ext:
        assert all(24 <= v <= 96 and v % 8 == 0 for v in e)
    for a in range(len(ext)):
        assert all(0 <= org[a][d] and org[a][d] + ext[a][d] <= w for d in range(3))
        for b in range(a):
            overlap = all(org[a][d] < org[b][d] + ext[b][d] and org[b][d] < org[a][d] + ext[a][d] for d in range(3))
            assert not overlap, (a, b)
    h = patches[0].h[0]
    assert all(abs(p.h[d] - h) < 1e-12 for p in patches for d in range(3))


def test_philox_noise_amplitude_and_determinism():
    import amssbssn as A
    cfg = A.config1(n=16)
    a = cfg.initial_state(0)
    b = cfg.initial_state(0)
    for v in range(8, 14):
        assert np.array_equal(a[v], b[v])
        assert 0.9e-8 < np.max(np.abs(a[v])) < 1e-8
    assert not np.array_equal(a[8], a[9])


def test_checkpoint_plan_fits_memory_and_minimises_recompute():
    import amssbssn.gradient as G
    # 400 steps of 2.5 GB states in 288 GB: 0.9 * 288 / 2.5 = 103 states fit;
    # k = 4 needs 100 + 4 > 103, k = 5 needs 80 + 5
    assert G.plan_checkpoints(400, 2_500_000_000, 288_000_000_000) == 5
    # everything fits: a checkpoint every step, nothing recomputed
    assert G.plan_checkpoints(200, 66_000_000, 280_000_000_000) == 1
    import pytest
    with pytest.raises(MemoryError):
        G.plan_checkpoints(400, 2_500_000_000, 10_000_000_000)
