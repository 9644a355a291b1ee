"""The reference's on-disk formats, written from (and read into) a level on
the GPU: Parallel::writefile's binary dump (Parallel.C:1049-1071) and the
one-level checkpoint of checkpoint::writecheck_cgh / readcheck_cgh
(checkpoint.C:132-292).  The files are parsed here in the reference readers'
own read order (plot_binary_data.py:37-43, checkpoint.C:208-250)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _two_patch_config():
    import amssbssn as A
    h = 0.4
    a = A.PatchGeom((18, 16, 20), (-3.6, -3.2, -4.0, -3.6 + 18 * h, -3.2 + 16 * h, -4.0 + 20 * h), 3)
    b = A.PatchGeom((14, 16, 20), (a.bbox[3] - 6 * h, -3.2, -4.0, a.bbox[3] + 8 * h, -3.2 + 16 * h,
                                   -4.0 + 20 * h), 3)
    return A.Config("two patches", [a, b], dt=0.25 * h, steps=1, punctures=[(0.5, 1.0, 0.0, 0.0)])


def test_write_bin_reads_back_with_the_reference_reader(tmp_path):
    import amssbssn as A
    cfg = _two_patch_config()
    lv = A.Level.from_config(cfg, gauge=0)
    assert lv.step(cfg.dt, 1) == 0
    path = str(tmp_path / "Lev00-01_Lap_00001.bin")
    lv.write_bin(path, 1, A.LIST_STATE, 17, 0.125)
    want = lv.get_field(1, A.LIST_STATE, 17)
    p = cfg.patches[1]
    with open(path, "rb") as f:
        t = np.fromfile(f, dtype=np.float64, count=1)
        nx, ny, nz = np.fromfile(f, dtype=np.int32, count=3)
        xmin, xmax = np.fromfile(f, dtype=np.float64, count=2)
        ymin, ymax = np.fromfile(f, dtype=np.float64, count=2)
        zmin, zmax = np.fromfile(f, dtype=np.float64, count=2)
        data = np.fromfile(f, dtype=np.float64)
    lv.close()
    assert t[0] == 0.125 and (nx, ny, nz) == p.shape
    assert (xmin, ymin, zmin, xmax, ymax, zmax) == p.bbox
    assert np.array_equal(data.reshape((nz, ny, nx)), want)


def _parse_checkpoint(path, shapes):
    """checkpoint.C:208-250 and 253-288, for one level."""
    with open(path, "rb") as f:
        time = np.fromfile(f, np.float64, 1)[0]
        levels, movls, nbh = np.fromfile(f, np.int32, 3)
        grids = np.fromfile(f, np.int32, levels)
        lt = np.fromfile(f, np.float64, levels)
        tab = []
        for _ in range(grids[0]):
            bb = np.fromfile(f, np.float64, 6)
            sh = np.fromfile(f, np.int32, 3)
            hd = np.fromfile(f, np.float64, 3)
            tab.append((tuple(bb), tuple(sh), tuple(hd)))
        porg = [tuple(np.fromfile(f, np.float64, 3)) for _ in range(nbh)]
        data = []
        for sh in shapes:
            nn = int(np.prod(sh))
            data.append([np.fromfile(f, np.float64, nn).reshape(sh[::-1]) for _ in range(48)])
        rest = f.read()
    return dict(time=time, levels=levels, movls=movls, grids=grids, lt=lt, tab=tab, porg=porg, data=data,
                rest=rest)


def test_checkpoint_layout_and_restart_continue_bit_for_bit(tmp_path):
    import amssbssn as A
    cfg = _two_patch_config()
    run = A.Level.from_config(cfg, gauge=0)
    assert run.step(cfg.dt, 2) == 0
    path = str(tmp_path / "bssn_cgh.CHK")
    porg = [(1.0, 0.0, 0.0), (-1.0, 0.25, 0.0)]
    handles = [(1.0, 0.0, 0.0), (0.0, 0.0, 0.0)]
    run.checkpoint_write(path, 2 * cfg.dt, movls=1, porg=porg, handles=handles)
    shapes = [p.shape for p in cfg.patches]
    c = _parse_checkpoint(path, shapes)
    assert c["time"] == 2 * cfg.dt and c["levels"] == 1 and c["movls"] == 1 and list(c["grids"]) == [2]
    assert c["lt"][0] == 2 * cfg.dt and c["porg"] == porg and c["rest"] == b""
    for (bb, sh, hd), p, h in zip(c["tab"], cfg.patches, handles):
        assert bb == p.bbox and sh == p.shape and hd == h
    for pi in range(2):
        for v in range(24):
            assert np.array_equal(c["data"][pi][v], run.get_field(pi, A.LIST_STATE, v))
            assert np.array_equal(c["data"][pi][24 + v], run.get_field(pi, A.LIST_OLD, v))
    # restart into a fresh level of the same geometry, then both continue
    rst = A.Level(cfg.patches, gauge=0, eps=cfg.eps, chitiny=cfg.chitiny)
    t, movls, pg = rst.checkpoint_read(path)
    assert t == 2 * cfg.dt and movls == 1 and pg == porg
    for lv in (run, rst):  # one kernel path for both (same contractions, so bitwise)
        assert lv.step(cfg.dt, 2) == 0
    for pi in range(2):
        for lst in (A.LIST_STATE, A.LIST_OLD):
            for v in range(24):
                assert np.array_equal(rst.get_field(pi, lst, v), run.get_field(pi, lst, v)), (pi, lst, v)
    run.close()
    rst.close()


def test_checkpoint_of_another_geometry_is_refused(tmp_path):
    import amssbssn as A
    cfg = _two_patch_config()
    lv = A.Level.from_config(cfg, gauge=0)
    path = str(tmp_path / "a_cgh.CHK")
    lv.checkpoint_write(path, 0.0)
    one = A.Level(cfg.patches[:1], gauge=0)
    with pytest.raises(A.AmssError, match="patch count"):
        one.checkpoint_read(path)
    with pytest.raises(A.AmssError, match="cannot open"):
        lv.checkpoint_read(str(tmp_path / "missing_cgh.CHK"))
    lv.checkpoint_write(path, 0.0, porg=[(1.0, 2.0, 3.0)] * 3)
    with pytest.raises(A.AmssError, match="more punctures"):
        lv.checkpoint_read(path, maxbh=2)
    lv.close()
    one.close()
