"""GPU parity of the inter-level transfer (prolong3_ / restrict3_ and the
device entry points) against the oracle, itself bit-identical to the
reference's prolongrestrict_cell.f90 on the same random 2:1 pairs and
regions (tests/test_oracle_reference.py).  The GPU contracts the 6x6x6
block with FMAs, so agreement is to rounding: max |gpu - oracle| <= 1e-14
max |oracle| per array, and the same points written."""
import numpy as np
import pytest

from helpers import transfer_cases

pytestmark = pytest.mark.gpu
TOL = 1e-14


def _close(a, b):
    return np.max(np.abs(a - b)) <= TOL * max(np.max(np.abs(b)), 1e-300)


def test_prolong3_restrict3_match_oracle(orc):
    from amssbssn import fortran_abi as fa
    n_p = n_r = 0
    for c in transfer_cases(11, 120):
        o = c["funf"].copy()
        if orc.prolong3(c["func"], c["llbc"], c["uubc"], o, c["llbf"], c["uubf"], c["llbp"], c["uubp"],
                        c["soa"]) == 0:
            g = c["funf"].copy()
            assert fa.prolong3(c["func"], c["llbc"], c["uubc"], g, c["llbf"], c["uubf"], c["llbp"], c["uubp"],
                               c["soa"]) == 0
            assert _close(g, o) and np.array_equal(g == c["funf"], o == c["funf"])
            n_p += 1
        if c["llbr"] is not None:
            o = c["func"].copy()
            if orc.restrict3(o, c["llbc"], c["uubc"], c["funf"], c["llbf"], c["uubf"], c["llbr"], c["uubr"],
                             c["soa"]) == 0:
                g = c["func"].copy()
                fa.restrict3(g, c["llbc"], c["uubc"], c["funf"], c["llbf"], c["uubf"], c["llbr"], c["uubr"],
                             c["soa"])
                assert _close(g, o) and np.array_equal(g == c["func"], o == c["func"])
                n_r += 1
    assert n_p > 50 and n_r > 50, (n_p, n_r)


def test_device_transfer_entry_points_match_oracle(orc):
    import ctypes as C
    import torch
    import amssbssn
    L = amssbssn.lib()
    d3 = lambda v: (C.c_double * 3)(*[float(x) for x in v])  # noqa: E731
    ex = lambda a: (C.c_int * 3)(*a.shape[::-1])  # noqa: E731
    dev = torch.device("cuda:0")
    done = 0
    for c in transfer_cases(12, 60):
        if c["llbr"] is None:
            continue
        o_f, o_c = c["funf"].copy(), c["func"].copy()
        if orc.prolong3(c["func"], c["llbc"], c["uubc"], o_f, c["llbf"], c["uubf"], c["llbp"], c["uubp"],
                        c["soa"]) or orc.restrict3(o_c, c["llbc"], c["uubc"], c["funf"], c["llbf"], c["uubf"],
                                                   c["llbr"], c["uubr"], c["soa"]):
            continue
        tc, tf = torch.from_numpy(c["func"]).to(dev), torch.from_numpy(c["funf"]).to(dev)
        s = torch.cuda.current_stream(dev)
        assert L.amss_prolong3_dev(d3(c["llbc"]), d3(c["uubc"]), ex(c["func"]), C.c_void_p(tc.data_ptr()),
                                   d3(c["llbf"]), d3(c["uubf"]), ex(c["funf"]), C.c_void_p(tf.data_ptr()),
                                   d3(c["llbp"]), d3(c["uubp"]), d3(c["soa"]), C.c_void_p(s.cuda_stream)) == 0
        s.synchronize()
        assert _close(tf.cpu().numpy(), o_f)
        tf = torch.from_numpy(c["funf"]).to(dev)
        assert L.amss_restrict3_dev(d3(c["llbc"]), d3(c["uubc"]), ex(c["func"]), C.c_void_p(tc.data_ptr()),
                                    d3(c["llbf"]), d3(c["uubf"]), ex(c["funf"]), C.c_void_p(tf.data_ptr()),
                                    d3(c["llbr"]), d3(c["uubr"]), d3(c["soa"]), C.c_void_p(s.cuda_stream)) == 0
        s.synchronize()
        assert _close(tc.cpu().numpy(), o_c)
        done += 1
    assert done > 10, done


def test_prolong_then_restrict_returns_a_quintic():
    """Size-independent property: a degree-5 polynomial prolonged onto a
    fine box and restricted back reproduces the coarse values (both
    transfers are exact on it), at a few hundred thousand points."""
    from amssbssn import fortran_abi as fa
    CD, FD = 0.1, 0.05
    extc = np.array([64, 60, 56])
    llbc = -CD * extc / 2
    extf = np.array([80, 72, 64])
    llbf = llbc + FD * np.array([24, 26, 22])
    cen = lambda lo, n, d: [lo[a] + (np.arange(n[a]) + 0.5) * d for a in range(3)]  # noqa: E731
    Xc, Xf = cen(llbc, extc, CD), cen(llbf, extf, FD)
    q = lambda x, y, z: 0.2 + x ** 5 - 0.3 * y ** 4 * z + x * y * z ** 3 - z ** 5  # noqa: E731
    zc, yc, xc = np.meshgrid(Xc[2], Xc[1], Xc[0], indexing="ij")
    func = q(xc, yc, zc)
    funf = np.zeros(tuple(extf[::-1]))
    uubc, uubf = llbc + CD * extc, llbf + FD * extf
    assert fa.prolong3(func, llbc, uubc, funf, llbf, uubf, llbf, uubf, (1, 1, 1)) == 0
    back = np.full(func.shape, 5.0)
    fa.restrict3(back, llbc, uubc, funf, llbf, uubf, llbf + 4 * FD, uubf - 4 * FD, (1, 1, 1))
    m = back != 5.0
    assert m.sum() > 30000
    assert np.max(np.abs(back[m] - func[m])) < 1e-12 * np.max(np.abs(func))
