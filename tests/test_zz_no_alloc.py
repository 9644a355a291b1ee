"""amss_rhs_bssn_dev_noalloc (the entry point that allocates nothing) against
amss_rhs_bssn_dev on the same device inputs, through the C ABI: two calls in
a row, the second on a smaller block, vacuum (matter NULL) and no caller
flag, so both of its scratch uses are exercised; then with matter arrays and
a caller flag; an oversized vacuum block is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_state

NF = 24


def _ptrs(t):
    return (C.c_void_p * t.shape[0])(*[t[i].data_ptr() for i in range(t.shape[0])])


def _call(name, S, h, co, gauge, matter=None, flag=None, symmetry=0, origin=(0.0, 0.0, 0.0)):
    import torch
    from amssbssn._lib import lib
    shape = tuple(S.shape[1:])
    dev = S.device
    out = {k: torch.full((n,) + shape, np.nan, dtype=torch.float64, device=dev)
           for k, n in (("rhs", NF), ("gam", 18), ("ric", 6), ("res", 7))}
    ex = (C.c_int * 3)(shape[2], shape[1], shape[0])
    s = torch.cuda.current_stream(dev)
    rc = getattr(lib(), name)(ex, (C.c_double * 3)(*h), _ptrs(S), _ptrs(out["rhs"]),
                              _ptrs(matter) if matter is not None else None, _ptrs(out["gam"]), _ptrs(out["ric"]),
                              _ptrs(out["res"]), symmetry, (C.c_double * 3)(*origin), 0.02, co, gauge,
                              C.c_void_p(flag.data_ptr()) if flag is not None else None, C.c_void_p(s.cuda_stream))
    s.synchronize()
    return rc, out


def _same(a, b, what):
    import torch
    for k in ("rhs", "gam", "ric", "res"):
        x, y = a[k], b[k]
        # NaN where neither wrote (res when co == 1, untouched edges) on both sides
        assert torch.equal(torch.isnan(x), torch.isnan(y)), f"{what}: {k} written at different points"
        m = ~torch.isnan(x)
        assert torch.equal(x[m], y[m]), f"{what}: {k} differs"


@pytest.mark.gpu
def test_noalloc_rhs_matches_allocating_entry_point():
    import torch
    dev = torch.device("cuda:0")
    for call, (nx, ny, nz), co, gauge in ((1, (20, 18, 22), 0, 0), (2, (12, 10, 11), 1, 2)):
        st = random_state((nz, ny, nx), seed=300 + call)
        S = torch.from_numpy(np.stack(st)).to(dev)
        h = (0.3, 0.27, 0.33)
        print(f"no-alloc call {call}: {nx}x{ny}x{nz} vacuum, no flag, co {co}, gauge {gauge}", flush=True)
        r0, a = _call("amss_rhs_bssn_dev", S, h, co, gauge)
        r1, b = _call("amss_rhs_bssn_dev_noalloc", S, h, co, gauge)
        assert r0 == 0 and r1 == 0
        _same(a, b, f"call {call}")
        assert not torch.isnan(b["rhs"][:, 3:-3, 3:-3, 3:-3]).any()


@pytest.mark.gpu
def test_noalloc_rhs_with_matter_and_flag():
    import torch
    dev = torch.device("cuda:0")
    nx, ny, nz = 14, 13, 12
    st = random_state((nz, ny, nx), seed=310)
    rng = np.random.default_rng(311)
    S = torch.from_numpy(np.stack(st)).to(dev)
    M = torch.from_numpy(0.01 * rng.standard_normal((10, nz, ny, nx))).to(dev)
    f0 = torch.zeros(1, dtype=torch.int32, device=dev)
    f1 = torch.zeros(1, dtype=torch.int32, device=dev)
    r0, a = _call("amss_rhs_bssn_dev", S, (0.3, 0.3, 0.3), 0, 0, matter=M, flag=f0)
    r1, b = _call("amss_rhs_bssn_dev_noalloc", S, (0.3, 0.3, 0.3), 0, 0, matter=M, flag=f1)
    assert r0 == 0 and r1 == 0 and int(f0.item()) == 0 and int(f1.item()) == 0
    _same(a, b, "matter")
    # a NaN input is reported through the caller's flag by both
    S[5, 6, 6, 6] = float("nan")
    _call("amss_rhs_bssn_dev", S, (0.3, 0.3, 0.3), 1, 0, matter=M, flag=f0)
    _call("amss_rhs_bssn_dev_noalloc", S, (0.3, 0.3, 0.3), 1, 0, matter=M, flag=f1)
    assert int(f0.item()) == 1 and int(f1.item()) == 1


@pytest.mark.gpu
def test_noalloc_rhs_refuses_an_oversized_vacuum_block():
    import torch
    from amssbssn._lib import lib
    n = 137  # 137^3 = 2,571,353 points > 2,500,000
    dev = torch.device("cuda:0")
    S = torch.zeros((NF, n, n, n), dtype=torch.float64, device=dev)
    R = torch.zeros((NF, n, n, n), dtype=torch.float64, device=dev)
    G = torch.zeros((31, n, n, n), dtype=torch.float64, device=dev)
    ex = (C.c_int * 3)(n, n, n)
    rc = lib().amss_rhs_bssn_dev_noalloc(ex, (C.c_double * 3)(0.3, 0.3, 0.3), _ptrs(S), _ptrs(R), None,
                                         _ptrs(G[:18]), _ptrs(G[18:24]), _ptrs(G[24:]), 0,
                                         (C.c_double * 3)(0.0, 0.0, 0.0), 0.02, 0, 0, None,
                                         C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    assert rc == 2
    assert b"exceeds the scratch" in lib().amss_last_error()
    torch.cuda.synchronize()
    assert float(R.abs().max()) == 0.0  # nothing was launched
