"""Shared test helpers: seeded grids and states, and the parity metric."""
import numpy as np

NF = 24


def grid(nx, ny, nz, h=(0.3, 0.27, 0.33), origin=(-1.1, -0.9, -1.3)):
    X = origin[0] + (np.arange(nx) + 0.5) * h[0]
    Y = origin[1] + (np.arange(ny) + 0.5) * h[1]
    Z = origin[2] + (np.arange(nz) + 0.5) * h[2]
    bbox = (origin[0], origin[1], origin[2], origin[0] + nx * h[0], origin[1] + ny * h[1], origin[2] + nz * h[2])
    return X, Y, Z, bbox


def random_state(shape, seed, amp=0.05, dtsf=True):
    """Smooth-ish random perturbation of flat space: every field varies, the
    shift takes both signs, chi and alpha sit well inside (-1, 0)."""
    rng = np.random.default_rng(seed)
    st = [amp * rng.standard_normal(shape) for _ in range(NF)]
    st[0] = -0.3 + amp * rng.standard_normal(shape)
    st[17] = -0.2 + amp * rng.standard_normal(shape)
    if not dtsf:
        for v in (21, 22, 23):
            st[v][:] = 0.0
    return st


def rel_err(a, b):
    """max |a - b| / max |b| (absolute when b is identically zero)."""
    m = float(np.max(np.abs(b)))
    d = float(np.max(np.abs(a - b)))
    return d / m if m > 0 else d


def transfer_cases(seed, n):
    """Random 2:1 cell-centred coarse/fine pairs and wanted regions for
    prolong3 / restrict3: fine boxes offset from the coarse by even and odd
    numbers of fine cells, on both sides of the coarse lower bound, regions
    touching the arrays' lower edges (mirrored halo reads).  Yields dicts."""
    rng = np.random.default_rng(seed)
    for _ in range(n):
        CD = float(rng.choice([0.4, 0.5, 0.25]))
        FD = CD / 2
        extc = rng.integers(10, 17, 3)
        llbc = np.array([-CD * extc[a] / 2 + CD * rng.integers(-1, 2) for a in range(3)])
        extf = rng.integers(8, 15, 3)
        llbf = llbc + FD * rng.integers(-2, 12, 3)
        func = rng.standard_normal(tuple(extc[::-1]))
        funf = rng.standard_normal(tuple(extf[::-1]))
        i0 = rng.integers(0, extf - 1)
        i1 = np.minimum(i0 + rng.integers(1, 8, 3), extf)
        soa = tuple(rng.choice([-1.0, 1.0], 3))
        c = dict(llbc=llbc, uubc=llbc + CD * extc, llbf=llbf, uubf=llbf + FD * extf, func=func, funf=funf,
                 llbp=llbf + FD * i0, uubp=llbf + FD * i1, soa=soa, llbr=None, uubr=None)
        lo = np.ceil((llbf + 2 * FD - llbc) / CD - 1e-9).astype(int)
        hi = np.floor((c["uubf"] - 2 * FD - llbc) / CD + 1e-9).astype(int)
        if np.all(hi > lo):
            j0 = np.array([rng.integers(a, b) for a, b in zip(lo, hi)])
            j1 = np.array([rng.integers(a + 1, b + 1) for a, b in zip(j0, hi)])
            c["llbr"], c["uubr"] = llbc + CD * j0, llbc + CD * j1
        yield c
