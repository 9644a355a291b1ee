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
