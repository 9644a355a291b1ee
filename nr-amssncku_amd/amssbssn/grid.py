"""Patch geometry and initial data of the BSSN benchmark configurations.

Grid functions are numpy float64 arrays of shape (nz, ny, nx): the
reference's Fortran (nx, ny, nz) layout with i fastest.  Coordinates are
cell centred, X[i] = xmin + (i + 1/2) h (Block.C:61-64).  Stored fields
follow the reference's offset convention: chi - 1, alpha - 1 and g_ii - 1
(bssn_rhs.f90:146-150, initial_puncture.f90:296-310).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np

NF = 24
FIELDS = ["chi", "trK", "dxx", "gxy", "gxz", "dyy", "gyz", "dzz",
          "Axx", "Axy", "Axz", "Ayy", "Ayz", "Azz", "Gamx", "Gamy", "Gamz",
          "Lap", "betax", "betay", "betaz", "dtSfx", "dtSfy", "dtSfz"]
CHI, TRK, DXX, AXX, LAP = 0, 1, 2, 8, 17


@dataclass
class PatchGeom:
    """One patch (= one Block): points per axis with ghosts, and the bbox of
    the cell-centred array (ghost cells included)."""
    shape: tuple            # (nx, ny, nz)
    bbox: tuple             # (xmin, ymin, zmin, xmax, ymax, zmax)
    gw: int = 3

    @property
    def h(self):
        return tuple((self.bbox[3 + a] - self.bbox[a]) / self.shape[a] for a in range(3))

    def coords(self):
        return [self.bbox[a] + (np.arange(self.shape[a]) + 0.5) * self.h[a] for a in range(3)]

    @property
    def npts(self):
        return int(np.prod(self.shape))

    @property
    def interior_pts(self):
        return int(np.prod([s - 2 * self.gw for s in self.shape]))


def cube_patch(n_interior: int, half_width: float, gw: int = 3) -> PatchGeom:
    """n_interior^3 points over [-half_width, half_width]^3 plus gw ghost cells per side."""
    h = 2.0 * half_width / n_interior
    lo = -half_width - gw * h
    hi = half_width + gw * h
    n = n_interior + 2 * gw
    return PatchGeom((n, n, n), (lo, lo, lo, hi, hi, hi), gw)


# ---------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), vectorised.
_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(x, np.uint64) & _MASK for x in (c0, c1, c2, c3))
    k0 = np.uint64(k0) & _MASK
    k1 = np.uint64(k1) & _MASK
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK, lo1, (hi0 ^ c3 ^ k1) & _MASK, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def philox_uniform(npts: int, field_id: int, seed: int) -> np.ndarray:
    """U[0,1) per point: counter (point index, field id, 0, 0), key (seed, 0);
    53 bits from the first two output words."""
    idx = np.arange(npts, dtype=np.uint64)
    r0, r1, _, _ = philox4x32_10(idx, np.full(npts, field_id, np.uint64), 0, 0, seed, 0)
    u53 = (r0 << np.uint64(21)) | (r1 >> np.uint64(11))
    return u53.astype(np.float64) * (2.0 ** -53)


# ---------------------------------------------------------------------------
def puncture_state(geom: PatchGeom, punctures, eps2: float = 0.01, a_noise: float = 0.0,
                   noise_seed: int = 0):
    """Brill-Lindquist conformal factor psi = 1 + sum m/(2 r), r = sqrt(|x - c|^2
    + eps2); chi = psi^-4, alpha = psi^-2, conformally flat metric, K = 0,
    shift = B = Gamma = 0; optional A_ij noise U(-a, a) from Philox."""
    X, Y, Z = geom.coords()
    z, y, x = np.meshgrid(Z, Y, X, indexing="ij")
    psi = np.ones_like(x)
    for (m, cx, cy, cz) in punctures:
        psi += m / (2.0 * np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 + eps2))
    st = [np.zeros(x.shape) for _ in range(NF)]
    st[CHI] = psi ** -4 - 1.0
    st[LAP] = psi ** -2 - 1.0
    if a_noise:
        for s in range(6):
            u = philox_uniform(geom.npts, AXX + s, noise_seed).reshape(x.shape)
            st[AXX + s] = a_noise * (2.0 * u - 1.0)
    return st


# ---------------------------------------------------------------------------
@dataclass
class Config:
    name: str
    patches: list
    dt: float
    steps: int
    eps: float = 0.02
    chitiny: float = 1e-5
    punctures: list = field(default_factory=list)
    a_noise: float = 0.0
    noise_seed: int = 0

    def initial_state(self, i: int):
        return puncture_state(self.patches[i], self.punctures, a_noise=self.a_noise, noise_seed=self.noise_seed)

    @property
    def interior_pts(self):
        return sum(p.interior_pts for p in self.patches)


def config0() -> Config:
    """One 48^3 patch (+3 ghosts) over [-8,8]^3, single puncture m=1, one RK4 step."""
    g = cube_patch(48, 8.0)
    return Config("single puncture 48^3 + 3 ghosts over [-8,8]^3", [g], dt=0.25 * g.h[0], steps=1,
                  punctures=[(1.0, 0.0, 0.0, 0.0)])


def config1(n: int = 128, steps: int = 100) -> Config:
    """One n^3 patch (+3 ghosts) over [-16,16]^3, two punctures m=0.5 at (+-3,0,0),
    A_ij noise 1e-8 (Philox seed 0)."""
    g = cube_patch(n, 16.0)
    return Config(f"two punctures {n}^3 + 3 ghosts over [-16,16]^3", [g], dt=0.25 * g.h[0], steps=steps,
                  punctures=[(0.5, 3.0, 0.0, 0.0), (0.5, -3.0, 0.0, 0.0)], a_noise=1e-8, noise_seed=0)


def pack_patches(ext, width_cells):
    """Shelf packing of boxes (cells per axis) into a cube of width_cells:
    slabs along z, rows along y, boxes along x.  Returns origins or None."""
    order = sorted(range(len(ext)), key=lambda i: (-ext[i][2], -ext[i][1], -ext[i][0]))
    org = [None] * len(ext)
    zs, hz, ys, hy, xc = 0, 0, 0, 0, 0
    for i in order:
        ex, ey, ez = ext[i]
        if xc + ex > width_cells:          # new row
            ys, xc, hy = ys + hy, 0, 0
        if ys + ey > width_cells:          # new slab
            zs, ys, xc, hy, hz = zs + hz, 0, 0, 0, 0
        if zs + ez > width_cells or xc + ex > width_cells or ys + ey > width_cells:
            return None
        org[i] = (xc, ys, zs)
        xc += ex
        hy, hz = max(hy, ey), max(hz, ez)
    return org


def uneven_level(npatch: int, lo: int, hi: int, seed: int, half_width: float, gw: int = 3):
    """npatch patches with seeded extents (multiples of 8 in [lo, hi]) packed
    without overlap in [-half_width, half_width]^3 at one grid spacing."""
    rng = np.random.default_rng(seed)
    ext = [tuple(int(8 * v) for v in rng.integers(lo // 8, hi // 8 + 1, size=3)) for _ in range(npatch)]
    vol = sum(e[0] * e[1] * e[2] for e in ext)
    w = 8 * int(np.ceil(vol ** (1 / 3) / 8))
    while True:
        org = pack_patches(ext, w)
        if org is not None:
            break
        w += 8
    h = 2.0 * half_width / w
    patches = []
    for e, o in zip(ext, org):
        bb = tuple(-half_width + (o[a] - gw) * h for a in range(3)) + \
             tuple(-half_width + (o[a] + e[a] + gw) * h for a in range(3))
        patches.append(PatchGeom(tuple(e[a] + 2 * gw for a in range(3)), bb, gw))
    return patches, dict(extents=ext, origins=org, width_cells=w, h=h)


def config2(steps: int = 50) -> Config:
    """37 uneven patches (multiples of 8 in [24,96], seed 1) in [-24,24]^3, the two-puncture data of config1()."""
    patches, _ = uneven_level(37, 24, 96, seed=1, half_width=24.0)
    h = patches[0].h[0]
    return Config("37 uneven patches in [-24,24]^3", patches, dt=0.25 * h, steps=steps,
                  punctures=[(0.5, 3.0, 0.0, 0.0), (0.5, -3.0, 0.0, 0.0)], a_noise=1e-8, noise_seed=0)


def layout_json(cfg: Config) -> str:
    return json.dumps([dict(shape=list(p.shape), bbox=list(p.bbox), gw=p.gw) for p in cfg.patches])
