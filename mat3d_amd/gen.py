"""Seeded synthetic microstructures for the homogenization workloads.

Phase-id convention follows the reference (include/FFT_analysis3D.hpp:44-45):
phase 1 = matrix, phase 0 = inclusion.  Arrays are returned as int32 of shape
``(Nx, Ny, Nz)``; flattening in C order gives the reference's voxel index
``i = z + Nz*(y + Ny*x)`` (FFT_analysis3D.cpp:115-123).  All geometry is
periodic with minimum-image distances.
"""
from __future__ import annotations

import hashlib

import numpy as np


def sha256(ids: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(ids, dtype=np.int32).tobytes()).hexdigest()


def sphere65() -> np.ndarray:
    """The reference fixture example/3d/rve65.txt: a centred sphere of radius 15
    (phase 0 where |x-(32,32,32)|^2 <= 225) in a 65^3 matrix (phase 1)."""
    g = np.arange(65) - 32
    r2 = g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2
    return np.where(r2 <= 225, 0, 1).astype(np.int32)


def _wrap(d, n):
    return (d + n // 2) % n - n // 2


def _stamp_ball(ids, c, r, val):
    n = np.array(ids.shape)
    rr = int(np.ceil(r))
    off = np.arange(-rr, rr + 1)
    ox, oy, oz = np.meshgrid(off, off, off, indexing="ij")
    m = (ox**2 + oy**2 + oz**2) <= r * r
    ix = (c[0] + ox[m]) % n[0]
    iy = (c[1] + oy[m]) % n[1]
    iz = (c[2] + oz[m]) % n[2]
    ids[ix, iy, iz] = val


class _Centres:
    """Centres and radii of placed balls with a periodic cell list, so that an
    overlap query looks at the 27 neighbouring cells only (same decisions as
    comparing against every centre)."""

    def __init__(self, n, reach):
        self.n = n
        self.nc = max(1, n // max(1, int(np.ceil(reach))))
        self.cells = {}
        self.count = 0

    def _cell(self, c):
        return tuple(int(v) * self.nc // self.n for v in c)

    def overlaps(self, c, r):
        if self.count == 0:
            return False
        if self.nc < 3:
            keys = list(self.cells.keys())
        else:
            k = self._cell(c)
            keys = {((k[0] + i) % self.nc, (k[1] + j) % self.nc, (k[2] + l) % self.nc)
                    for i in (-1, 0, 1) for j in (-1, 0, 1) for l in (-1, 0, 1)}
        for key in keys:
            lst = self.cells.get(key)
            if not lst:
                continue
            arr = np.array(lst, dtype=np.int64)
            d = _wrap(arr[:, :3] - c, self.n)
            if np.any((d * d).sum(1) <= (arr[:, 3] + r) ** 2):
                return True
        return False

    def add(self, c, r):
        self.cells.setdefault(self._cell(c), []).append((int(c[0]), int(c[1]), int(c[2]), int(r)))
        self.count += 1


def _device_balls(n, balls, matrix):
    """The balls' voxels rasterised on the GPU (m3d_gen_balls)."""
    import ctypes as ct

    from . import lib

    ids = np.empty((n, n, n), dtype=np.int32)
    b = np.ascontiguousarray(np.array(balls, dtype=np.int32).reshape(-1, 5))
    i32p = ct.POINTER(ct.c_int32)
    rc = lib().m3d_gen_balls(n, n, n, b.ctypes.data_as(i32p), len(b), int(matrix), ids.ctypes.data_as(i32p))
    if rc != 0:
        raise RuntimeError(lib().m3d_gen_last_error().decode())
    return ids


def spheres(n, radius, frac, seed, incl=0, matrix=1, max_tries=2_000_000, device=False):
    """Random sequential addition of non-overlapping balls |d|^2 <= r^2 (integer
    centres, periodic) until the inclusion fraction reaches ``frac``.
    device=True rasterises the placed balls on the GPU (same ids)."""
    rng = np.random.default_rng(seed)
    ids = None if device else np.full((n, n, n), matrix, dtype=np.int32)
    balls = []
    placed = _Centres(n, 2 * radius + 1)
    target = frac * n**3
    count = 0
    vol1 = _ball_volume(radius)
    tries = 0
    while count < target and tries < max_tries:
        tries += 1
        c = rng.integers(0, n, size=3)
        if placed.overlaps(c, radius):
            continue
        placed.add(c, radius)
        if device:
            balls.append((int(c[0]), int(c[1]), int(c[2]), int(radius), int(incl)))
        else:
            _stamp_ball(ids, c, radius, incl)
        count += vol1
    return _device_balls(n, balls, matrix) if device else ids


def multi_spheres(n, groups, seed, matrix=1, max_tries=5_000_000, device=False):
    """Non-overlapping balls of several phases; ``groups`` is a list of
    (phase, fraction, rmin, rmax).  Radii are drawn uniformly (integer).
    device=True rasterises the placed balls on the GPU (same ids)."""
    rng = np.random.default_rng(seed)
    ids = None if device else np.full((n, n, n), matrix, dtype=np.int32)
    balls = []
    placed = _Centres(n, 2 * max(g[3] for g in groups) + 1)
    for phase, frac, rmin, rmax in groups:
        target = frac * n**3
        got = 0
        tries = 0
        while got < target and tries < max_tries:
            tries += 1
            r = int(rng.integers(rmin, rmax + 1))
            c = rng.integers(0, n, size=3)
            if placed.overlaps(c, r):
                continue
            placed.add(c, r)
            if device:
                balls.append((int(c[0]), int(c[1]), int(c[2]), int(r), int(phase)))
            else:
                _stamp_ball(ids, c, r, phase)
            got += _ball_volume(r)
    return _device_balls(n, balls, matrix) if device else ids


def _ball_volume(r):
    o = np.arange(-r, r + 1)
    return int(((o[:, None, None] ** 2 + o[None, :, None] ** 2 + o[None, None, :] ** 2) <= r * r).sum())


def _fibre_draw(rng, n, radius, half):
    """One fibre's random draws and box, in the order fibres() makes them."""
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    c = rng.uniform(0, n, size=3)
    ext = np.abs(v) * half + radius + 1
    lo = np.floor(c - ext).astype(int)
    hi = np.ceil(c + ext).astype(int)
    return v, c, lo, hi


def _device_fibres(n, radius, frac, seed, length, incl, matrix):
    """fibres() with the cylinders stamped and counted on the GPU
    (m3d_gen_cylinders); candidates are drawn ahead in the same order."""
    import ctypes as ct

    from . import lib

    rng = np.random.default_rng(seed)
    ids = np.full((n, n, n), matrix, dtype=np.int32)
    half = length / 2.0
    target = int(np.ceil(frac * n**3))
    i32p, dp = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double)
    while True:
        cyl = np.empty((32 * 64, 12))
        for k in range(len(cyl)):
            v, c, lo, hi = _fibre_draw(rng, n, radius, half)
            cyl[k, :3], cyl[k, 3:6], cyl[k, 6:9], cyl[k, 9:12] = v, c, lo, hi
        used, count = ct.c_int(0), ct.c_longlong(0)
        rc = lib().m3d_gen_cylinders(n, n, n, cyl.ctypes.data_as(dp), len(cyl), half, float(radius), int(incl), 32,
                                     target, ids.ctypes.data_as(i32p), ct.byref(used), ct.byref(count))
        if rc != 0:
            raise RuntimeError(lib().m3d_gen_last_error().decode())
        if count.value >= target:
            return ids


def fibres(n, radius, frac, seed, length=64, incl=0, matrix=1, device=False):
    """Straight finite cylinders (radius ``radius``, length ``length``) with
    axes uniform on S^2 and centres uniform in the box, overlaps allowed,
    added until the inclusion fraction reaches ``frac``.  device=True stamps
    and counts on the GPU (same ids)."""
    if device:
        return _device_fibres(n, radius, frac, seed, length, incl, matrix)
    rng = np.random.default_rng(seed)
    ids = np.full((n, n, n), matrix, dtype=np.int32)
    half = length / 2.0
    target = frac * n**3
    count = 0
    while count < target:
        for _ in range(32):
            v, c, lo, hi = _fibre_draw(rng, n, radius, half)
            ax = [np.arange(lo[k], hi[k] + 1) for k in range(3)]
            X, Y, Z = np.meshgrid(ax[0] - c[0], ax[1] - c[1], ax[2] - c[2], indexing="ij")
            t = X * v[0] + Y * v[1] + Z * v[2]
            d2 = X * X + Y * Y + Z * Z - t * t
            m = (np.abs(t) <= half) & (d2 <= radius * radius)
            ii = np.nonzero(m)
            ids[(ax[0][ii[0]]) % n, (ax[1][ii[1]]) % n, (ax[2][ii[2]]) % n] = incl
        count = int((ids == incl).sum())
    return ids


def voronoi(n, grains, seed, device=False):
    """Periodic Voronoi polycrystal on an n^3 grid (or an (Nx, Ny, Nz) grid
    when n is a tuple): phase id = index of the nearest seed (minimum image).
    Uses scipy's periodic k-d tree, or with device=True the GPU's exhaustive
    search (m3d_gen_voronoi; same ids up to exact distance ties)."""
    shape = (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)
    rng = np.random.default_rng(seed)
    box = np.array(shape, dtype=np.float64)
    pts = rng.uniform(0, 1, size=(grains, 3)) * box
    if device:
        import ctypes as ct

        from . import lib

        ids = np.empty(shape, dtype=np.int32)
        p = np.ascontiguousarray(pts)
        rc = lib().m3d_gen_voronoi(*shape, p.ctypes.data_as(ct.POINTER(ct.c_double)), grains,
                                   ids.ctypes.data_as(ct.POINTER(ct.c_int32)))
        if rc != 0:
            raise RuntimeError(lib().m3d_gen_last_error().decode())
        return ids
    from scipy.spatial import cKDTree

    tree = cKDTree(pts, boxsize=box)
    X, Y, Z = np.meshgrid(*(np.arange(v) + 0.0 for v in shape), indexing="ij")
    q = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) % box
    _, idx = tree.query(q, k=1)
    return idx.reshape(shape).astype(np.int32)


# ---------------------------------------------------------------- workloads
MAT_C1 = np.array([[400e9, 0.20, np.inf, 0, 1, 0, 0],   # phase 0: inclusion
                   [70e9, 0.30, np.inf, 0, 1, 0, 0]])   # phase 1: matrix
MAT_C2 = np.array([[300e9, 0.25, np.inf, 0, 1, 0, 0],   # phase 0: fibre
                   [3e9, 0.35, np.inf, 0, 1, 0, 0]])    # phase 1: matrix
MAT_C4 = np.array([[400e9, 0.20, np.inf, 0, 1, 0, 0],   # phase 0: elastic inclusion
                   [70e9, 0.30, 100e6, 1e9, 1, 0, 0]])  # phase 1: J2 matrix


def c1_rve(n=64, device=False):
    return spheres(n, 6, 0.20, seed=1, device=device)


def c2_rve(n=256, device=False):
    return fibres(n, 4, 0.30, seed=2, device=device)


def c4_rve(n=192, device=False):
    return spheres(n, 8, 0.25, seed=4, device=device)


def c5_rve(n=512, device=False):
    return multi_spheres(n, [(2, 0.15, 3, 10), (0, 0.10, 3, 10)], seed=5, matrix=1, device=device)


def c3_batch(count=64, sizes=(32, 48, 64, 96, 128), device=False):
    rng = np.random.default_rng(1000)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        g = int(rng.integers(20, 201))
        r = np.random.default_rng(1000 + i)
        E = r.uniform(50e9, 300e9, size=g)
        nu = r.uniform(0.2, 0.35, size=g)
        mat = np.zeros((g, 7))
        mat[:, 0], mat[:, 1], mat[:, 2], mat[:, 4] = E, nu, np.inf, 1
        out.append((n, voronoi(n, g, 1000 + i, device=device), mat))
    return out
