"""Seeded synthetic CT-like registration pairs (test and benchmark inputs).

Volumes are numpy arrays of shape (o, n, m) in C order, so that ``vol.ravel()``
is the reference's voxel order ``i + j*m + k*m*n`` with i along NIfTI dim 1
(deedsBCV0.cpp:126-137, index convention of every reference kernel).

Each workload below is one of the benchmark configurations:

* ``small``      96x96x80, 8-12 ellipsoids, noise 20 HU, B-spline warp (spacing 24, max |u| 6)
* ``full``       256x256x128, 20 ellipsoids, noise 25 HU, warp (spacing 32, max 12), gamma remap
* ``crossmodal`` 192x160x192, non-monotone 6-knot transfer, Rician-like noise 15, warp (32, 10)
* ``atlas``      16 pairs, dims U[128,224]xU[128,192]xU[96,176], generated as ``crossmodal``
* ``soft``       160^3 float32 pair generated as ``crossmodal``

The generator is this project's own; nothing of it comes from the reference,
which ships no test data of its own (its two label volumes have no scans).
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

AIR = -1000.0
BODY = 40.0


def _body_and_organs(shape_mno, n_organs, rng):
    m, n, o = shape_mno
    vol = np.full((o, n, m), AIR, dtype=np.float64)
    cx, cy = (m - 1) / 2.0, (n - 1) / 2.0
    ax, ay = 0.44 * m, 0.36 * n
    zz = np.arange(o, dtype=np.float64)[:, None, None]
    yy = np.arange(n, dtype=np.float64)[None, :, None]
    xx = np.arange(m, dtype=np.float64)[None, None, :]
    body = (((xx - cx) / ax) ** 2 + ((yy - cy) / ay) ** 2 <= 1.0) & (zz >= 0.04 * o) & (zz <= 0.96 * o)
    vol[body] = BODY
    for _ in range(n_organs):
        # centre inside the body's inner part, random radii and orientation
        r = rng.uniform(0.0, 0.65)
        t = rng.uniform(0, 2 * np.pi)
        c = np.array([cx + r * ax * np.cos(t), cy + r * ay * np.sin(t), rng.uniform(0.2, 0.8) * o])
        radii = rng.uniform(0.05, 0.16, size=3) * np.array([m, n, o])
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        # the ellipsoid lies within its largest radius of the centre: evaluate that box only
        # (elementwise, so every voxel gets the value a whole-volume evaluation would give)
        R = radii.max() + 2
        lo = [max(0, int(np.floor(c[a] - R))) for a in range(3)]
        hi = [min(s, int(np.ceil(c[a] + R)) + 1) for a, s in enumerate((m, n, o))]
        sl = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
        Z, Y, X = np.meshgrid(np.arange(lo[2], hi[2], dtype=np.float64), np.arange(lo[1], hi[1], dtype=np.float64),
                              np.arange(lo[0], hi[0], dtype=np.float64), indexing="ij")
        d = np.stack([X - c[0], Y - c[1], Z - c[2]], axis=-1) @ q
        inside = ((d / radii) ** 2).sum(-1) <= 1.0
        sub = vol[sl]
        sub[inside & body[sl]] = rng.uniform(-200.0, 300.0)
    return vol


def _bspline_basis(npts, spacing, ncp):
    """Cubic B-spline basis matrix (npts x ncp) on a uniform knot grid."""
    t = np.arange(npts, dtype=np.float64) / spacing
    cell = np.floor(t).astype(int)
    f = t - cell
    w = [(1 - f) ** 3 / 6.0, (3 * f ** 3 - 6 * f ** 2 + 4) / 6.0,
         (-3 * f ** 3 + 3 * f ** 2 + 3 * f + 1) / 6.0, f ** 3 / 6.0]
    B = np.zeros((npts, ncp))
    for a in range(4):
        np.add.at(B, (np.arange(npts), cell + a), w[a])
    return B


def bspline_field(shape_mno, spacing, max_u, rng):
    """Random smooth displacement field (3, o, n, m) with max |u| == max_u voxels."""
    m, n, o = shape_mno
    ncp = [int(np.ceil(s / spacing)) + 3 for s in (m, n, o)]
    Bm, Bn, Bo = (_bspline_basis(s, spacing, c) for s, c in zip((m, n, o), ncp))
    comps = []
    for _ in range(3):
        cp = rng.normal(size=(ncp[2], ncp[1], ncp[0]))
        f = np.einsum("kji,zk->zji", cp, Bo)
        f = np.einsum("zji,yj->zyi", f, Bn)
        f = np.einsum("zyi,xi->zyx", f, Bm)
        comps.append(f)
    u = np.stack(comps)
    mag = np.sqrt((u ** 2).sum(0)).max()
    return u * (max_u / max(mag, 1e-12))


def warp_volume(vol, field, order=1):
    o, n, m = vol.shape
    grid = np.meshgrid(np.arange(o), np.arange(n), np.arange(m), indexing="ij")
    # field components are displacements along (m, n, o) = (x, y, z)
    coords = [grid[0] + field[2], grid[1] + field[1], grid[2] + field[0]]
    return ndimage.map_coordinates(vol, coords, order=order, mode="nearest")


def _to_int16(v):
    return np.clip(np.rint(v), -32768, 32767).astype(np.int16)


def _gamma_remap(v, rng):
    g = rng.uniform(0.75, 1.35)
    lo, hi = AIR, 400.0
    x = np.clip((v - lo) / (hi - lo), 0.0, None)
    return lo + (hi - lo) * x ** g


def _piecewise_transfer(v, rng):
    lo, hi = AIR, 400.0
    knots_x = np.sort(rng.uniform(lo, hi, size=6))
    knots_y = rng.uniform(lo, hi, size=6)
    xs = np.concatenate([[lo - 1e4], knots_x, [hi + 1e4]])
    ys = np.concatenate([[knots_y[0]], knots_y, [knots_y[-1]]])
    return np.interp(v, xs, ys)


def make_pair(shape_mno, seed, *, n_organs=None, noise=20.0, spacing=24, max_u=6.0,
              contrast="none", rician=False, dtype="int16"):
    """Return (fixed, moving) arrays of shape (o, n, m)."""
    rng = np.random.default_rng(seed)
    if n_organs is None:
        n_organs = int(rng.integers(8, 13))
    clean = _body_and_organs(shape_mno, n_organs, rng)
    field = bspline_field(shape_mno, spacing, max_u, rng)
    mov = warp_volume(clean, field)
    if contrast == "gamma":
        mov = _gamma_remap(mov, rng)
    elif contrast == "piecewise":
        mov = _piecewise_transfer(mov, rng)
    if rician:
        def noisy(v):
            off = v - AIR
            a = rng.normal(0.0, noise, size=v.shape)
            b = rng.normal(0.0, noise, size=v.shape)
            return AIR + np.sqrt((off + a) ** 2 + b ** 2)
        fixed, moving = noisy(clean), noisy(mov)
    else:
        fixed = clean + rng.normal(0.0, noise, size=clean.shape)
        moving = mov + rng.normal(0.0, noise, size=mov.shape)
    if dtype == "float32":
        return fixed.astype(np.float32), moving.astype(np.float32)
    return _to_int16(fixed), _to_int16(moving)


# --- the benchmark workloads -------------------------------------------------

SCHEDULE_SMALL = dict(grid=(6, 5, 4), radius=(6, 5, 4), quant=(3, 2, 1), alpha=1.6)
SCHEDULE_FULL = dict(grid=(8, 7, 6, 5, 4), radius=(8, 7, 6, 5, 4), quant=(5, 4, 3, 2, 1), alpha=1.6)


def workload_small(seed=0):
    return make_pair((96, 96, 80), seed, noise=20.0, spacing=24, max_u=6.0)


def workload_full(seed=1):
    return make_pair((256, 256, 128), seed, n_organs=20, noise=25.0, spacing=32, max_u=12.0,
                     contrast="gamma")


def workload_crossmodal(seed=2, shape_mno=(192, 160, 192), dtype="int16"):
    return make_pair(shape_mno, seed, n_organs=20, noise=15.0, spacing=32, max_u=10.0,
                     contrast="piecewise", rician=True, dtype=dtype)


def atlas_dims(seed=100, count=16):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(128, 225)), int(rng.integers(128, 193)), int(rng.integers(96, 177)))
            for _ in range(count)]


def workload_atlas(first_seed=100, count=16):
    dims = atlas_dims(first_seed, count)
    return [workload_crossmodal(first_seed + i, d) for i, d in enumerate(dims)]


def workload_soft(seed=4, crop=None):
    """The 160^3 float32 pair of the soft variant; `crop` = edge of a centred cube cut from it."""
    f, m = workload_crossmodal(seed, (160, 160, 160), dtype="float32")
    if crop:
        a = (160 - crop) // 2
        sl = (slice(a, a + crop),) * 3
        f, m = np.ascontiguousarray(f[sl]), np.ascontiguousarray(m[sl])
    return f, m
