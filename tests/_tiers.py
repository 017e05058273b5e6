"""Sub-cell boundary geometry for the tier tests: a cylinder (or a disk in a
film) sampled on the grid the way crates/fullmag-plan/src/boundary_geometry.rs
describes it (R^3 sub-samples per cell for the volume fraction, R^2 per face
for the face links, bisection along each axis for the intersection
distances), plus synthetic sparse demag correction rows on the boundary
shell.  These are inputs only; the planner's exact numbers are not needed.
"""

import numpy as np

R = 4


def cylinder_sdf(nx, ny, dx, dy, radius_frac=0.42):
    cx, cy = 0.5 * nx * dx, 0.5 * ny * dy
    rad = radius_frac * min(nx * dx, ny * dy)

    def sdf(x, y, z):
        return np.hypot(x - cx, y - cy) - rad

    return sdf


def _distance(sdf, x, y, z, vx, vy, vz):
    """Distance from the cell centre to the first zero of sdf along the step
    (vx, vy, vz), or the step length without a sign change."""
    step = np.sqrt(vx * vx + vy * vy + vz * vz)
    f0 = sdf(x, y, z)
    f1 = sdf(x + vx, y + vy, z + vz)
    lo = np.zeros_like(f0)
    hi = np.ones_like(f0)
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        fm = sdf(x + mid * vx, y + mid * vy, z + mid * vz)
        same = fm * f0 > 0
        lo = np.where(same, mid, lo)
        hi = np.where(same, hi, mid)
    return np.where(f0 * f1 > 0, step, 0.5 * (lo + hi) * step)


def geometry(nx, ny, nz, dx, dy, dz, radius_frac=0.42):
    """Returns phi (n,), face links (6, n), distances (6, n) in the order
    xp, xm, yp, ym, zp, zm."""
    sdf = cylinder_sdf(nx, ny, dx, dy, radius_frac)
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = ix.ravel().astype(float), iy.ravel().astype(float), iz.ravel().astype(float)
    t = (np.arange(R) + 0.5) / R
    inside = np.zeros(ix.size)
    for a in t:
        for b in t:
            for c in t:
                inside += sdf((ix + a) * dx, (iy + b) * dy, (iz + c) * dz) <= 0
    phi = inside / R**3

    def face(x0, y0, z0, sx, sy, sz):
        cnt = np.zeros(ix.size)
        for a in t:
            for b in t:
                if sx == 0:
                    p = (x0, y0 + a * sy, z0 + b * sz)
                elif sy == 0:
                    p = (x0 + a * sx, y0, z0 + b * sz)
                else:
                    p = (x0 + a * sx, y0 + b * sy, z0)
                cnt += sdf(*p) <= 0
        return cnt / R**2

    x0, y0, z0 = ix * dx, iy * dy, iz * dz
    faces = np.stack([
        face(x0 + dx, y0, z0, 0, dy, dz), face(x0, y0, z0, 0, dy, dz),
        face(x0, y0 + dy, z0, dx, 0, dz), face(x0, y0, z0, dx, 0, dz),
        face(x0, y0, z0 + dz, dx, dy, 0), face(x0, y0, z0, dx, dy, 0),
    ])
    cx, cy, cz = x0 + 0.5 * dx, y0 + 0.5 * dy, z0 + 0.5 * dz
    deltas = np.stack([
        _distance(sdf, cx, cy, cz, dx, 0, 0), _distance(sdf, cx, cy, cz, -dx, 0, 0),
        _distance(sdf, cx, cy, cz, 0, dy, 0), _distance(sdf, cx, cy, cz, 0, -dy, 0),
        _distance(sdf, cx, cy, cz, 0, 0, dz), _distance(sdf, cx, cy, cz, 0, 0, -dz),
    ])
    return phi, faces, deltas


def correction_rows(nx, ny, nz, phi, seed=3, scale=0.02):
    """Rows on the boundary shell (0 < phi < 1): the cell itself and its six
    neighbours as sources (-1 beyond the grid or where phi = 0), random
    symmetric tensors; a few cells get a second row and one row has target
    -1, as the kernel must skip it."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    shell = np.nonzero((phi > 0) & (phi < 1))[0]
    targets = list(shell) + list(shell[::5]) + [-1]
    st = 7
    src = np.full((len(targets), st), -1, dtype=np.int32)
    for r, t in enumerate(targets):
        if t < 0:
            src[r] = rng.integers(0, n, st)
            continue
        x, y, z = t % nx, (t // nx) % ny, t // (nx * ny)
        cand = [(x, y, z), (x + 1, y, z), (x - 1, y, z), (x, y + 1, z), (x, y - 1, z), (x, y, z + 1),
                (x, y, z - 1)]
        for s, (a, b, c) in enumerate(cand):
            if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz:
                j = a + nx * (b + ny * c)
                if phi[j] > 0:
                    src[r, s] = j
    ten = scale * rng.standard_normal((len(targets), st, 6))
    return np.array(targets, dtype=np.int32), src, ten
