"""Sub-cell boundary geometry for the boundary tiers, as fullmag's planner
computes it from a signed distance function
(crates/fullmag-plan/src/boundary_geometry.rs): per-cell volume fractions,
face-link fractions and intersection distances, and the sparse demag
correction rows.  Host-side numpy; the SDF takes numpy arrays of x, y, z
[m] and is negative inside the material.

The arrays feed fullmag_fdm_plan_desc unchanged (FdmPlan.boundary_geometry).
Note that the planner writes each correction tensor as (xx, yy, zz, xy, xz,
yz) while the device kernel reads (xx, xy, xz, yy, yz, zz)
(native/backends/fdm/src/demag_boundary_fp64.cu:52-58); both are kept as the
reference has them, so a plan gives the reference's result.
"""

from __future__ import annotations

from typing import Callable, Optional

import numpy as np

from .native_fdm import BoundaryGeometry, compute_newell_kernel_spectra

SUB_SAMPLE_R = 4  # boundary_geometry.rs:14
CORR_RADIUS = 3   # boundary_geometry.rs:302
PHI_EMPTY = 0.001  # boundary_geometry.rs:331-332
PHI_FULL = 0.999


def cylinder_sdf(radius: float, cx: float, cy: float) -> Callable:
    """boundary_geometry.rs:271-281: a z-infinite cylinder."""
    return lambda x, y, z: np.sqrt((x - cx) ** 2 + (y - cy) ** 2) - radius


def cylinder_hole_sdf(base_radius: float, hole_radius: float, cx: float, cy: float) -> Callable:
    """boundary_geometry.rs:283-298: the outer cylinder minus the inner."""
    outer, inner = cylinder_sdf(base_radius, cx, cy), cylinder_sdf(hole_radius, cx, cy)
    return lambda x, y, z: np.maximum(outer(x, y, z), -inner(x, y, z))


def _face_fraction(sdf, x0, y0, z0, s1, s2, plane: str) -> np.ndarray:
    """boundary_geometry.rs:196-233: R x R samples of one face; `plane` names
    the two axes the face spans ("yz", "xz" or "xy")."""
    r = SUB_SAMPLE_R
    inside = np.zeros(np.shape(x0))
    for a in range(r):
        t1 = (a + 0.5) / r
        for b in range(r):
            t2 = (b + 0.5) / r
            if plane == "yz":
                p = (x0, y0 + t1 * s1, z0 + t2 * s2)
            elif plane == "xz":
                p = (x0 + t1 * s1, y0, z0 + t2 * s2)
            else:
                p = (x0 + t1 * s1, y0 + t2 * s2, z0)
            inside += sdf(*p) <= 0.0
    return inside / (r * r)


def _boundary_distance(sdf, x0, y0, z0, vx, vy, vz) -> np.ndarray:
    """boundary_geometry.rs:239-268: distance to the zero of the SDF along the
    step by 20 bisections, or the step length without a sign change."""
    step = np.sqrt(vx * vx + vy * vy + vz * vz)
    if step == 0.0:
        return np.zeros(np.shape(x0))
    f0 = sdf(x0, y0, z0)
    f1 = sdf(x0 + vx, y0 + vy, z0 + vz)
    lo = np.zeros(np.shape(x0))
    hi = np.ones(np.shape(x0))
    for _ in range(20):
        mid = 0.5 * (lo + hi)
        fm = sdf(x0 + mid * vx, y0 + mid * vy, z0 + mid * vz)
        left = f0 * fm <= 0.0
        hi = np.where(left, mid, hi)
        lo = np.where(left, lo, mid)
    return np.where(f0 * f1 > 0.0, step, 0.5 * (lo + hi) * step)


def device_newell_kernels(n: int, dx: float, dy: float, dz: float) -> np.ndarray:
    """Real-space padded Newell kernels of an n x n x n grid, shape
    (6, 2n, 2n, 2n) in the order xx, yy, zz, xy, xz, yz, recovered from the
    library's device-built spectra by an inverse FFT."""
    spec = compute_newell_kernel_spectra(n, n, n, dx, dy, dz, thin_film_2d=False)
    c = spec[:, 0::2] + 1j * spec[:, 1::2]
    return np.fft.ifftn(c.reshape(6, 2 * n, 2 * n, 2 * n), axes=(1, 2, 3)).real


def compute_sparse_demag_correction(volume_fraction: np.ndarray, nx: int, ny: int, nz: int, dx: float, dy: float,
                                    dz: float, newell: Optional[Callable] = None):
    """boundary_geometry.rs:320-455.  Rows for the boundary cells
    (0.001 < phi < 0.999) in cell order; per row the sources within +-3
    cells (z outermost, x innermost, the cell itself excluded) with
    phi_s >= 0.001 and |phi_t phi_s - 1| >= 1e-12, each with
    dN = (phi_t phi_s - 1) N(offset) as (xx, yy, zz, xy, xz, yz); rows padded
    to the longest with source -1 and zero tensors.  `newell(n, dx, dy, dz)`
    returns the padded real-space kernels of an n^3 grid (default: the
    library's, which needs the GPU).  Returns (target, source, tensor,
    stencil) or empty arrays and 0."""
    phi = np.asarray(volume_fraction, dtype=np.float64).reshape(-1)
    empty = (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), 0)
    targets = np.nonzero((phi > PHI_EMPTY) & (phi < PHI_FULL))[0]
    if targets.size == 0:
        return empty
    kr = CORR_RADIUS + 1
    nk = (newell or device_newell_kernels)(kr, dx, dy, dz)
    kp = 2 * kr
    tx, ty, tz = targets % nx, (targets // nx) % ny, targets // (nx * ny)
    offsets = [(di, dj, dk) for dk in range(-CORR_RADIUS, CORR_RADIUS + 1)
               for dj in range(-CORR_RADIUS, CORR_RADIUS + 1) for di in range(-CORR_RADIUS, CORR_RADIUS + 1)
               if (di, dj, dk) != (0, 0, 0)]
    no = len(offsets)
    valid = np.zeros((targets.size, no), bool)
    src = np.zeros((targets.size, no), np.int64)
    fac = np.zeros((targets.size, no))
    lut = np.zeros((no, 6))
    for o, (di, dj, dk) in enumerate(offsets):
        sx, sy, sz = tx + di, ty + dj, tz + dk
        inb = (sx >= 0) & (sx < nx) & (sy >= 0) & (sy < ny) & (sz >= 0) & (sz < nz)
        s = np.where(inb, sx + nx * (sy + ny * sz), 0)
        ps = np.where(inb, phi[s], 0.0)
        cf = phi[targets] * ps - 1.0
        valid[:, o] = inb & (ps >= PHI_EMPTY) & (np.abs(cf) >= 1e-12)
        src[:, o] = s
        fac[:, o] = cf
        ix = di if di >= 0 else kp + di
        iy = dj if dj >= 0 else kp + dj
        iz = dk if dk >= 0 else kp + dk
        lut[o] = nk[:, iz, iy, ix]
    keep = valid.any(axis=1)
    if not keep.any():
        return empty
    valid, src, fac, targets = valid[keep], src[keep], fac[keep], targets[keep]
    counts = valid.sum(axis=1)
    st = int(counts.max())
    rows = targets.size
    pos = np.cumsum(valid, axis=1) - 1
    r_idx = np.nonzero(valid)
    source = np.full((rows, st), -1, np.int32)
    tensor = np.zeros((rows, st, 6))
    source[r_idx[0], pos[r_idx]] = src[r_idx]
    tensor[r_idx[0], pos[r_idx]] = fac[r_idx][:, None] * lut[r_idx[1]]
    return targets.astype(np.int32), source.reshape(-1), tensor.reshape(-1), st


def compute_boundary_geometry(sdf: Callable, nx: int, ny: int, nz: int, dx: float, dy: float, dz: float,
                              compute_delta: bool, newell: Optional[Callable] = None) -> BoundaryGeometry:
    """boundary_geometry.rs:23-192 for an SDF over the grid (cell i spans
    [i d, (i + 1) d) on each axis)."""
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = (v.ravel().astype(np.float64) for v in (ix, iy, iz))
    r = SUB_SAMPLE_R
    inside = np.zeros(ix.size)
    for sz in range(r):
        for sy in range(r):
            for sx in range(r):
                inside += sdf(ix * dx + (sx + 0.5) / r * dx, iy * dy + (sy + 0.5) / r * dy,
                              iz * dz + (sz + 0.5) / r * dz) <= 0.0
    vf = inside / r**3
    x0, y0, z0 = ix * dx, iy * dy, iz * dz
    g = BoundaryGeometry(
        volume_fraction=vf,
        face_link_xp=_face_fraction(sdf, x0 + dx, y0, z0, dy, dz, "yz"),
        face_link_xm=_face_fraction(sdf, x0, y0, z0, dy, dz, "yz"),
        face_link_yp=_face_fraction(sdf, x0, y0 + dy, z0, dx, dz, "xz"),
        face_link_ym=_face_fraction(sdf, x0, y0, z0, dx, dz, "xz"),
        face_link_zp=_face_fraction(sdf, x0, y0, z0 + dz, dx, dy, "xy"),
        face_link_zm=_face_fraction(sdf, x0, y0, z0, dx, dy, "xy"))
    n = ix.size
    if compute_delta:
        cx, cy, cz = x0 + 0.5 * dx, y0 + 0.5 * dy, z0 + 0.5 * dz
        g.delta_xp = _boundary_distance(sdf, cx, cy, cz, dx, 0.0, 0.0)
        g.delta_xm = _boundary_distance(sdf, cx, cy, cz, -dx, 0.0, 0.0)
        g.delta_yp = _boundary_distance(sdf, cx, cy, cz, 0.0, dy, 0.0)
        g.delta_ym = _boundary_distance(sdf, cx, cy, cz, 0.0, -dy, 0.0)
        g.delta_zp = _boundary_distance(sdf, cx, cy, cz, 0.0, 0.0, dz)
        g.delta_zm = _boundary_distance(sdf, cx, cy, cz, 0.0, 0.0, -dz)
    else:
        g.delta_xp = g.delta_xm = np.full(n, dx)
        g.delta_yp = g.delta_ym = np.full(n, dy)
        g.delta_zp = g.delta_zm = np.full(n, dz)
    tgt, src, ten, st = compute_sparse_demag_correction(vf, nx, ny, nz, dx, dy, dz, newell)
    g.demag_corr_target_idx, g.demag_corr_source_idx, g.demag_corr_tensor = tgt, src, ten
    g.demag_corr_stencil_size = st
    return g
