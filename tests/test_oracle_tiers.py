"""Pins of the oracle's sub-cell boundary tiers (restated from
native/backends/fdm/src/exchange_t0_fp64.cu, exchange_t1_fp64.cu,
demag_boundary_fp64.cu and the volume-fraction packing of demag_fp64.cu).
The reference holds no known-answer data for them, so each check is a case
the stencils reduce to independently of the tier code, or a closed form:

* T0 with unit volume fractions and face links, and T1 with distances equal
  to the cell size, are the standard six-point exchange, and zero links /
  zero neighbours around a hole are the masked standard exchange;
* T0 scales as f / max(phi, phi_floor) (a dropped or inverted factor, a
  wrong floor fails); T1's coefficient 2 / (d (d + delta')) takes the
  distance of the opposite side (a swapped side fails);
* the phi-weighted demag source equals the standard one with per-cell
  Ms' = phi Ms, and every correction row adds -dN . (phi Ms m) for its own
  target and sources (a transposed tensor entry or a wrong index fails);
* the exchange energy stays that of the standard stencil, as the
  reference's pair-sum reduction ignores the tiers.
"""

import dataclasses
import math

import numpy as np
import pytest

from tests import _tiers

MU0 = 4.0 * math.pi * 1e-7
DX, DY, DZ = 3e-9, 4e-9, 2e-9


def base(O, nx, ny, nz, **kw):
    c = dict(nx=nx, ny=ny, nz=nz, dx=DX, dy=DY, dz=DZ, ms=8e5, a=1.3e-11, alpha=0.1)
    c.update(kw)
    return O.OracleConfig(**c)


def observe(O, cfg, m, raw=False):
    p = O.OracleProblem(cfg, m)
    if raw:
        p.set_m_raw(m)
    return p.observe()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("cells", [(9, 7, 4), (12, 10, 1)])
def test_unit_tiers_are_the_standard_exchange(oracle_mod, cells):
    O = oracle_mod
    nx, ny, nz = cells
    n = nx * ny * nz
    cfg = base(O, nx, ny, nz)
    m = O.random_unit_vectors(11, n)
    ref = observe(O, cfg, m)
    t0 = observe(O, dataclasses.replace(cfg, tier=1, phi=np.ones(n), face_links=np.ones((6, n))), m)
    full = np.array([[DX] * n, [DX] * n, [DY] * n, [DY] * n, [DZ] * n, [DZ] * n])
    t1 = observe(O, dataclasses.replace(cfg, tier=2, phi=np.ones(n), deltas=full), m)
    for o in (t0, t1):
        assert rel(o["h_ex"], ref["h_ex"]) <= 1e-14
        np.testing.assert_array_equal(o["h_demag"], ref["h_demag"])
        assert o["e_ex"] == ref["e_ex"]
    # tier 2 without distances and tier 1 without links keep the standard stencil
    for t, kw in ((1, dict(deltas=full)), (2, dict(face_links=np.ones((6, n))))):
        o = observe(O, dataclasses.replace(cfg, tier=t, phi=np.ones(n), **kw), m)
        np.testing.assert_array_equal(o["h_ex"], ref["h_ex"])


def test_hole_matches_masked_standard_exchange(oracle_mod):
    """Cells of a hole hold m = 0 and phi = 0.  T0 with the links into the
    hole cut, and T1 with full distances (it skips zero neighbours), equal
    the standard exchange with the hole masked out."""
    O = oracle_mod
    nx, ny, nz = 10, 8, 3
    n = nx * ny * nz
    x, y, z = np.arange(n) % nx, (np.arange(n) // nx) % ny, np.arange(n) // (nx * ny)
    hole = (np.abs(x - 4.5) < 2) & (np.abs(y - 3.5) < 2) & (z >= 1)
    m = O.random_unit_vectors(5, n)
    m[hole] = 0.0
    ref = observe(O, base(O, nx, ny, nz, active_mask=(~hole).astype(np.uint8)), m, raw=True)
    phi = (~hole).astype(float)
    faces = np.ones((6, n))
    for k, (ddx, ddy, ddz) in enumerate([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]):
        xn, yn, zn = x + ddx, y + ddy, z + ddz
        ok = (xn >= 0) & (xn < nx) & (yn >= 0) & (yn < ny) & (zn >= 0) & (zn < nz)
        j = np.where(ok, xn + nx * (yn + ny * zn), 0)
        faces[k][ok & hole[j]] = 0.0
    full = np.array([[DX] * n, [DX] * n, [DY] * n, [DY] * n, [DZ] * n, [DZ] * n])
    cfg = base(O, nx, ny, nz)
    t0 = observe(O, dataclasses.replace(cfg, tier=1, phi=phi, face_links=faces), m, raw=True)
    t1 = observe(O, dataclasses.replace(cfg, tier=2, phi=phi, deltas=full), m, raw=True)
    for o in (t0, t1):
        assert rel(o["h_ex"], ref["h_ex"]) <= 1e-14
        assert np.all(o["h_ex"][hole] == 0.0)


def test_t0_scales_with_links_and_volume(oracle_mod):
    O = oracle_mod
    nx, ny, nz = 8, 6, 3
    n = nx * ny * nz
    cfg = base(O, nx, ny, nz)
    m = O.random_unit_vectors(21, n)
    ref = observe(O, cfg, m)["h_ex"]
    one = np.ones((6, n))
    cases = [(0.3 * one, 1.0, 0.0, 0.3), (one, 0.5, 0.0, 2.0), (one, 0.01, 0.0, 1 / 0.05),
             (one, 0.1, 0.2, 1 / 0.2), (0.5 * one, 0.25, 0.0, 2.0)]
    for faces, phi, floor, factor in cases:
        o = observe(O, dataclasses.replace(cfg, tier=1, phi=np.full(n, phi), face_links=faces, phi_floor=floor), m)
        assert rel(o["h_ex"], factor * ref) <= 1e-14, (phi, floor, factor)
    phi = np.ones(n)
    phi[17] = 0.0
    o = observe(O, dataclasses.replace(cfg, tier=1, phi=phi, face_links=one), m)
    assert np.all(o["h_ex"][17] == 0.0)
    keep = np.arange(n) != 17
    assert rel(o["h_ex"][keep], ref[keep]) <= 1e-14


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_t1_uses_the_opposite_sides_distance(oracle_mod, axis):
    """A raw state equal everywhere but at one neighbour of an interior cell:
    the cell's field is the standard one times 2 d / (d + delta_far), where
    delta_far is max(delta, delta_min) on the side away from that neighbour."""
    O = oracle_mod
    nx, ny, nz = 7, 7, 7
    n = nx * ny * nz
    d = (DX, DY, DZ)[axis]
    stride = (1, nx, nx * ny)[axis]
    i = 3 + nx * (3 + ny * 3)
    cfg = base(O, nx, ny, nz, demag=False)
    dmin = 0.1 * min(DX, DY, DZ)
    for side in (-1, 1):  # the neighbour that differs
        m = np.tile([0.0, 0.0, 1.0], (n, 1))
        m[i + side * stride] = [1.0, 0.0, 0.0]
        ref = observe(O, cfg, m, raw=True)["h_ex"][i]
        near, far = (2 * axis + (1 if side < 0 else 0)), (2 * axis + (0 if side < 0 else 1))
        for dn, df in ((d, 0.5 * d), (0.5 * d, d), (d, 0.0), (0.0, d)):
            deltas = np.array([[DX] * n, [DX] * n, [DY] * n, [DY] * n, [DZ] * n, [DZ] * n])
            deltas[near, i] = dn
            deltas[far, i] = df
            o = observe(O, dataclasses.replace(cfg, tier=2, phi=np.ones(n), deltas=deltas), m, raw=True)
            factor = 2 * d / (d + max(df, dmin))
            assert rel(o["h_ex"][i], factor * ref) <= 1e-14, (side, dn, df)


def test_phi_weighted_demag_source(oracle_mod):
    """The tier's source phi Ms m is the standard source with Ms' = phi Ms;
    the mask no longer gates the source but still zeroes the field."""
    O = oracle_mod
    nx, ny, nz = 12, 10, 3
    n = nx * ny * nz
    phi, faces, _ = _tiers.geometry(nx, ny, nz, DX, DY, DZ)
    m = O.random_unit_vectors(8, n)
    mask = (phi > 0.3).astype(np.uint8)
    m[mask == 0] = 0.0  # a masked state is zero outside the mask
    cfg = base(O, nx, ny, nz, exchange=False)
    ref = observe(O, dataclasses.replace(cfg, ms_cell=phi * 8e5), m, raw=True)["h_demag"]
    o = observe(O, dataclasses.replace(cfg, tier=1, phi=phi, face_links=faces, active_mask=mask), m, raw=True)
    act = mask.astype(bool)
    assert rel(o["h_demag"][act], ref[act]) <= 1e-14
    assert np.all(o["h_demag"][~act] == 0.0)


def test_demag_correction_rows(oracle_mod):
    O = oracle_mod
    nx, ny, nz = 8, 6, 2
    n = nx * ny * nz
    rng = np.random.default_rng(4)
    phi = rng.uniform(0.2, 1.0, n)
    m = O.random_unit_vectors(9, n)
    cfg = base(O, nx, ny, nz, exchange=False, tier=1, phi=phi, face_links=np.ones((6, n)))
    h0 = observe(O, cfg, m)["h_demag"]
    i0, i1, i2, i3 = 5, 30, 31, 77
    c1, c2, c3 = 0.07, -0.03, 0.11
    tgt = np.array([i0, i1, -1, i0], dtype=np.int32)
    src = np.array([[i0, -1], [i2, -1], [3, 4], [-1, i3]], dtype=np.int32)
    ten = np.zeros((4, 2, 6))
    ten[0, 0, [0, 3, 5]] = c1  # c1 I at the cell itself
    ten[1, 0, 1] = c2  # Nxy only
    ten[2] = 1.0  # target -1: skipped
    ten[3, 1, 5] = c3  # Nzz only, second row of i0
    o = observe(O, dataclasses.replace(cfg, corr_target=tgt, corr_source=src, corr_tensor=ten), m)["h_demag"]
    M = lambda j: phi[j] * 8e5 * m[j]
    want = np.zeros((n, 3))
    want[i0] = -c1 * M(i0) + np.array([0.0, 0.0, -c3 * M(i3)[2]])
    want[i1] = -c2 * np.array([M(i2)[1], M(i2)[0], 0.0])
    d = o - h0
    np.testing.assert_allclose(d, want, rtol=1e-12, atol=1e-12 * np.abs(h0).max())
    others = np.ones(n, bool)
    others[[i0, i1]] = False
    assert np.all(d[others] == 0.0)


def test_tier_exchange_energy_is_the_standard_one(oracle_mod):
    O = oracle_mod
    nx, ny, nz = 12, 10, 2
    n = nx * ny * nz
    phi, faces, deltas = _tiers.geometry(nx, ny, nz, DX, DY, DZ)
    m = O.random_unit_vectors(31, n)
    cfg = base(O, nx, ny, nz, demag=False)
    ref = observe(O, cfg, m)
    for kw in (dict(tier=1, phi=phi, face_links=faces), dict(tier=2, phi=phi, deltas=deltas)):
        o = observe(O, dataclasses.replace(cfg, **kw), m)
        assert o["e_ex"] == ref["e_ex"]
        assert rel(o["h_ex"], ref["h_ex"]) > 1e-3  # the field itself is the tier's
