"""The planner's boundary geometry (fullmag_amd/boundary.py, restating
crates/fullmag-plan/src/boundary_geometry.rs) on cases with closed forms:
a half space cutting one column of cells at a known plane (volume and face
fractions, intersection distances to the bisection tolerance) and one
partial cell in a full block (its correction row is -0.5 N(offset) for every
neighbour within reach, in z-major order).  The Newell kernels come from the
oracle here; on the GPU the library's own are checked against them."""

import numpy as np
import pytest

from fullmag_amd import boundary as B

DX, DY, DZ = 2e-9, 3e-9, 1.5e-9


def oracle_newell(O):
    return lambda n, dx, dy, dz: O.newell_kernels(n, n, n, dx, dy, dz)


def test_half_space_fractions_and_distances(oracle_mod):
    nx, ny, nz = 7, 3, 2
    xp = 3.3 * DX  # material for x <= xp
    g = B.compute_boundary_geometry(lambda x, y, z: x - xp, nx, ny, nz, DX, DY, DZ, True,
                                    newell=oracle_newell(oracle_mod))
    x = np.arange(nx * ny * nz) % nx
    want_vf = np.where(x < 3, 1.0, np.where(x == 3, 0.25, 0.0))
    np.testing.assert_array_equal(g.volume_fraction, want_vf)
    np.testing.assert_array_equal(g.face_link_xm, np.where(x <= 3, 1.0, 0.0))
    np.testing.assert_array_equal(g.face_link_xp, np.where(x < 3, 1.0, 0.0))
    np.testing.assert_array_equal(g.face_link_yp, want_vf)
    np.testing.assert_array_equal(g.face_link_zm, want_vf)
    tol = DX * 2.0**-19
    np.testing.assert_allclose(g.delta_xp[x == 2], 0.8 * DX, atol=tol)   # centre 2.5 dx to 3.3 dx
    np.testing.assert_allclose(g.delta_xm[x == 3], 0.2 * DX, atol=tol)   # centre 3.5 dx back to 3.3 dx
    np.testing.assert_array_equal(g.delta_xp[x == 0], DX)                # no crossing within one step
    np.testing.assert_array_equal(g.delta_yp, DY)
    np.testing.assert_array_equal(g.delta_zm, DZ)


def test_single_partial_cell_correction_row(oracle_mod):
    O = oracle_mod
    n = 5
    phi = np.ones(n**3)
    c = 2 + n * (2 + n * 2)
    phi[c] = 0.5
    phi[0] = 0.0005  # an empty source is skipped
    tgt, src, ten, st = B.compute_sparse_demag_correction(phi, n, n, n, DX, DY, DZ, newell=oracle_newell(O))
    np.testing.assert_array_equal(tgt, [c])
    assert st == n**3 - 2
    nk = O.newell_kernels(4, 4, 4, DX, DY, DZ)
    src = src.reshape(1, st)
    ten = ten.reshape(1, st, 6)
    k = 0
    for dk in range(-2, 3):
        for dj in range(-2, 3):
            for di in range(-2, 3):
                j = (2 + di) + n * ((2 + dj) + n * (2 + dk))
                if j == c or j == 0:
                    continue
                assert src[0, k] == j
                want = -0.5 * nk[:, dk % 8, dj % 8, di % 8]  # (xx, yy, zz, xy, xz, yz)
                np.testing.assert_array_equal(ten[0, k], want)
                k += 1
    assert k == st


def test_no_boundary_cells_give_no_rows(oracle_mod):
    tgt, src, ten, st = B.compute_sparse_demag_correction(np.ones(27), 3, 3, 3, DX, DY, DZ,
                                                          newell=oracle_newell(oracle_mod))
    assert tgt.size == 0 and src.size == 0 and ten.size == 0 and st == 0


def test_cylinder_volume(oracle_mod):
    nx = ny = 40
    r = 15.3 * DX
    g = B.compute_boundary_geometry(B.cylinder_sdf(r, 20 * DX, 20 * DY), nx, ny, 1, DX, DY, DZ, False,
                                    newell=oracle_newell(oracle_mod))
    area = g.volume_fraction.sum() * DX * DY
    assert abs(area - np.pi * r * r) <= 0.01 * np.pi * r * r
    assert np.all((g.volume_fraction >= 0) & (g.volume_fraction <= 1))
    rows = g.demag_corr_target_idx.size
    assert rows > 0 and g.demag_corr_source_idx.size == rows * g.demag_corr_stencil_size
    hole = B.cylinder_hole_sdf(r, 5 * DX, 20 * DX, 20 * DY)
    gh = B.compute_boundary_geometry(hole, nx, ny, 1, DX, DY, DZ, False, newell=oracle_newell(oracle_mod))
    assert gh.volume_fraction.sum() < g.volume_fraction.sum()


@pytest.mark.gpu
def test_device_newell_kernels_match_oracle(fa, oracle_mod):
    got = B.device_newell_kernels(4, DX, DY, DZ)
    want = oracle_mod.newell_kernels(4, 4, 4, DX, DY, DZ)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
