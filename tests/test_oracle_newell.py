"""Oracle Newell tensor vs the reference's own known-answer tests.

Restates crates/fullmag-fdm-demag/src/newell.rs:676-907 plus closed forms.
"""

import math

import numpy as np
import pytest


def test_f_and_g_vanish_at_origin(oracle_mod):
    # newell.rs:676-684
    assert oracle_mod.newell_f(0.0, 0.0, 0.0) == 0.0
    assert oracle_mod.newell_g(0.0, 0.0, 0.0) == 0.0


@pytest.mark.parametrize("cell", [(1.0, 1.0, 1.0), (5e-9, 5e-9, 2e-9), (5e-9, 5e-9, 1e-9), (3e-9, 4e-9, 5e-9)])
def test_self_term_trace_is_one(oracle_mod, cell, kats):
    # newell.rs:686-717, 870-905: trace(N_self) = 1
    k = oracle_mod.newell_kernels(1, 1, 1, *cell)
    assert abs(k[0, 0, 0, 0] + k[1, 0, 0, 0] + k[2, 0, 0, 0] - 1.0) < kats["newell"]["trace_tol"]


def test_cubic_self_term_is_one_third(oracle_mod, kats):
    # newell.rs:700-707; closed form N = I/3 for a cube
    k = oracle_mod.newell_kernels(1, 1, 1, 1.0, 1.0, 1.0)
    v, tol = kats["newell"]["cubic_nxx"]
    for c in range(3):
        assert abs(k[c, 0, 0, 0] - v) < tol


def test_flat_cell_ordering_and_xx_eq_yy(oracle_mod):
    # newell.rs:719-737 and 883-905
    k = oracle_mod.newell_kernels(1, 1, 1, 5e-9, 5e-9, 2e-9)
    assert k[2, 0, 0, 0] > k[0, 0, 0, 0]
    assert abs(k[0, 0, 0, 0] - k[1, 0, 0, 0]) < 1e-12
    k = oracle_mod.newell_kernels(1, 1, 1, 5e-9, 5e-9, 1e-9)
    assert k[2, 0, 0, 0] > 0.5


def test_off_diagonal_self_term_is_zero(oracle_mod):
    # newell.rs:739-758
    k = oracle_mod.newell_kernels(1, 1, 1, 3e-9, 4e-9, 5e-9)
    for c in (3, 4, 5):
        assert abs(k[c, 0, 0, 0]) < 1e-12


def test_kernel_symmetries(oracle_mod):
    # newell.rs:760-780: N_xx even in x, N_xy odd in x
    k = oracle_mod.newell_kernels(4, 4, 4, 1.0, 1.0, 1.0)
    px = 8
    assert abs(k[0, 3, 2, 1] - k[0, 3, 2, px - 1]) < 1e-15
    assert abs(k[3, 3, 2, 1] + k[3, 3, 2, px - 1]) < 1e-15
    # also z parity of xz and yz (placement signs, newell.rs:397-426)
    assert abs(k[4, 1, 2, 1] + k[4, 8 - 1, 2, 1]) < 1e-15
    assert abs(k[5, 1, 2, 1] + k[5, 8 - 1, 2, 1]) < 1e-15
    assert abs(k[2, 1, 2, 1] - k[2, 8 - 1, 2, 1]) < 1e-15


def test_nxy_absolute_values(oracle_mod, kats):
    # newell.rs:782-841
    spec = kats["newell"]["nxy_10x10x1_unit"]
    k = oracle_mod.newell_kernels(*spec["grid"], *spec["cell"])
    for i, j, kk, v, tol in spec["points"]:
        assert abs(k[3, kk, j, i] - v) < tol
    assert abs(k[3, 0, 0, 1]) < 1e-15
    for i in range(1, 5):
        for j in range(1, 5):
            assert k[3, 0, j, i] < 0


def test_nxy_converges_to_dipole(oracle_mod, kats):
    # newell.rs:843-868
    k = oracle_mod.newell_kernels(12, 12, 1, 1.0, 1.0, 1.0)
    x = y = 8.0
    n_asym = -3.0 * x * y / (4.0 * math.pi * (x * x + y * y) ** 2.5)
    assert abs(k[3, 0, 8, 8] - n_asym) / abs(n_asym) < kats["newell"]["asymptotic_8_8_rel"]


def test_far_field_switch_uses_point_dipole(oracle_mod):
    # newell.rs:371-375: displacement >= 40 cells uses the point-dipole formula
    dx, dy, dz = 2e-9, 3e-9, 1e-9
    k = oracle_mod.newell_kernels(42, 2, 1, dx, dy, dz)
    i, j = 41, 1
    x, y, z = i * dx, j * dy, 0.0
    r2 = x * x + y * y + z * z
    r = math.sqrt(r2)
    vol = dx * dy * dz
    nxx = (1.0 / (r2 * r) - 3.0 * x * x / (r2 * r * r2)) / (4.0 * math.pi) * vol
    nxy = -3.0 * x * y / (4.0 * math.pi * r2 * r2 * r) * vol
    assert k[0, 0, j, i] == pytest.approx(nxx, rel=1e-14)
    assert k[3, 0, j, i] == pytest.approx(nxy, rel=1e-14)
    # just inside the threshold the near-field stencil agrees with the
    # dipole formula to well under a percent
    i2 = 39
    x2 = i2 * dx
    r2b = x2 * x2 + y * y
    rb = math.sqrt(r2b)
    nxx_dip = (1.0 / (r2b * rb) - 3.0 * x2 * x2 / (r2b * rb * r2b)) / (4.0 * math.pi) * vol
    assert k[0, 0, j, i2] == pytest.approx(nxx_dip, rel=1e-2)


def test_spectra_are_real_up_to_cancellation_noise(oracle_mod):
    # The padded real-space kernels are even/odd in each axis, so their
    # spectra are real; the imaginary parts are cancellation noise only.
    for (nx, ny, nz, tf) in [(8, 6, 1, True), (6, 4, 3, False)]:
        s = oracle_mod.newell_spectra(nx, ny, nz, 5e-9, 5e-9, 3e-9, tf)
        re, im = s[:, 0::2], s[:, 1::2]
        assert np.abs(im).max() < 1e-10 * np.abs(re).max()


def test_thin_film_2d_spectra_equal_3d_plane(oracle_mod):
    # lib.rs:553-610 vs 511-551: for nz=1 the 2-D spectrum equals the kz=0
    # plane of the 3-D spectrum with pz=2 (the kz=1 plane is identical).
    nx, ny = 6, 5
    s2 = oracle_mod.newell_spectra(nx, ny, 1, 4e-9, 4e-9, 2e-9, True).reshape(6, 2 * ny, 2 * nx, 2)
    s3 = oracle_mod.newell_spectra(nx, ny, 1, 4e-9, 4e-9, 2e-9, False).reshape(6, 2, 2 * ny, 2 * nx, 2)
    np.testing.assert_allclose(s2, s3[:, 0], atol=1e-14)
    np.testing.assert_allclose(s3[:, 1], s3[:, 0], atol=1e-14)
