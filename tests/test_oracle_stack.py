"""Pins of the oracle's multilayer stack (restated from
crates/fullmag-engine/src/multilayer.rs, crates/fullmag-fdm-demag/src/
{shifted_kernel.rs, transfer.rs, multiply.rs}, newell.rs:528-651 and the
runner's Heun in crates/fullmag-runner/src/multilayer_reference.rs).

* The shifted kernel at zero shift is the self kernel (the reference's own
  test, shifted_kernel.rs:60-77), and for thin layers a shift of s cells is
  the single-grid kernel of a taller grid at displacement s (so a wrong sign
  or a missing offset fails); its mirrored z slots keep slot k's value and
  sign, as the reference places them.
* A stack of one layer is the single-grid problem; thin layers stacked with
  or without a gap are one grid with the gap masked, per-cell Ms per plane.
* push_m averages by overlap volume and pull_h interpolates trilinearly
  (exact for a linear field inside the grid); a stack whose layer needs a
  transfer equals push, single-grid demag on the common grid, pull.
"""

import math

import numpy as np
import pytest

MU0 = 4.0 * math.pi * 1e-7
DX, DY, DZ = 3e-9, 2.5e-9, 1e-9


def problem(O, cells, cell=(DX, DY, DZ), ms=8e5, demag=False, m=None, seed=1, **kw):
    nx, ny, nz = cells
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], ms=ms, a=1.3e-11, alpha=0.3,
                         demag=demag, **kw)
    m0 = O.random_unit_vectors(seed, nx * ny * nz) if m is None else m
    return O.OracleProblem(cfg, m0)


def test_zero_shift_thin_kernel_is_the_self_kernel(oracle_mod):
    O = oracle_mod
    np.testing.assert_array_equal(O.newell_kernels_shifted(4, 4, 1, 2e-9, 2e-9, 1e-9, 0.0),
                                  O.newell_kernels(4, 4, 1, 2e-9, 2e-9, 1e-9))


@pytest.mark.parametrize("s", [1, 2, 3, -1, -3])
def test_thin_shift_is_the_taller_grids_kernel(oracle_mod, s):
    O = oracle_mod
    nx, ny = 6, 5
    k = O.newell_kernels_shifted(nx, ny, 1, DX, DY, DZ, s * DZ)  # (6, 2, 2ny, 2nx)
    tall = O.newell_kernels(nx, ny, abs(s) + 1, DX, DY, DZ)       # (6, 2(|s|+1), ...)
    slot = s if s > 0 else 2 * (abs(s) + 1) + s
    # equal to rounding of the table coordinates ((kk - 1) dz + s dz against (kk + s - 1) dz)
    for c in range(6):
        np.testing.assert_allclose(k[c, 0], tall[c, slot], rtol=0, atol=1e-12 * np.abs(tall[c]).max())
    assert np.all(k[:, 1] == 0.0)
    assert np.abs(k[4, 0]).max() > 0 and np.abs(k[5, 0]).max() > 0  # xz, yz couple shifted layers


def test_shifted_kernel_mirror_slots_keep_slot_k(oracle_mod):
    O = oracle_mod
    nz = 3
    k = O.newell_kernels_shifted(4, 3, nz, DX, DY, DZ, 2.5 * DZ)
    for kk in range(1, nz):
        np.testing.assert_array_equal(k[:, 2 * nz - kk], k[:, kk])
    assert np.all(k[:, nz] == 0.0)


def test_single_layer_stack_is_the_single_grid_problem(oracle_mod):
    O = oracle_mod
    cells = (10, 8, 2)
    ref = problem(O, cells, demag=True, h_ext=(1e4, 0, 2e3))
    lay = problem(O, cells, demag=False, h_ext=(1e4, 0, 2e3))
    st = O.OracleStack([lay], [(0.0, 0.0, 0.0)], cells, (DX, DY, DZ))
    o, r = st.observe(), ref.observe()
    for key in ("h_ex", "h_demag", "h_ext", "h_eff", "rhs"):
        np.testing.assert_allclose(o[key][0], r[key], rtol=0, atol=1e-15 * np.abs(r["h_eff"]).max())
    for key in ("e_ex", "e_demag", "e_ext", "e_total", "max_h_eff", "max_h_demag", "max_dm_dt"):
        assert abs(o[key] - r[key]) <= 1e-13 * max(abs(r[key]), abs(r["e_total"])), key
    for _ in range(5):
        st.step(1e-13)
        ref.step(1e-13)
    np.testing.assert_allclose(lay.m, ref.m, rtol=0, atol=1e-14)


@pytest.mark.parametrize("gap", [0, 2])
def test_thin_layers_are_one_grid_with_the_gap_masked(oracle_mod, gap):
    """Layers at z = 0 and (1 + gap) dz of one 9x7 film each, Ms 8e5 and
    1.1e6: the stack's demag equals a (2 + gap)-plane grid's at its first and
    last plane (the middle planes empty), per-cell Ms per plane."""
    O = oracle_mod
    nx, ny = 9, 7
    n = nx * ny
    nz = 2 + gap
    m = O.random_unit_vectors(3, 2 * n)
    la = problem(O, (nx, ny, 1), ms=8e5, m=m[:n], exchange=False)
    lb = problem(O, (nx, ny, 1), ms=1.1e6, m=m[n:], exchange=False)
    st = O.OracleStack([la, lb], [(0, 0, 0), (0, 0, (1 + gap) * DZ)], (nx, ny, 1), (DX, DY, DZ))
    o = st.observe()
    mask = np.zeros(n * nz, np.uint8)
    mask[:n] = 1
    mask[-n:] = 1
    full = np.zeros((n * nz, 3))
    full[:n] = m[:n]
    full[-n:] = m[n:]
    ms_cell = np.full(n * nz, 8e5)
    ms_cell[-n:] = 1.1e6
    grid = problem(O, (nx, ny, nz), demag=True, exchange=False, active_mask=mask, ms_cell=ms_cell, m=full)
    hd = grid.observe()["h_demag"]
    scale = np.abs(hd).max()
    np.testing.assert_allclose(o["h_demag"][0], hd[:n], rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(o["h_demag"][1], hd[-n:], rtol=0, atol=1e-12 * scale)


def test_transfer_operators(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(2)
    # equal grids: copies
    m = rng.standard_normal((24, 3))
    np.testing.assert_array_equal(O.push_m(m, (4, 3, 2), (1, 1, 1), (4, 3, 2), (2, 2, 2)), m)
    np.testing.assert_array_equal(O.pull_h(m, (4, 3, 2), (1, 1, 1), (4, 3, 2), (2, 2, 2)), m)
    # coarsening 4x4x1 -> 2x2x1: each coarse cell is the mean of its four
    m = rng.standard_normal((16, 3))
    c = O.push_m(m, (4, 4, 1), (1e-9, 1e-9, 1e-9), (2, 2, 1), (2e-9, 2e-9, 1e-9))
    g = m.reshape(4, 4, 3)
    want = np.stack([g[2 * j:2 * j + 2, 2 * i:2 * i + 2].mean(axis=(0, 1)) for j in range(2) for i in range(2)])
    np.testing.assert_allclose(c, want, rtol=1e-14, atol=1e-15)
    # refinement 2x2x1 -> 4x4x1: piecewise constant
    c = O.push_m(want, (2, 2, 1), (2e-9, 2e-9, 1e-9), (4, 4, 1), (1e-9, 1e-9, 1e-9))
    np.testing.assert_allclose(c.reshape(4, 4, 3), np.repeat(np.repeat(want.reshape(2, 2, 3), 2, 0), 2, 1),
                               rtol=1e-14)
    # trilinear pull of a linear field is exact at native centres inside the
    # convolution grid's centre hull
    cc, cd = (5, 4, 3), (2.0, 3.0, 1.5)
    zc, yc, xc = np.meshgrid(*[(np.arange(k) + 0.5) * d for k, d in zip(cc[::-1], cd[::-1])], indexing="ij")
    lin = lambda x, y, z: np.stack([1 + 2 * x - y, 3 * z + 0.5 * x, x + y + z], -1)
    h = lin(xc.ravel(), yc.ravel(), zc.ravel())
    nc, nd = (10, 6, 3), (1.0, 2.0, 1.5)
    out = O.pull_h(h, cc, cd, nc, nd)
    zn, yn, xn = np.meshgrid(*[(np.arange(k) + 0.5) * d for k, d in zip(nc[::-1], nd[::-1])], indexing="ij")
    inside = ((xn >= cd[0] / 2) & (xn <= (cc[0] - 0.5) * cd[0]) & (yn >= cd[1] / 2) & (yn <= (cc[1] - 0.5) * cd[1])
              & (zn >= cd[2] / 2) & (zn <= (cc[2] - 0.5) * cd[2])).ravel()
    np.testing.assert_allclose(out[inside], lin(xn.ravel(), yn.ravel(), zn.ravel())[inside], rtol=1e-13,
                               atol=1e-12)


def test_stack_transfer_is_push_demag_pull(oracle_mod):
    """A layer on a 2x finer native grid than the common grid: its field is
    pull(single-grid demag on the common grid of push(m))."""
    O = oracle_mod
    cc, cd = (6, 5, 1), (2 * DX, 2 * DY, DZ)
    nc, nd = (12, 10, 1), (DX, DY, DZ)
    m = O.random_unit_vectors(7, 120)
    lay = problem(O, nc, cell=nd, m=m, exchange=False)
    st = O.OracleStack([lay], [(0, 0, 0)], cc, cd)
    got = st.observe()["h_demag"][0]
    pm = O.push_m(m, nc, nd, cc, cd)
    coarse = problem(O, cc, cell=cd, demag=True, exchange=False)
    coarse.set_m_raw(pm)
    want = O.pull_h(coarse.observe()["h_demag"], cc, cd, nc, nd)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14 * np.abs(want).max())
