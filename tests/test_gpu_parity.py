"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Tolerances: FP64 paths are held to the reference's own Tier-A bar
(native/backends/fdm/tests/tier_a_compare.cu:261-262, m 1e-12 for exchange
only) or 1e-10 in m with demag; single precision to the
reference's Tier-B bar (tier_b_compare.cu:113-114, m 1e-5, E 1e-4 relative);
mixed precision to max|dm| <= 1e-5 against the FP64 GPU path.
"""

import json
import math
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

MU0 = 4.0 * math.pi * 1e-7


def make_pair(fa, O, *, cells, cell, ms, a, alpha, m0, h_ext=None, exchange=True, demag=True,
              integrator="heun", precision="double", mask=None, kernel="spectra", disable_precession=False,
              adaptive=None, ms_cell=None, region=None, pairs=()):
    nx, ny, nz = cells
    spectra = None
    if demag and kernel == "spectra":
        spectra = O.newell_spectra(nx, ny, nz, *cell, thin_film_2d=(nz == 1))
    adaptive = adaptive or {}
    plan = fa.FdmPlan(cells=cells, cell_size=cell, initial_magnetization=m0, saturation_magnetisation=ms,
                      exchange_stiffness=a, damping=alpha, precision=precision, integrator=integrator,
                      enable_exchange=exchange, enable_demag=demag, external_field=h_ext, active_mask=mask,
                      demag_kernel=spectra if spectra is not None else kernel,
                      disable_precession=disable_precession, ms_cell=ms_cell, region_mask=region,
                      inter_region_exchange=list(pairs), **adaptive)
    gpu = fa.NativeFdmBackend(plan)
    lut = None
    if region is not None:
        lut = np.zeros((256, 256))
        lut[np.arange(256), np.arange(256)] = a
        for ri, rj, aij in pairs:
            lut[ri, rj] = lut[rj, ri] = aij
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], ms=ms, a=a, alpha=alpha,
                         integrator=integrator, exchange=exchange, demag=demag, h_ext=h_ext,
                         active_mask=mask, precession=not disable_precession, ms_cell=ms_cell,
                         max_error=adaptive.get("adaptive_max_error", 0.0),
                         dt_min=adaptive.get("adaptive_dt_min", 0.0), dt_max=adaptive.get("adaptive_dt_max", 0.0),
                         headroom=adaptive.get("adaptive_headroom", 0.0), region_mask=region, exchange_lut=lut)
    cpu = O.OracleProblem(cfg, m0)
    return gpu, cpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_config1_thin_film_heun_fp64_1000_steps(fa, oracle_mod):
    """64x64x1 of 5x5x3 nm, Ms 8e5, A 1.3e-11, alpha 0.5, B = 10 mT x,
    exchange + Newell demag (2-D padded 128x128) + Zeeman, dt 1e-13, 1000
    steps, m0 = planner generator seed 42: max|dm| <= 1e-10."""
    m0 = fa.generate_random_unit_vectors(42, 64 * 64)
    gpu, cpu = make_pair(fa, oracle_mod, cells=(64, 64, 1), cell=(5e-9, 5e-9, 3e-9), ms=8e5, a=1.3e-11,
                         alpha=0.5, m0=m0, h_ext=(10e-3 / MU0, 0.0, 0.0))
    worst = 0.0
    for s in range(1000):
        g = gpu.step(1e-13)
        r = cpu.step(1e-13)
        if s % 100 == 99:
            worst = max(worst, float(np.abs(gpu.copy_m() - cpu.m).max()))
    dm = float(np.abs(gpu.copy_m() - cpu.m).max())
    print(f"config1: max|dm|={dm:.3e} worst={worst:.3e} E_gpu={g.e_total:.9e} E_cpu={r['e_total']:.9e}")
    assert dm <= 1e-10
    assert rel(g.e_total, r["e_total"]) <= 1e-9
    assert rel(g.e_ex, r["e_ex"]) <= 1e-9
    assert rel(g.e_demag, r["e_demag"]) <= 1e-9
    assert rel(g.e_ext, r["e_ext"]) <= 1e-9
    assert rel(g.max_h_eff, r["max_h_eff"]) <= 1e-9
    assert rel(g.max_h_demag, r["max_h_demag"]) <= 1e-9
    assert rel(g.max_dm_dt, r["max_dm_dt"]) <= 1e-8
    assert g.step == 1000 and abs(g.time - 1000 * 1e-13) < 1e-22


def tier_generator(seed, n):
    """Initial state generator of native/backends/fdm/tests/tier_a_compare.cu:121-135."""
    s = seed
    mask64 = (1 << 64) - 1
    out = np.zeros((n, 3))
    for i in range(n):
        v = []
        for _ in range(3):
            s ^= (s << 13) & mask64
            s ^= s >> 7
            s ^= (s << 17) & mask64
            v.append((s % 10000) / 5000.0 - 1.0)
        nrm = math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
        if nrm < 1e-10:
            v, nrm = [1.0, 0.0, 0.0], 1.0
        out[i] = [v[0] / nrm, v[1] / nrm, v[2] / nrm]
    return out


def test_tier_a_exchange_only_fp64(fa, oracle_mod, kats):
    k = kats["tier_a"]
    nx, ny, nz = k["grid"]
    m0 = tier_generator(12345, nx * ny * nz)
    gpu, cpu = make_pair(fa, oracle_mod, cells=(nx, ny, nz), cell=tuple(k["cell"]), ms=k["ms"], a=k["a"],
                         alpha=k["alpha"], m0=m0, demag=False)
    for _ in range(k["steps"]):
        g = gpu.step(k["dt"])
        r = cpu.step(k["dt"])
        assert rel(g.e_ex, r["e_ex"]) <= k["e_rel_tol"]
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= k["m_tol"]


def test_tier_b_single_vs_double(fa, oracle_mod, kats):
    k, kb = kats["tier_a"], kats["tier_b"]
    nx, ny, nz = k["grid"]
    m0 = tier_generator(12345, nx * ny * nz)
    common = dict(cells=(nx, ny, nz), cell=tuple(k["cell"]), ms=k["ms"], a=k["a"], alpha=k["alpha"], m0=m0,
                  demag=False)
    g64, _ = make_pair(fa, oracle_mod, precision="double", **common)
    g32, _ = make_pair(fa, oracle_mod, precision="single", **common)
    for _ in range(k["steps"]):
        a = g64.step(k["dt"])
        b = g32.step(k["dt"])
        assert rel(b.e_ex, a.e_ex) <= kb["e_rel_tol"]
    assert float(np.abs(g32.copy_m() - g64.copy_m()).max()) <= kb["m_tol"]


def assert_field_close(actual, expected, rel_tol, abs_tol):
    """native_fdm.rs:983-1016: |a - e| <= max(abs_tol, rel_tol max(|e|, |a|, 1))."""
    scale = np.maximum(np.maximum(np.abs(expected), np.abs(actual)), 1.0)
    assert np.all(np.abs(actual - expected) <= np.maximum(abs_tol, rel_tol * scale))


@pytest.mark.parametrize("demag", [False, True])
def test_runner_masked_plan_reference_tolerances(fa, oracle_mod, kats, demag):
    """crates/fullmag-runner/src/native_fdm.rs:1105-1270 as written: the plan's
    m0 rows are not exactly unit length; the CPU engine normalises them
    (lib.rs:701-721) while the GPU backend uploads them as given
    (native/backends/fdm/src/context.cu:733-782), hence the test's 5e-6 bar."""
    k = kats["runner_masked_plan"]
    t = k["tolerances"]
    m0 = np.array(k["m0"])
    mask = np.array(k["active_mask"], dtype=np.uint8)
    gpu, cpu = make_pair(fa, oracle_mod, cells=tuple(k["grid"]), cell=tuple(k["cell"]), ms=k["ms"], a=k["a"],
                         alpha=k["alpha"], m0=m0, h_ext=tuple(k["h_ext"]), demag=demag, mask=mask)
    g = gpu.step(k["dt"])
    r = cpu.step(k["dt"])
    o = cpu.observe()
    assert_field_close(gpu.copy_m(), cpu.m, t["m_rel"], t["m_abs"])
    assert_field_close(gpu.copy_h_ex(), o["h_ex"], t["field_rel"], t["field_abs"])
    assert_field_close(gpu.copy_h_eff(), o["h_eff"], t["field_rel"], t["field_abs"])
    assert_field_close(gpu.copy_h_ext(), o["h_ext"], 1e-12, 1e-12)
    for key, gv in (("e_ex", g.e_ex), ("e_demag", g.e_demag), ("e_ext", g.e_ext), ("e_total", g.e_total)):
        assert_field_close(np.array([gv]), np.array([r[key]]), t["energy_rel"], 0.0)


@pytest.mark.parametrize("demag", [False, True])
def test_runner_masked_plan_single_step(fa, oracle_mod, kats, demag):
    """Same plan with unit-length m0: FP64 parity to rounding."""
    k = kats["runner_masked_plan"]
    m0 = np.array(k["m0"])
    m0 = m0 / np.linalg.norm(m0, axis=1, keepdims=True)
    mask = np.array(k["active_mask"], dtype=np.uint8)
    gpu, cpu = make_pair(fa, oracle_mod, cells=tuple(k["grid"]), cell=tuple(k["cell"]), ms=k["ms"], a=k["a"],
                         alpha=k["alpha"], m0=m0, h_ext=tuple(k["h_ext"]), demag=demag, mask=mask)
    g = gpu.step(k["dt"])
    r = cpu.step(k["dt"])
    o = cpu.observe()
    np.testing.assert_allclose(gpu.copy_m(), cpu.m, rtol=0, atol=1e-13)
    scale = np.abs(o["h_eff"]).max()
    np.testing.assert_allclose(gpu.copy_h_ex(), o["h_ex"], rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(gpu.copy_h_eff(), o["h_eff"], rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(gpu.copy_h_demag(), o["h_demag"], rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(gpu.copy_h_ext(), o["h_ext"], rtol=0, atol=0)
    for key, gv in (("e_ex", g.e_ex), ("e_demag", g.e_demag), ("e_ext", g.e_ext), ("e_total", g.e_total)):
        assert abs(gv - r[key]) <= 1e-11 * max(abs(r["e_total"]), abs(r[key]))
    # masked cells stay exactly zero (native_fdm.rs:1209-1270)
    for arr in (gpu.copy_m(), gpu.copy_h_demag(), gpu.copy_h_eff(), gpu.copy_h_ex()):
        assert np.all(arr[mask == 0] == 0.0)


def test_thin_film_demag_field_matches_oracle(fa, oracle_mod, kats):
    """8x6x1 thin film (native_fdm.rs:902-947, 1422-1492), FP64."""
    nx, ny, nz = kats["thin_film_plan"]["grid"]
    m0 = fa.generate_random_unit_vectors(7, nx * ny * nz)
    gpu, cpu = make_pair(fa, oracle_mod, cells=(nx, ny, nz), cell=(5e-9, 5e-9, 2e-9), ms=8e5, a=1.3e-11,
                         alpha=0.1, m0=m0)
    hd = cpu.observe()["h_demag"]
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=1e-12 * np.abs(hd).max())


@pytest.mark.parametrize("cells", [(1, 1, 1), (3, 3, 1), (6, 5, 4), (5, 7, 3), (12, 10, 1), (8, 8, 2), (16, 4, 6)])
@pytest.mark.parametrize("kernel", ["spectra", "device"])
def test_demag_field_all_shapes(fa, oracle_mod, cells, kernel):
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(11 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(4e-9, 3e-9, 2e-9), ms=1e6, a=1e-11, alpha=0.1,
                         m0=m0, exchange=False, kernel=kernel)
    hd = cpu.observe()["h_demag"]
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=1e-12 * np.abs(hd).max())
    if cells == (1, 1, 1):
        # closed form: H = -N_self M, with N_self the cell's Newell self term
        k = oracle_mod.newell_kernels(1, 1, 1, 4e-9, 3e-9, 2e-9)
        np.testing.assert_allclose(gpu.copy_h_demag()[0], -1e6 * np.diag([k[0, 0, 0, 0], k[1, 0, 0, 0],
                                                                              k[2, 0, 0, 0]]) @ m0[0],
                                   rtol=1e-12)


@pytest.mark.parametrize("cells", [(17, 6, 1), (6, 19, 3), (5, 4, 17), (37, 3, 2), (3, 29, 1), (34, 17, 4)])
@pytest.mark.parametrize("kernel", ["spectra", "device"])
def test_demag_field_prime_factor_above_13(fa, oracle_mod, cells, kernel):
    """Padded lengths 34, 38, 58, 68 and 74 carry a prime factor above 13,
    transformed by the direct-sum prime pass (z at 34 through the LDS line
    path), at FP64 against the oracle."""
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(23 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(4e-9, 3e-9, 2e-9), ms=1e6, a=1e-11, alpha=0.1,
                         m0=m0, exchange=False, kernel=kernel)
    hd = cpu.observe()["h_demag"]
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=1e-12 * np.abs(hd).max())


@pytest.mark.parametrize("cells,precision", [((16, 1024, 1), "double"), ((8, 2048, 1), "mixed"),
                                             ((2100, 2, 1), "double"), ((4, 2048, 2), "double")])
@pytest.mark.parametrize("kernel", ["spectra", "device"])
def test_demag_field_lines_above_64k_lds(fa, oracle_mod, cells, precision, kernel):
    """Lines whose LDS footprint passes 64 KB (the thin-film y convolution's
    three 2048- and 4096-point lines, a 4200-point FP64 x line, a 4096-point
    FP64 y line and the FP64 Newell transforms) use the device's larger
    per-workgroup LDS, against the oracle.  The FP64 bound is 1e-11 of the
    largest field: rounding of 2048- to 4200-point transforms over up to 32K
    cells reaches 1.5e-12, where a wrong line layout is off by O(1)."""
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(31 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(4e-9, 3e-9, 2e-9), ms=1e6, a=1e-11, alpha=0.1,
                         m0=m0, exchange=False, kernel=kernel, precision=precision)
    hd = cpu.observe()["h_demag"]
    tol = 1e-11 if precision == "double" else 2e-6
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=tol * np.abs(hd).max())


@pytest.mark.parametrize("cells,precision,tune", [((12, 10, 1), "double", "y_split=1"),
                                                  ((17, 6, 1), "double", "y_split=1"),
                                                  ((3, 29, 1), "mixed", "y_split=1"),
                                                  ((64, 48, 1), "mixed", "y_split=1"),
                                                  ((64, 256, 1), "double", "y_split=1"),
                                                  ((8, 2048, 1), "double", ""), ((4, 4096, 1), "mixed", "")])
@pytest.mark.parametrize("kernel", ["spectra", "device"])
def test_thin_film_split_y_matches_oracle(fa, oracle_mod, monkeypatch, cells, precision, tune, kernel):
    """Thin films whose three y lines do not fit one workgroup (FP64 ny = 2048,
    FP32 ny = 4096), and small films forced onto the same path, run the y
    convolution as forward lines, a tensor pass and inverse lines."""
    monkeypatch.setenv("FULLMAG_FDM_TUNE", tune)
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(41 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(4e-9, 3e-9, 2e-9), ms=1e6, a=1e-11, alpha=0.1,
                         m0=m0, exchange=False, kernel=kernel, precision=precision)
    hd = cpu.observe()["h_demag"]
    tol = 1e-11 if precision == "double" else 2e-6
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=tol * np.abs(hd).max())


@pytest.mark.parametrize("cells", [(12, 10, 1), (17, 6, 1), (64, 32, 1)])
@pytest.mark.parametrize("precision", ["double", "mixed"])
def test_single_layer_with_3d_spectra_matches_oracle(fa, oracle_mod, cells, precision):
    """nz = 1 with spectra of the 3-D padded grid (pz = 2) takes the 3-D path,
    as native/backends/fdm/src/api.cpp:184-187 routes it; the field equals
    the thin-film field, since the padded z = 1 plane only meets padding.
    The oracle's 2-D and 3-D spectra agree to 1e-14 (test_oracle_newell);
    summed over 8 nx ny padded modes that leaves up to 2.4e-12 of the largest
    field at 64x32, hence the FP64 bound of 1e-11."""
    nx, ny, nz = cells
    n = nx * ny
    cell = (4e-9, 3e-9, 2e-9)
    m0 = fa.generate_random_unit_vectors(51 + n, n)
    spectra = oracle_mod.newell_spectra(nx, ny, 1, *cell, thin_film_2d=False)
    assert spectra.size == 6 * 2 * (2 * nx) * (2 * ny) * 2
    gpu = fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=cell, initial_magnetization=m0,
                                         saturation_magnetisation=1e6, exchange_stiffness=1e-11, damping=0.1,
                                         precision=precision, enable_exchange=False, demag_kernel=spectra))
    _, cpu = make_pair(fa, oracle_mod, cells=cells, cell=cell, ms=1e6, a=1e-11, alpha=0.1, m0=m0,
                       exchange=False, kernel="device", precision=precision)
    hd = cpu.observe()["h_demag"]
    tol = 1e-11 if precision == "double" else 2e-6
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=tol * np.abs(hd).max())


@pytest.mark.parametrize("tune", ["", "z_tile=2"])
@pytest.mark.parametrize("cells", [(4, 3, 12), (6, 4, 9), (5, 4, 16), (3, 5, 11), (16, 12, 64)])
def test_demag_field_long_z(fa, oracle_mod, monkeypatch, cells, tune):
    """nz > 8: the z convolution runs as LDS line transforms (padded lengths
    24, 18, 32, 22 and 128, with radices 3 and 11), at the automatic and at
    a narrow column tile."""
    monkeypatch.setenv("FULLMAG_FDM_TUNE", tune)
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(5 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(3e-9, 4e-9, 2e-9), ms=9e5, a=1e-11, alpha=0.1,
                         m0=m0, exchange=False, kernel="device")
    hd = cpu.observe()["h_demag"]
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=1e-12 * np.abs(hd).max())


# Launch-schedule knobs change only how the transforms are scheduled; every
# schedule must give the oracle's field and trajectory.
SCHEDULES = ["", "fixed=0", "z_quad=1", "z_quad=0", "x_rows=4,y_lpw=1", "xcd=0", "x_radix=4,y_radix=8",
             "x_lanes=128,y_lanes=64", "y_pipe=0", "y_pipe=3", "zc=1", "zc=3,xcd=0", "x_rows=4", "x_rows=2",
             "y_pipe_f=1", "y_pipe_f=4,y_pipe=1", "y_pipe=4,y_pipe_f=2", "x_radix=8", "y_radix=4,fixed=0"]


@pytest.mark.parametrize("cells", [(32, 16, 4), (64, 32, 8), (16, 16, 6)])
@pytest.mark.parametrize("tune", SCHEDULES)
def test_demag_schedule_variants_match_oracle(fa, oracle_mod, monkeypatch, cells, tune):
    monkeypatch.setenv("FULLMAG_FDM_TUNE", tune)
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(17 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(4e-9, 4e-9, 2e-9), ms=8.6e5, a=1.3e-11, alpha=0.05,
                         m0=m0, h_ext=(1e4, 0.0, 2e3), kernel="device")
    hd = cpu.observe()["h_demag"]
    # other FFT factorisations round differently: 1e-11 of max|H| (the
    # trajectory bar is max|dm| <= 1e-10)
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=1e-11 * np.abs(hd).max())
    for _ in range(3):
        gpu.step(5e-14)
        cpu.step(5e-14)
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-12


@pytest.mark.parametrize("cells,tune", [((13, 7, 1), ""), ((10, 9, 3), ""), ((10, 9, 3), "x_rows=2"),
                                        ((64, 48, 8), "z_quad=1"), ((33, 17, 5), "x_rows=4"),
                                        ((19, 17, 2), ""), ((7, 5, 1), "")])
def test_mixed_demag_odd_shapes_match_oracle(fa, oracle_mod, monkeypatch, cells, tune):
    """FP32 demag (mixed precision) on odd grids, including odd thin-film row
    pitches (ny odd), where the x forward unpack stores rows one at a time,
    against the FP64 oracle field to FP32 rounding."""
    monkeypatch.setenv("FULLMAG_FDM_TUNE", tune)
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(61 + n, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(4e-9, 3e-9, 2e-9), ms=8.6e5, a=1.3e-11, alpha=0.05,
                         m0=m0, exchange=False, kernel="device", precision="mixed")
    hd = cpu.observe()["h_demag"]
    np.testing.assert_allclose(gpu.copy_h_demag(), hd, rtol=0, atol=2e-6 * np.abs(hd).max())


@pytest.mark.parametrize("cells,tf", [((8, 6, 1), True), ((6, 5, 4), False), ((45, 3, 2), False),
                                      ((7, 9, 1), False), ((17, 5, 1), True), ((5, 19, 3), False)])
def test_in_library_newell_spectra_match_oracle(fa, oracle_mod, cells, tf):
    s_gpu = fa.compute_newell_kernel_spectra(*cells, 3e-9, 4e-9, 2e-9, thin_film_2d=tf)
    s_cpu = oracle_mod.newell_spectra(*cells, 3e-9, 4e-9, 2e-9, thin_film_2d=tf)
    scale = np.abs(s_cpu).max()
    np.testing.assert_allclose(s_gpu, s_cpu, rtol=0, atol=1e-11 * scale)


@pytest.mark.parametrize("integrator", ["heun", "rk4", "rk23", "dp45", "abm3"])
def test_integrators_fp64_match_oracle(fa, oracle_mod, integrator):
    cells = (16, 8, 2)
    n = 16 * 8 * 2
    m0 = fa.generate_random_unit_vectors(3, n)
    adaptive = dict(adaptive_max_error=1e-4, adaptive_dt_min=1e-16, adaptive_dt_max=4e-13, adaptive_headroom=0.8)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(3e-9, 3e-9, 3e-9), ms=8e5, a=1.3e-11, alpha=0.2,
                         m0=m0, h_ext=(2e4, -1e4, 5e3), integrator=integrator, adaptive=adaptive)
    dt = 1e-13
    for _ in range(25):
        g = gpu.step(dt)
        r = cpu.step(dt)
        assert abs(g.dt - r["dt_used"]) <= 1e-9 * r["dt_used"]
        if r["suggested_next_dt"] > 0:
            assert rel(g.dt_suggested, r["suggested_next_dt"]) <= 1e-6
            dt = r["suggested_next_dt"]
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-10
    assert rel(g.e_total, r["e_total"]) <= 1e-9


def test_relax_mode_disables_precession(fa, oracle_mod):
    """native_fdm.rs:1494-1531: damping-only RHS."""
    m0 = np.tile([1.0, 0.0, 0.0], (4, 1))
    gpu, cpu = make_pair(fa, oracle_mod, cells=(4, 1, 1), cell=(1e-9, 1e-9, 1e-9), ms=8e5, a=1e-11, alpha=0.5,
                         m0=m0, h_ext=(0.0, 0.0, 1e5), demag=False, disable_precession=True)
    gpu.step(1e-13)
    cpu.step(1e-13)
    m = gpu.copy_m()
    assert np.abs(m[:, 1]).max() <= 1e-12
    assert m[0, 2] > 0.0
    np.testing.assert_allclose(m, cpu.m, atol=1e-14)


def smooth_state(nx, ny, nz):
    """Uniform magnetisation with a gentle in-plane twist: no sub-cell-scale
    gradients, so the dynamics do not amplify rounding differences."""
    x = np.arange(nx)[None, None, :] * np.ones((nz, ny, 1))
    th = 0.3 * np.sin(2 * np.pi * x / nx)
    m = np.stack([np.cos(th), np.sin(th), 0.05 * np.ones_like(th)], -1).reshape(-1, 3)
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def test_mixed_precision_tracks_fp64_smooth_state(fa, oracle_mod):
    cells = (64, 64, 4)
    common = dict(cells=cells, cell=(4e-9, 4e-9, 2e-9), ms=8.6e5, a=1.3e-11, alpha=0.02,
                  m0=smooth_state(*cells), h_ext=(1e4, 2e4, -5e3), kernel="device")
    g64, _ = make_pair(fa, oracle_mod, precision="double", **common)
    gmx, _ = make_pair(fa, oracle_mod, precision="mixed", **common)
    assert gmx.precision_mode() == 3 and g64.precision_mode() == 2
    for _ in range(1000):
        a = g64.step(5e-14)
        b = gmx.step(5e-14)
    dm = float(np.abs(gmx.copy_m() - g64.copy_m()).max())
    print(f"mixed vs fp64, smooth state, 1000 steps: max|dm|={dm:.3e} dE/E={rel(b.e_total, a.e_total):.3e}")
    assert dm <= 1e-5
    assert rel(b.e_total, a.e_total) <= 1e-6
    np.testing.assert_allclose(np.linalg.norm(gmx.copy_m(), axis=1), 1.0, atol=1e-14)


def test_mixed_precision_random_state_within_fp64_chaos_envelope(fa, oracle_mod):
    """From seeded random m0 at alpha 0.02 the dynamics amplify any
    perturbation (an FP64 perturbation of 1e-12 grows ~700x in 100 steps),
    so max|dm| <= 1e-5 between FP32 and FP64 arithmetic holds only over a
    finite horizon; check it there and check the energy stays close."""
    cells = (64, 64, 4)
    n = 64 * 64 * 4
    m0 = fa.generate_random_unit_vectors(5, n)
    common = dict(cells=cells, cell=(4e-9, 4e-9, 2e-9), ms=8.6e5, a=1.3e-11, alpha=0.02, m0=m0,
                  h_ext=(1e4, 2e4, -5e3), kernel="device")
    g64, _ = make_pair(fa, oracle_mod, precision="double", **common)
    gmx, _ = make_pair(fa, oracle_mod, precision="mixed", **common)
    rng = np.random.default_rng(1)
    mp = m0 + 1e-12 * rng.standard_normal(m0.shape)
    common["m0"] = mp / np.linalg.norm(mp, axis=1, keepdims=True)
    gpt, _ = make_pair(fa, oracle_mod, precision="double", **common)
    for _ in range(50):
        a = g64.step(5e-14)
        b = gmx.step(5e-14)
        gpt.step(5e-14)
    dm = float(np.abs(gmx.copy_m() - g64.copy_m()).max())
    amp = float(np.abs(gpt.copy_m() - g64.copy_m()).max()) / 1e-12
    print(f"mixed vs fp64, random state, 50 steps: max|dm|={dm:.3e}, fp64 amplification of 1e-12: {amp:.1f}x")
    assert dm <= 1e-5
    assert amp > 1.0
    assert rel(b.e_total, a.e_total) <= 1e-6


@pytest.mark.parametrize("precision,per_cell_ms", [("double", False), ("double", True), ("mixed", False),
                                                   ("single", False)])
def test_masked_disc_heun_matches_oracle(fa, oracle_mod, precision, per_cell_ms):
    """A disc with a hole and scattered inactive cells on 32x32x4 cells of
    3x3x2 nm: power-of-two x lines, so the fixed-length x transforms read the
    mask (and per-cell Ms) beside m, and without per-cell Ms the Heun stages
    and the stats pass run the uniform stage kernel's masked form.  40 Heun
    steps of 1e-13 s from a random m0 against the oracle: FP64 max|dm| <=
    1e-11 and energies <= 1e-10 relative; mixed and single 1e-5 and 1e-4."""
    nx, ny, nz = 32, 32, 4
    n = nx * ny * nz
    yy, xx = np.mgrid[0:ny, 0:nx]
    r2 = (xx - 15.5) ** 2 + (yy - 15.5) ** 2
    disc = ((r2 <= 15.0 ** 2) & (r2 >= 4.0 ** 2)).astype(np.uint8)
    mask = np.tile(disc.ravel(), nz)
    mask[np.random.default_rng(11).choice(n, 40, replace=False)] = 0
    m0 = fa.generate_random_unit_vectors(2024, n)
    m0[mask == 0] = 0.0
    ms_cell = None
    if per_cell_ms:
        ms_cell = np.random.default_rng(12).uniform(6e5, 1e6, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=(nx, ny, nz), cell=(3e-9, 3e-9, 2e-9), ms=8e5, a=1.3e-11,
                         alpha=0.1, m0=m0, h_ext=(2e4, -1e4, 5e3), precision=precision, mask=mask,
                         ms_cell=ms_cell)
    for _ in range(40):
        g = gpu.step(1e-13)
        r = cpu.step(1e-13)
    dm = float(np.abs(gpu.copy_m() - cpu.m).max())
    print(f"masked disc {precision} per-cell Ms={per_cell_ms}: max|dm|={dm:.3e}")
    tol_m, tol_e = (1e-11, 1e-10) if precision == "double" else (1e-5, 1e-4)
    assert dm <= tol_m
    for key, gv in (("e_ex", g.e_ex), ("e_demag", g.e_demag), ("e_ext", g.e_ext), ("e_total", g.e_total)):
        assert rel(gv, r[key]) <= tol_e, key
    assert rel(g.max_h_eff, r["max_h_eff"]) <= tol_e
    gpu.close()


def test_single_precision_runner_plan_close_to_double(fa, oracle_mod, kats):
    """native_fdm.rs:1272-1346: single vs double m <= 1e-5, E <= 1e-4 rel."""
    k = kats["runner_masked_plan"]
    m0 = np.array(k["m0"])
    mask = np.array(k["active_mask"], dtype=np.uint8)
    common = dict(cells=tuple(k["grid"]), cell=tuple(k["cell"]), ms=k["ms"], a=k["a"], alpha=k["alpha"], m0=m0,
                  h_ext=tuple(k["h_ext"]), mask=mask)
    g64, _ = make_pair(fa, oracle_mod, precision="double", **common)
    g32, _ = make_pair(fa, oracle_mod, precision="single", **common)
    a = g64.step(k["dt"])
    b = g32.step(k["dt"])
    assert float(np.abs(g32.copy_m() - g64.copy_m()).max()) <= 1e-5
    assert rel(b.e_total, a.e_total) <= 1e-4
    f32 = g32.copy_field_f32("m")
    np.testing.assert_array_equal(f32.astype(np.float64), g32.copy_m())


def test_upload_refresh_preview_snapshot(fa, oracle_mod):
    cells = (10, 8, 3)
    n = 240
    m0 = fa.generate_random_unit_vectors(9, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=cells, cell=(2e-9, 2e-9, 2e-9), ms=8e5, a=1.3e-11, alpha=0.3,
                         m0=m0, h_ext=(0.0, 1e4, 0.0))
    np.testing.assert_array_equal(gpu.copy_m(), m0)
    hd0 = gpu.copy_h_demag()
    m1 = fa.generate_random_unit_vectors(10, n)
    gpu.upload_magnetization(m1)
    np.testing.assert_array_equal(gpu.copy_m(), m1)
    np.testing.assert_array_equal(gpu.copy_h_demag(), hd0)  # stale until refresh
    gpu.refresh_observables()
    cpu.m = m1
    o = cpu.observe()
    np.testing.assert_allclose(gpu.copy_h_eff(), o["h_eff"], atol=1e-12 * np.abs(o["h_eff"]).max())
    # preview: box average (context.cu:222-286)
    pv = gpu.copy_live_preview_field("m", (5, 4, 1), z_origin=1, z_stride=2)
    full = m1.reshape(3, 8, 10, 3)[1:3].reshape(2, 4, 2, 5, 2, 3).mean(axis=(0, 2, 4)).reshape(-1, 3)
    np.testing.assert_allclose(pv, full, atol=1e-15)
    snap = gpu.begin_field_snapshot("m")
    soa = snap.wait()
    np.testing.assert_array_equal(soa.T, m1)
    snap.close()
    info = gpu.device_info()
    assert "MI355" in info["name"] or info["compute_capability"][0] == 9


def test_error_paths_report_last_error(fa):
    m0 = np.tile([1.0, 0.0, 0.0], (8, 1))
    with pytest.raises(fa.RunError, match="active_mask_len"):
        fa.NativeFdmBackend(fa.FdmPlan(cells=(2, 2, 2), cell_size=(1e-9,) * 3, initial_magnetization=m0,
                                       saturation_magnetisation=8e5, exchange_stiffness=1e-11, damping=0.1,
                                       active_mask=np.ones(7, np.uint8), enable_demag=False))
    with pytest.raises(fa.RunError, match="demag_kernel_spectrum_len"):
        fa.NativeFdmBackend(fa.FdmPlan(cells=(2, 2, 2), cell_size=(1e-9,) * 3, initial_magnetization=m0,
                                       saturation_magnetisation=8e5, exchange_stiffness=1e-11, damping=0.1,
                                       demag_kernel=np.zeros((6, 10))))
    b = fa.NativeFdmBackend(fa.FdmPlan(cells=(2, 2, 2), cell_size=(1e-9,) * 3, initial_magnetization=m0,
                                       saturation_magnetisation=8e5, exchange_stiffness=1e-11, damping=0.1,
                                       enable_demag=False))
    with pytest.raises(fa.RunError, match="dt must be positive"):
        b.step(-1.0)
    # a 3-D y line of 16384 points would need 1024 lanes: rejected at set-up
    m1 = np.tile([1.0, 0.0, 0.0], (2 * 8192 * 2, 1))
    with pytest.raises(fa.RunError, match="y line FFT does not fit"):
        fa.NativeFdmBackend(fa.FdmPlan(cells=(2, 8192, 2), cell_size=(1e-9,) * 3, initial_magnetization=m1,
                                       saturation_magnetisation=8e5, exchange_stiffness=1e-11, damping=0.1,
                                       demag_kernel="device"))


def test_region_table_decoupled_equals_masked_neighbour(fa, oracle_mod):
    """Two regions with A_12 = 0: the exchange field in region 1 equals the
    field of the same cells when region 2 is inactive (both make the cross
    link a free surface, exchange_fp64.cu:61-76 and :80-114)."""
    nx, ny, nz = 8, 6, 2
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(21, n)
    region = np.zeros(n, np.uint32)
    region.reshape(nz, ny, nx)[:, :, 4:] = 1
    common = dict(cells=(nx, ny, nz), cell_size=(2e-9, 2e-9, 2e-9), initial_magnetization=m0,
                  saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1, enable_demag=False)
    a = fa.NativeFdmBackend(fa.FdmPlan(region_mask=region, **common))
    b = fa.NativeFdmBackend(fa.FdmPlan(active_mask=(region == 0).astype(np.uint8), **common))
    sel = region == 0
    np.testing.assert_allclose(a.copy_h_ex()[sel], b.copy_h_ex()[sel], rtol=1e-13, atol=1e-9)
    # a uniform region table reduces to the plain exchange path
    c = fa.NativeFdmBackend(fa.FdmPlan(region_mask=np.zeros(n, np.uint32), **common))
    d = fa.NativeFdmBackend(fa.FdmPlan(**common))
    np.testing.assert_array_equal(c.copy_h_ex(), d.copy_h_ex())


@pytest.mark.parametrize("precision,masked", [("double", False), ("double", True), ("mixed", False)])
def test_region_table_coupled_regions_match_oracle(fa, oracle_mod, precision, masked):
    """Three regions with A_01 = 0.3 A, A_12 = 1.7 A and A_02 = 0 (per-pair
    table of native/backends/fdm/src/exchange_fp64.cu:80-114) against the
    oracle's table exchange: fields, energies and 20 Heun steps with demag."""
    nx, ny, nz = 12, 9, 3
    n = nx * ny * nz
    a = 1.3e-11
    region = np.zeros(n, np.uint32)
    r3 = region.reshape(nz, ny, nx)
    r3[:, :, 4:8] = 1
    r3[:, :, 8:] = 2
    r3[2, :, :] = np.where(r3[2] == 0, 2, r3[2])
    mask = None
    if masked:
        mask = np.ones(n, np.uint8)
        mask[np.random.default_rng(4).choice(n, 30, replace=False)] = 0
    m0 = fa.generate_random_unit_vectors(314, n)
    gpu, cpu = make_pair(fa, oracle_mod, cells=(nx, ny, nz), cell=(3e-9, 2e-9, 4e-9), ms=8e5, a=a, alpha=0.1,
                         m0=m0, h_ext=(5e3, 0.0, 1e4), precision=precision, mask=mask, region=region,
                         pairs=[(0, 1, 0.3 * a), (1, 2, 1.7 * a), (0, 2, 0.0)])
    o = cpu.observe()
    tol_f = 1e-11 if precision == "double" else 2e-6
    scale = np.abs(o["h_ex"]).max()
    np.testing.assert_allclose(gpu.copy_h_ex(), o["h_ex"], rtol=0, atol=tol_f * scale)
    for _ in range(20):
        g = gpu.step(1e-13)
        r = cpu.step(1e-13)
    dm = float(np.abs(gpu.copy_m() - cpu.m).max())
    print(f"region table {precision} masked={masked}: max|dm|={dm:.3e}")
    assert dm <= (1e-10 if precision == "double" else 1e-5)
    tol_e = 1e-9 if precision == "double" else 1e-5
    assert rel(g.e_ex, r["e_ex"]) <= tol_e
    assert rel(g.e_total, r["e_total"]) <= tol_e


def test_config2_full_size_properties(fa):
    """512x512x8 thin film at full size (the headline workload): demag is
    linear in M (uploads are not renormalised), |m| stays 1, mixed and FP64
    stay within 1e-5 over a short run."""
    nx, ny, nz = 512, 512, 8
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(2024, n)
    rng = np.random.default_rng(0)
    plan = dict(cells=(nx, ny, nz), cell_size=(4e-9, 4e-9, 2e-9), saturation_magnetisation=8.6e5,
                exchange_stiffness=1.3e-11, damping=0.02)
    g = fa.NativeFdmBackend(fa.FdmPlan(initial_magnetization=m0, enable_exchange=False, **plan))
    m1 = rng.standard_normal((n, 3))
    m2 = rng.standard_normal((n, 3))
    g.upload_magnetization(m1)
    g.refresh_observables()
    h1 = g.copy_h_demag()
    g.upload_magnetization(m2)
    g.refresh_observables()
    h2 = g.copy_h_demag()
    g.upload_magnetization(0.5 * m1 - 2.0 * m2)
    g.refresh_observables()
    h12 = g.copy_h_demag()
    np.testing.assert_allclose(h12, 0.5 * h1 - 2.0 * h2, atol=1e-11 * np.abs(h12).max())
    g.close()
    bext = (0.03 / MU0) * np.array([0.6, 0.0, 0.8])
    g64 = fa.NativeFdmBackend(fa.FdmPlan(initial_magnetization=m0, external_field=bext, **plan))
    gmx = fa.NativeFdmBackend(fa.FdmPlan(initial_magnetization=m0, external_field=bext, precision="mixed", **plan))
    for _ in range(50):
        a = g64.step(5e-14)
        b = gmx.step(5e-14)
    ma, mb = g64.copy_m(), gmx.copy_m()
    np.testing.assert_allclose(np.linalg.norm(ma, axis=1), 1.0, atol=1e-14)
    dm = float(np.abs(ma - mb).max())
    print(f"512x512x8 random m0, 50 steps: mixed vs FP64 max|dm| {dm:.3e}, rel dE {rel(b.e_total, a.e_total):.3e}")
    assert dm <= 1e-5
    assert rel(b.e_total, a.e_total) <= 1e-7


def test_config2_full_size_smooth_m0_1000_steps(fa):
    """The headline accuracy bar at full size over a longer horizon: from a
    smooth m0 (uniform with a gentle in-plane twist along x), 1000 Heun
    steps of 5e-14 s on the 512x512x8 film with a random-direction field;
    mixed precision stays within max|dm| <= 1e-5 of the FP64 path and its
    total energy within 1e-6 of the energy terms' magnitude, with |m| = 1."""
    nx, ny, nz = 512, 512, 8
    x = np.arange(nx)[None, None, :] * np.ones((nz, ny, 1))
    th = 0.3 * np.sin(2 * np.pi * x / nx)
    m0 = np.stack([np.cos(th), np.sin(th), 0.05 * np.ones_like(th)], -1).reshape(-1, 3)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    bext = (0.02 / MU0) * np.array([0.36, -0.48, 0.8])
    plan = dict(cells=(nx, ny, nz), cell_size=(4e-9, 4e-9, 2e-9), saturation_magnetisation=8.6e5,
                exchange_stiffness=1.3e-11, damping=0.02, initial_magnetization=m0, external_field=bext)
    g64 = fa.NativeFdmBackend(fa.FdmPlan(**plan))
    gmx = fa.NativeFdmBackend(fa.FdmPlan(precision="mixed", **plan))
    a = g64.step_many(5e-14, 1000)
    b = gmx.step_many(5e-14, 1000)
    ma, mb = g64.copy_m(), gmx.copy_m()
    np.testing.assert_allclose(np.linalg.norm(mb, axis=1), 1.0, atol=1e-14)
    dm = float(np.abs(ma - mb).max())
    # the Zeeman term nearly cancels the others here, so the energy error is
    # taken relative to the size of the terms
    scale = abs(a.e_ex) + abs(a.e_demag) + abs(a.e_ext)
    de = abs(b.e_total - a.e_total) / scale
    print(f"512x512x8 smooth m0, 1000 steps: mixed vs FP64 max|dm| {dm:.3e}, |dE| / sum|E_i| {de:.3e} "
          f"(E_total {a.e_total:.4e} J of terms summing to {scale:.4e} J)")
    assert dm <= 1e-5
    assert de <= 1e-6
    g64.close()
    gmx.close()


def _box_means(field, cells, pcells, z0, zs):
    """Box averages with the bin edges of native/backends/fdm/src/context.cu:
    222-286: [b n / pn, (b + 1) n / pn), at least one cell, clamped; z planes
    [z0 + b zs, z0 + (b + 1) zs) clamped to the grid."""
    nx, ny, nz = cells
    pnx, pny, pnz = pcells
    f = field.reshape(nz, ny, nx, 3)

    def edges(b, n, pn):
        lo, hi = b * n // pn, (b + 1) * n // pn
        hi = max(hi, lo + 1)
        return lo, min(hi, n)

    out = np.empty((pnz, pny, pnx, 3))
    for bz in range(pnz):
        za = min(z0 + bz * zs, nz - 1)
        zb = min(max(z0 + (bz + 1) * zs, za + 1), nz)
        for by in range(pny):
            ya, yb = edges(by, ny, pny)
            for bx in range(pnx):
                xa, xb = edges(bx, nx, pnx)
                out[bz, by, bx] = f[za:zb, ya:yb, xa:xb].reshape(-1, 3).mean(axis=0)
    return out.reshape(-1, 3)


@pytest.mark.parametrize("pcells,z0,zs", [((3, 2, 1), 0, 4), ((5, 3, 2), 1, 2), ((7, 11, 4), 0, 1), ((48, 40, 4), 0, 1),
                                          ((1, 1, 1), 2, 9)])
def test_preview_box_means_large_and_uneven_bins(fa, pcells, z0, zs):
    """Preview bins of up to 1920 cells (a wavefront's lanes loop over them),
    uneven bin edges, z strides past the grid, and full resolution, against
    numpy box means; FP64 and FP32 outputs."""
    cells = (48, 40, 4)
    n = 48 * 40 * 4
    m0 = fa.generate_random_unit_vectors(31, n)
    b = fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=(4e-9, 4e-9, 2e-9), initial_magnetization=m0,
                                       saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1,
                                       precision="double", demag_kernel="device"))
    hd = b.copy_h_demag()
    for name, field in (("m", m0), ("H_demag", hd)):
        ref = _box_means(field, cells, pcells, z0, zs)
        got = b.copy_live_preview_field(name, pcells, z0, zs)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(field).max())
        got32 = b.copy_live_preview_field(name, pcells, z0, zs, f32=True)
        np.testing.assert_allclose(got32, ref, rtol=0, atol=1e-6 * np.abs(field).max())
    b.close()
