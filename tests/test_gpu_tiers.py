"""Sub-cell boundary tiers through the C ABI against the oracle: T0 face-link
exchange (boundary_correction "volume"), T1 intersection-distance exchange
("full"), the volume-fraction demag source and the sparse demag correction
rows, on a cylinder sampled as the planner samples it (tests/_tiers.py).
FP64 fields and trajectories to rounding; mixed precision to FP32 rounding;
single precision ignores the tiers, as the reference's FP32 kernels do."""

import numpy as np
import pytest

from tests import _tiers

pytestmark = pytest.mark.gpu

CELL = (3e-9, 4e-9, 2e-9)
MS, A_EX = 8e5, 1.3e-11


def problem(fa, O, cells, mode, precision="double", integrator="heun", corr=True, adaptive=None, seed=7,
            tiers=True):
    nx, ny, nz = cells
    n = nx * ny * nz
    phi, faces, deltas = _tiers.geometry(nx, ny, nz, *CELL)
    mask = (phi > 0).astype(np.uint8)
    m0 = fa.generate_random_unit_vectors(seed, n)
    m0[mask == 0] = 0.0
    tgt = src = ten = None
    if corr:
        tgt, src, ten = _tiers.correction_rows(nx, ny, nz, phi)
    names = ("xp", "xm", "yp", "ym", "zp", "zm")
    bg = fa.BoundaryGeometry(volume_fraction=phi, **{f"face_link_{k}": faces[i] for i, k in enumerate(names)},
                             **{f"delta_{k}": deltas[i] for i, k in enumerate(names)},
                             demag_corr_target_idx=tgt, demag_corr_source_idx=src, demag_corr_tensor=ten,
                             demag_corr_stencil_size=0 if src is None else src.shape[1])
    adaptive = adaptive or {}
    plan = fa.FdmPlan(cells=cells, cell_size=CELL, initial_magnetization=m0, saturation_magnetisation=MS,
                      exchange_stiffness=A_EX, damping=0.2, precision=precision, integrator=integrator,
                      external_field=(2e4, -1e4, 5e3), active_mask=mask, demag_kernel="device",
                      boundary_correction=mode if tiers else None, boundary_geometry=bg if tiers else None,
                      **adaptive)
    gpu = fa.NativeFdmBackend(plan)
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=nz, dx=CELL[0], dy=CELL[1], dz=CELL[2], ms=MS, a=A_EX, alpha=0.2,
                         integrator=integrator, h_ext=(2e4, -1e4, 5e3), active_mask=mask,
                         tier={"volume": 1, "full": 2}[mode] if tiers else 0, phi=phi, face_links=faces,
                         deltas=deltas, corr_target=tgt, corr_source=src, corr_tensor=ten,
                         max_error=adaptive.get("adaptive_max_error", 0.0),
                         dt_min=adaptive.get("adaptive_dt_min", 0.0), dt_max=adaptive.get("adaptive_dt_max", 0.0),
                         headroom=adaptive.get("adaptive_headroom", 0.0))
    cpu = O.OracleProblem(cfg, m0)
    return gpu, cpu


def check_fields(gpu, o, tol):
    scale = np.abs(o["h_eff"]).max()
    np.testing.assert_allclose(gpu.copy_h_ex(), o["h_ex"], rtol=0, atol=tol * scale)
    np.testing.assert_allclose(gpu.copy_h_demag(), o["h_demag"], rtol=0, atol=tol * scale)
    np.testing.assert_allclose(gpu.copy_h_eff(), o["h_eff"], rtol=0, atol=tol * scale)


@pytest.mark.parametrize("cells", [(24, 20, 3), (32, 28, 1)])
@pytest.mark.parametrize("mode", ["volume", "full"])
def test_tier_fields_and_heun_match_oracle(fa, oracle_mod, cells, mode):
    gpu, cpu = problem(fa, oracle_mod, cells, mode)
    check_fields(gpu, cpu.observe(), 1e-11)
    for _ in range(20):
        g = gpu.step(5e-14)
        r = cpu.step(5e-14)
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-12
    for key, gv in (("e_ex", g.e_ex), ("e_demag", g.e_demag), ("e_ext", g.e_ext), ("e_total", g.e_total)):
        assert abs(gv - r[key]) <= 1e-11 * max(abs(r["e_total"]), abs(r[key])), key
    assert abs(g.max_h_eff - r["max_h_eff"]) <= 1e-11 * r["max_h_eff"]
    assert abs(g.max_h_demag - r["max_h_demag"]) <= 1e-11 * r["max_h_demag"]
    check_fields(gpu, cpu.observe(), 1e-11)
    gpu.close()


@pytest.mark.parametrize("mode", ["volume", "full"])
def test_tier_dp45_matches_oracle(fa, oracle_mod, mode):
    adaptive = dict(adaptive_max_error=1e-4, adaptive_dt_min=1e-16, adaptive_dt_max=4e-13, adaptive_headroom=0.8)
    gpu, cpu = problem(fa, oracle_mod, (20, 16, 2), mode, integrator="dp45", adaptive=adaptive)
    dt = 1e-13
    for _ in range(12):
        g = gpu.step(dt)
        r = cpu.step(dt)
        assert abs(g.dt - r["dt_used"]) <= 1e-9 * r["dt_used"]
        if r["suggested_next_dt"] > 0:
            dt = r["suggested_next_dt"]
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-10
    gpu.close()


@pytest.mark.parametrize("mode", ["volume", "full"])
def test_tier_mixed_precision_matches_oracle(fa, oracle_mod, mode):
    gpu, cpu = problem(fa, oracle_mod, (24, 20, 3), mode, precision="mixed")
    check_fields(gpu, cpu.observe(), 3e-6)
    for _ in range(10):
        gpu.step(5e-14)
        cpu.step(5e-14)
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-6
    gpu.close()


def test_single_precision_ignores_tiers(fa, oracle_mod):
    a, _ = problem(fa, oracle_mod, (24, 20, 3), "full", precision="single")
    b, _ = problem(fa, oracle_mod, (24, 20, 3), "full", precision="single", tiers=False)
    for _ in range(5):
        sa, sb = a.step(5e-14), b.step(5e-14)
    np.testing.assert_array_equal(a.copy_m(), b.copy_m())
    np.testing.assert_array_equal(a.copy_h_ex(), b.copy_h_ex())
    assert sa.e_total == sb.e_total
    a.close()
    b.close()


def test_tier_rejections(fa, oracle_mod):
    nx, ny, nz = 12, 10, 2
    n = nx * ny * nz
    phi, faces, _ = _tiers.geometry(nx, ny, nz, *CELL)
    for tgt, src in ((np.array([n], np.int32), np.array([[0]], np.int32)),
                     (np.array([3], np.int32), np.array([[n + 5]], np.int32))):
        bg = fa.BoundaryGeometry(volume_fraction=phi, demag_corr_target_idx=tgt, demag_corr_source_idx=src,
                                 demag_corr_tensor=np.zeros(6), demag_corr_stencil_size=1)
        plan = fa.FdmPlan(cells=(nx, ny, nz), cell_size=CELL, initial_magnetization=np.tile([1.0, 0, 0], (n, 1)),
                          saturation_magnetisation=MS, exchange_stiffness=A_EX, damping=0.1,
                          boundary_correction="volume", boundary_geometry=bg)
        with pytest.raises(fa.RunError, match="is not a cell index"):
            fa.NativeFdmBackend(plan)
    # the Heun adjoint does not cover the tiers
    import torch

    gpu, _ = problem(fa, oracle_mod, (nx, ny, nz), "volume")
    buf = torch.zeros(3, 3 * n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(fa.RunError, match="boundary tiers"):
        gpu.heun_adjoint(1e-13, buf[0].data_ptr(), buf[1].data_ptr(), buf[2].data_ptr())
    gpu.close()


def test_tier_member_in_batch_matches_eager(fa, oracle_mod):
    """A batch holding tier members (their correction pass has no
    multi-instance form, so they run their own launches) and a plain member
    ends bit-identical to eager stepping."""
    def members():
        a, _ = problem(fa, oracle_mod, (24, 20, 3), "volume", seed=1)
        b, _ = problem(fa, oracle_mod, (24, 20, 3), "full", seed=2, precision="mixed")
        c, _ = problem(fa, oracle_mod, (24, 20, 3), "full", seed=3, tiers=False)
        return [a, b, c]

    batch, eager = members(), members()
    run = fa.BatchRun(batch, [5e-14] * 3, steps_per_graph=3)
    run.run(7)
    for e in eager:
        for _ in range(7):
            e.step(5e-14)
    for b, e in zip(batch, eager):
        np.testing.assert_array_equal(b.copy_m(), e.copy_m())
    run.close()
    for x in batch + eager:
        x.close()


def test_t1_per_cell_ms_dp45_mixed_matches_oracle(fa, oracle_mod):
    """T1 with per-cell Ms (extension: it enters the tier prefactor and the
    correction source as Ms_i), mixed precision, DP45 with dt feedback."""
    O = oracle_mod
    nx, ny, nz = 20, 16, 2
    n = nx * ny * nz
    phi, faces, deltas = _tiers.geometry(nx, ny, nz, *CELL)
    tgt, src, ten = _tiers.correction_rows(nx, ny, nz, phi, seed=5)
    mask = (phi > 0).astype(np.uint8)
    ms_cell = np.random.default_rng(3).uniform(6e5, 1e6, n)
    m0 = fa.generate_random_unit_vectors(17, n)
    m0[mask == 0] = 0.0
    names = ("xp", "xm", "yp", "ym", "zp", "zm")
    bg = fa.BoundaryGeometry(volume_fraction=phi, **{f"delta_{k}": deltas[i] for i, k in enumerate(names)},
                             demag_corr_target_idx=tgt, demag_corr_source_idx=src, demag_corr_tensor=ten,
                             demag_corr_stencil_size=src.shape[1])
    adaptive = dict(adaptive_max_error=1e-4, adaptive_dt_min=1e-16, adaptive_dt_max=4e-13, adaptive_headroom=0.8)
    gpu = fa.NativeFdmBackend(fa.FdmPlan(cells=(nx, ny, nz), cell_size=CELL, initial_magnetization=m0,
                                         saturation_magnetisation=MS, exchange_stiffness=A_EX, damping=0.2,
                                         precision="mixed", integrator="dp45", external_field=(2e4, -1e4, 5e3),
                                         active_mask=mask, demag_kernel="device", ms_cell=ms_cell,
                                         boundary_correction="full", boundary_geometry=bg, **adaptive))
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=nz, dx=CELL[0], dy=CELL[1], dz=CELL[2], ms=MS, a=A_EX, alpha=0.2,
                         integrator="dp45", h_ext=(2e4, -1e4, 5e3), active_mask=mask, ms_cell=ms_cell, tier=2,
                         phi=phi, deltas=deltas, corr_target=tgt, corr_source=src, corr_tensor=ten,
                         max_error=1e-4, dt_min=1e-16, dt_max=4e-13, headroom=0.8)
    cpu = O.OracleProblem(cfg, m0)
    check_fields(gpu, cpu.observe(), 3e-6)
    dt = 1e-13
    for _ in range(8):
        g = gpu.step(dt)
        r = cpu.step(dt)
        assert abs(g.dt - r["dt_used"]) <= 1e-5 * r["dt_used"]
        dt = r["suggested_next_dt"] or dt
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-6
    gpu.close()


@pytest.mark.parametrize("mode", ["volume", "full"])
def test_planner_ring_geometry_matches_oracle(fa, oracle_mod, mode):
    """A 1 um permalloy ring with a 150 nm hole on 200x200 cells of
    5x5x10 nm (the planner's cylinder-minus-cylinder SDF), geometry and
    correction rows sampled by fullmag_amd.boundary with the library's Newell
    kernels; Ms 8e5, A 1.3e-11, alpha 0.5, m0 = x, 30 FP64 Heun steps of
    1e-13 s against the oracle fed the same arrays."""
    from fullmag_amd import boundary

    O = oracle_mod
    nx = ny = 200
    cell = (5e-9, 5e-9, 10e-9)
    n = nx * ny
    g = boundary.compute_boundary_geometry(boundary.cylinder_hole_sdf(480e-9, 75e-9, 500e-9, 500e-9), nx, ny, 1,
                                           *cell, mode == "full")
    phi = g.volume_fraction
    mask = (phi > 0).astype(np.uint8)
    m0 = np.tile([1.0, 0.0, 0.0], (n, 1))
    m0[mask == 0] = 0.0
    gpu = fa.NativeFdmBackend(fa.FdmPlan(cells=(nx, ny, 1), cell_size=cell, initial_magnetization=m0,
                                         saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.5,
                                         active_mask=mask, demag_kernel="device", boundary_correction=mode,
                                         boundary_geometry=g))
    names = ("xp", "xm", "yp", "ym", "zp", "zm")
    faces = np.stack([getattr(g, f"face_link_{k}") for k in names])
    deltas = np.stack([getattr(g, f"delta_{k}") for k in names])
    st = g.demag_corr_stencil_size
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=1, dx=cell[0], dy=cell[1], dz=cell[2], ms=8e5, a=1.3e-11, alpha=0.5,
                         active_mask=mask, tier={"volume": 1, "full": 2}[mode], phi=phi, face_links=faces,
                         deltas=deltas, corr_target=g.demag_corr_target_idx,
                         corr_source=g.demag_corr_source_idx.reshape(-1, st),
                         corr_tensor=g.demag_corr_tensor.reshape(-1, st, 6))
    cpu = O.OracleProblem(cfg, m0)
    check_fields(gpu, cpu.observe(), 1e-11)
    for _ in range(30):
        gs = gpu.step(1e-13)
        r = cpu.step(1e-13)
    assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-12
    assert abs(gs.e_total - r["e_total"]) <= 1e-11 * abs(r["e_total"])
    gpu.close()
