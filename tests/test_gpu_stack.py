"""Multilayer stacks (fullmag_fdm_ext_stack_*) against the oracle's stack:
thin layers with a gap, 3-D layers (whose shifted kernels keep the
reference's z placement), a layer on a finer native grid (push/pull
transfer), FP64 to rounding, mixed and single precision to FP32 rounding;
a one-layer stack equals the single backend; layer handles refuse to be
stepped on their own."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DX, DY, DZ = 3e-9, 2.5e-9, 1e-9
H_EXT = (1.5e4, -5e3, 2e3)


def layer(fa, O, grid, cell, origin, ms, seed, mask=None, a=1.3e-11, alpha=0.2):
    n = grid[0] * grid[1] * grid[2]
    m0 = fa.generate_random_unit_vectors(seed, n)
    if mask is not None:
        m0[mask == 0] = 0.0
    lp = fa.FdmLayerPlan(magnet_name=f"L{seed}", native_grid=grid, native_cell_size=cell, native_origin=origin,
                         initial_magnetization=m0, saturation_magnetisation=ms, exchange_stiffness=a, damping=alpha,
                         native_active_mask=mask)
    cfg = O.OracleConfig(nx=grid[0], ny=grid[1], nz=grid[2], dx=cell[0], dy=cell[1], dz=cell[2], ms=ms, a=a,
                         alpha=alpha, demag=False, h_ext=H_EXT, active_mask=mask)
    return lp, O.OracleProblem(cfg, m0)


def build(fa, O, specs, common, conv_cell, precision="double"):
    lps, probs = zip(*[layer(fa, O, *sp) for sp in specs])
    plan = fa.FdmMultilayerPlan(common_cells=common, convolution_cell_size=conv_cell, layers=list(lps),
                                external_field=H_EXT, precision=precision)
    stack = fa.MultilayerStack(plan)
    cpu = O.OracleStack(list(probs), [lp.native_origin for lp in lps], common, conv_cell)
    return stack, cpu, probs


def check_fields(stack, o, tol):
    scale = max(np.abs(h).max() for h in o["h_eff"])
    for i, lay in enumerate(stack.layers):
        np.testing.assert_allclose(lay.copy_h_demag(), o["h_demag"][i], rtol=0, atol=tol * scale, err_msg=f"L{i}")
        np.testing.assert_allclose(lay.copy_h_ex(), o["h_ex"][i], rtol=0, atol=tol * scale, err_msg=f"L{i}")
        np.testing.assert_allclose(lay.copy_h_eff(), o["h_eff"][i], rtol=0, atol=tol * scale, err_msg=f"L{i}")


def run_and_compare(stack, cpu, probs, steps, dt, m_tol, e_tol):
    for _ in range(steps):
        g = stack.step(dt)
        r = cpu.step(dt)
    for lay, p in zip(stack.layers, probs):
        assert float(np.abs(lay.copy_m() - p.m).max()) <= m_tol
    for key, gv in (("e_ex", g.e_ex), ("e_demag", g.e_demag), ("e_ext", g.e_ext), ("e_total", g.e_total)):
        assert abs(gv - r[key]) <= e_tol * max(abs(r["e_total"]), abs(r[key])), key
    assert abs(g.max_h_eff - r["max_h_eff"]) <= e_tol * r["max_h_eff"]
    assert abs(g.time - r["time"]) <= 1e-12 * r["time"]


def thin_gap(fa, O, precision="double"):
    specs = [((32, 24, 1), (DX, DY, DZ), (0.0, 0.0, 0.0), 8e5, 1),
             ((32, 24, 1), (DX, DY, DZ), (0.0, 0.0, 3 * DZ), 1.1e6, 2),
             ((32, 24, 1), (DX, DY, DZ), (0.0, 0.0, 5 * DZ), 6e5, 3)]
    return build(fa, O, specs, (32, 24, 1), (DX, DY, DZ), precision)


def test_thin_layers_with_gaps_match_oracle(fa, oracle_mod):
    stack, cpu, probs = thin_gap(fa, oracle_mod)
    check_fields(stack, cpu.observe(), 1e-11)
    run_and_compare(stack, cpu, probs, 20, 1e-13, 1e-12, 1e-11)
    check_fields(stack, cpu.observe(), 1e-11)
    stack.close()


def test_three_d_layers_match_oracle(fa, oracle_mod):
    """Two 16x12x2 layers 3 dz apart on a 16x12x2 common grid: the shifted
    kernels' mirrored z slots keep slot k's sign, as the reference places
    them, on both sides."""
    specs = [((16, 12, 2), (DX, DY, DZ), (0.0, 0.0, 0.0), 8e5, 4),
             ((16, 12, 2), (DX, DY, DZ), (0.0, 0.0, 3 * DZ), 9e5, 5)]
    stack, cpu, probs = build(fa, oracle_mod, specs, (16, 12, 2), (DX, DY, DZ))
    check_fields(stack, cpu.observe(), 1e-11)
    run_and_compare(stack, cpu, probs, 10, 1e-13, 1e-12, 1e-11)
    stack.close()


def test_transfer_layer_matches_oracle(fa, oracle_mod):
    """A 32x24 layer of (dx, dy) cells over a 16x12 common grid of (2dx, 2dy)
    (push/pull), a masked 16x12 layer on the common grid itself."""
    mask = np.ones(16 * 12, np.uint8)
    mask[::5] = 0
    specs = [((32, 24, 1), (DX, DY, DZ), (0.0, 0.0, 0.0), 8e5, 6),
             ((16, 12, 1), (2 * DX, 2 * DY, DZ), (0.0, 0.0, 2 * DZ), 1e6, 7, mask)]
    stack, cpu, probs = build(fa, oracle_mod, specs, (16, 12, 1), (2 * DX, 2 * DY, DZ))
    check_fields(stack, cpu.observe(), 1e-11)
    run_and_compare(stack, cpu, probs, 10, 1e-13, 1e-12, 1e-11)
    assert np.all(stack.layers[1].copy_m()[mask == 0] == 0.0)
    stack.close()


@pytest.mark.parametrize("precision,field_tol,m_tol", [("mixed", 3e-6, 1e-6), ("single", 3e-6, 2e-5)])
def test_stack_reduced_precision_matches_oracle(fa, oracle_mod, precision, field_tol, m_tol):
    stack, cpu, probs = thin_gap(fa, oracle_mod, precision)
    check_fields(stack, cpu.observe(), field_tol)
    for _ in range(10):
        stack.step(1e-13)
        cpu.step(1e-13)
    for lay, p in zip(stack.layers, probs):
        assert float(np.abs(lay.copy_m() - p.m).max()) <= m_tol
    stack.close()


def test_one_layer_stack_is_the_single_backend(fa, oracle_mod):
    cells = (20, 16, 3)
    n = 20 * 16 * 3
    m0 = fa.generate_random_unit_vectors(9, n)
    lp = fa.FdmLayerPlan(magnet_name="only", native_grid=cells, native_cell_size=(DX, DY, DZ),
                         native_origin=(0.0, 0.0, 0.0), initial_magnetization=m0, saturation_magnetisation=8e5,
                         exchange_stiffness=1.3e-11, damping=0.2)
    stack = fa.MultilayerStack(fa.FdmMultilayerPlan(common_cells=cells, convolution_cell_size=(DX, DY, DZ),
                                                    layers=[lp], external_field=H_EXT))
    single = fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=(DX, DY, DZ), initial_magnetization=m0,
                                            saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.2,
                                            external_field=H_EXT, demag_kernel="device"))
    hd = single.copy_h_demag()
    np.testing.assert_allclose(stack.layers[0].copy_h_demag(), hd, rtol=0, atol=1e-12 * np.abs(hd).max())
    for _ in range(10):
        g = stack.step(1e-13)
        r = single.step(1e-13)
    assert float(np.abs(stack.layers[0].copy_m() - single.copy_m()).max()) <= 1e-12
    assert abs(g.e_total - r.e_total) <= 1e-11 * abs(r.e_total)
    stack.close()
    single.close()


def test_stack_layer_handles(fa, oracle_mod):
    stack, cpu, probs = thin_gap(fa, oracle_mod)
    lay = stack.layers[1]
    with pytest.raises(fa.RunError, match="layer of a multilayer stack"):
        lay.step(1e-13)
    with pytest.raises(fa.RunError, match="layer of a multilayer stack"):
        fa.BatchRun([lay], [1e-13])
    # an upload into one layer changes every layer's field: refresh, then step
    m_new = fa.generate_random_unit_vectors(77, 32 * 24)
    lay.upload_magnetization(m_new)
    probs[1].m = m_new
    stack.refresh()
    check_fields(stack, cpu.observe(), 1e-11)
    run_and_compare(stack, cpu, probs, 5, 1e-13, 1e-12, 1e-11)
    # an upload without refresh: the next step recomputes the fields
    lay.upload_magnetization(fa.generate_random_unit_vectors(78, 32 * 24))
    probs[1].m = fa.generate_random_unit_vectors(78, 32 * 24)
    run_and_compare(stack, cpu, probs, 3, 1e-13, 1e-12, 1e-11)
    stack.close()


def test_stack_rejections(fa):
    m0 = fa.generate_random_unit_vectors(1, 64)
    lp = fa.FdmLayerPlan(magnet_name="a", native_grid=(8, 8, 1), native_cell_size=(DX, DY, DZ),
                         native_origin=(0, 0, 0), initial_magnetization=m0, saturation_magnetisation=8e5,
                         exchange_stiffness=1.3e-11, damping=0.1)
    with pytest.raises(fa.RunError, match="positive"):
        fa.MultilayerStack(fa.FdmMultilayerPlan(common_cells=(8, 0, 1), convolution_cell_size=(DX, DY, DZ),
                                                layers=[lp]))
    with pytest.raises(fa.RunError, match="at least one layer"):
        fa.MultilayerStack(fa.FdmMultilayerPlan(common_cells=(8, 8, 1), convolution_cell_size=(DX, DY, DZ),
                                                layers=[]))


def test_five_layers_match_oracle(fa, oracle_mod):
    """Five thin layers (more than the register-held source count, so every
    destination reloads its sources), one of them 1.5 dz off the cell grid."""
    z = [0.0, 2 * DZ, 3.5 * DZ, 6 * DZ, 9 * DZ]
    specs = [((12, 10, 1), (DX, DY, DZ), (0.0, 0.0, zi), 7e5 + 1e5 * i, 20 + i) for i, zi in enumerate(z)]
    stack, cpu, probs = build(fa, oracle_mod, specs, (12, 10, 1), (DX, DY, DZ))
    check_fields(stack, cpu.observe(), 1e-11)
    run_and_compare(stack, cpu, probs, 10, 1e-13, 1e-12, 1e-11)
    stack.close()


def test_stack_without_demag_is_independent_layers(fa, oracle_mod):
    """enable_demag off: every layer steps exactly as its own backend with
    demag off (the same kernels, bit for bit)."""
    lps = [layer(fa, oracle_mod, (20, 12, 2), (DX, DY, DZ), (0.0, 0.0, 3 * DZ * i), 8e5 + 2e5 * i, 30 + i)[0]
           for i in range(2)]
    stack = fa.MultilayerStack(fa.FdmMultilayerPlan(common_cells=(20, 12, 2), convolution_cell_size=(DX, DY, DZ),
                                                    layers=lps, external_field=H_EXT, enable_demag=False))
    singles = [fa.NativeFdmBackend(fa.FdmPlan(cells=lp.native_grid, cell_size=lp.native_cell_size,
                                              initial_magnetization=lp.initial_magnetization,
                                              saturation_magnetisation=lp.saturation_magnetisation,
                                              exchange_stiffness=lp.exchange_stiffness, damping=lp.damping,
                                              external_field=H_EXT, enable_demag=False)) for lp in lps]
    for _ in range(6):
        g = stack.step(1e-13)
        rs = [b.step(1e-13) for b in singles]
    for lay, b in zip(stack.layers, singles):
        np.testing.assert_array_equal(lay.copy_m(), b.copy_m())
        assert np.all(lay.copy_h_demag() == 0.0)
    assert g.e_total == rs[0].e_total + rs[1].e_total
    stack.close()
    for b in singles:
        b.close()
