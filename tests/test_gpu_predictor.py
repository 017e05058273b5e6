"""Heun's first slope and predictor are stored by the stats pass of the
previous step (same m, same H, same dt) and reused by the next step.  The
reuse must be bit-exact against recomputing the first stage, and every event
that changes m, H or dt must drop the stored predictor."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H_EXT = (1.2e4, -3e3, 5e3)


def backend(fa, cells, prec, seed, integrator="heun"):
    n = cells[0] * cells[1] * cells[2]
    m0 = fa.generate_random_unit_vectors(seed, n)
    return fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=(4e-9, 4e-9, 2e-9), initial_magnetization=m0,
                                          saturation_magnetisation=8.6e5, exchange_stiffness=1.3e-11, damping=0.05,
                                          external_field=H_EXT, precision=prec, integrator=integrator,
                                          demag_kernel="device"))


@pytest.mark.parametrize("prec", ["double", "mixed", "single"])
@pytest.mark.parametrize("cells", [(64, 48, 4), (40, 24, 1)])
def test_predictor_reuse_is_bit_exact(fa, prec, cells):
    a = backend(fa, cells, prec, 11)
    b = backend(fa, cells, prec, 11)
    for _ in range(6):
        sa = a.step(5e-14)
        b.set_external_field(H_EXT)  # same field: only drops b's stored predictor
        sb = b.step(5e-14)
        np.testing.assert_array_equal(a.copy_m(), b.copy_m())
        assert sa.e_total == sb.e_total and sa.max_dm_dt == sb.max_dm_dt
    sa = a.step_many(5e-14, 5)
    b.set_external_field(H_EXT)
    sb = b.step_many(5e-14, 5)
    np.testing.assert_array_equal(a.copy_m(), b.copy_m())
    a.close()
    b.close()


@pytest.mark.parametrize("prec", ["double", "mixed"])
@pytest.mark.parametrize("kind", ["masked", "tier_volume", "tier_full"])
def test_predictor_reuse_is_bit_exact_masked_and_tiers(fa, oracle_mod, prec, kind):
    """The same on a masked cylinder, without and with the sub-cell boundary
    tiers (the stats pass and the plain stage are then the uniform stage
    kernel's masked and tier forms)."""
    from tests import _tiers

    cells = (24, 20, 3)
    nx, ny, nz = cells
    n = nx * ny * nz
    phi, faces, deltas = _tiers.geometry(nx, ny, nz, 3e-9, 4e-9, 2e-9)
    mask = (phi > 0).astype(np.uint8)
    m0 = fa.generate_random_unit_vectors(17, n)
    m0[mask == 0] = 0.0
    names = ("xp", "xm", "yp", "ym", "zp", "zm")
    kw = {}
    if kind != "masked":
        kw = dict(boundary_correction="volume" if kind == "tier_volume" else "full",
                  boundary_geometry=fa.BoundaryGeometry(
                      volume_fraction=phi, **{f"face_link_{k}": faces[i] for i, k in enumerate(names)},
                      **{f"delta_{k}": deltas[i] for i, k in enumerate(names)}))

    def make():
        return fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=(3e-9, 4e-9, 2e-9), initial_magnetization=m0,
                                              saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.2,
                                              external_field=H_EXT, precision=prec, active_mask=mask,
                                              demag_kernel="device", **kw))

    a, b = make(), make()
    for _ in range(5):
        sa = a.step(5e-14)
        b.set_external_field(H_EXT)
        sb = b.step(5e-14)
        np.testing.assert_array_equal(a.copy_m(), b.copy_m())
        assert sa.e_total == sb.e_total
    a.step(5e-14)
    b.step_many(5e-14, 1)
    np.testing.assert_array_equal(a.copy_m(), b.copy_m())
    a.close()
    b.close()


def test_predictor_dropped_on_dt_field_and_state_changes(fa, oracle_mod):
    cells = (24, 20, 3)
    nx, ny, nz = cells
    n = nx * ny * nz
    gpu = backend(fa, cells, "double", 5)
    cfg = oracle_mod.OracleConfig(nx=nx, ny=ny, nz=nz, dx=4e-9, dy=4e-9, dz=2e-9, ms=8.6e5, a=1.3e-11, alpha=0.05,
                                  h_ext=H_EXT)
    cpu = oracle_mod.OracleProblem(cfg, fa.generate_random_unit_vectors(5, n))
    rng = np.random.default_rng(3)
    for i, dt in enumerate([5e-14, 5e-14, 2e-14, 2e-14, 5e-14, 5e-14, 5e-14, 5e-14]):
        if i == 4:
            h = (8e3, 2e3, -1e3)
            gpu.set_external_field(h)
            cpu = oracle_mod.OracleProblem(oracle_mod.OracleConfig(nx=nx, ny=ny, nz=nz, dx=4e-9, dy=4e-9, dz=2e-9,
                                                                   ms=8.6e5, a=1.3e-11, alpha=0.05, h_ext=h),
                                           cpu.m)
        if i == 6:
            m = rng.standard_normal((n, 3))
            m /= np.linalg.norm(m, axis=1, keepdims=True)
            gpu.upload_magnetization(m)
            cpu.set_m_raw(m)
        gpu.step(dt)
        cpu.step(dt)
        assert float(np.abs(gpu.copy_m() - cpu.m).max()) <= 1e-12, i
    gpu.close()


@pytest.mark.parametrize("prec", ["double", "mixed"])
@pytest.mark.parametrize("cells", [(48, 40, 4), (40, 24, 1)])
def test_next_predictor_transforms_survive_interleaved_calls(fa, prec, cells):
    """A Heun step enqueues the next predictor's demag transforms behind its
    stats and returns once the stats are in.  Observable copies, previews,
    snapshots, a refresh, step_many and a dt change between steps must see the
    fields of the step just taken, and every later step must be bit-identical
    to a backend whose stored predictor is dropped before each step."""
    a = backend(fa, cells, prec, 23)
    b = backend(fa, cells, prec, 23)

    def snap(x, name):
        s = x.begin_field_snapshot(name)
        out = s.wait()
        s.close()
        return out

    ops = [lambda x: x.copy_h_demag(), lambda x: x.copy_h_eff(), lambda x: x.copy_h_ex(),
           lambda x: x.copy_live_preview_field("H_demag", (8, 8, 1)), lambda x: snap(x, "H_demag"),
           lambda x: snap(x, "m"), None, lambda x: x.refresh_observables(), None, lambda x: x.copy_m()]
    dts = [5e-14, 5e-14, 5e-14, 4e-14, 4e-14, 5e-14, 5e-14, 5e-14, 5e-14, 5e-14]
    for i, (op, dt) in enumerate(zip(ops, dts)):
        if op is not None:
            ra, rb = op(a), op(b)
            if ra is not None:
                np.testing.assert_array_equal(ra, rb, err_msg=str(i))
        b.set_external_field(H_EXT)  # same field: drops b's stored predictor
        sa, sb = a.step(dt), b.step(dt)
        np.testing.assert_array_equal(a.copy_m(), b.copy_m(), err_msg=str(i))
        assert sa.e_total == sb.e_total and sa.max_h_demag == sb.max_h_demag, i
        if i == 5:
            sa, sb = a.step_many(5e-14, 3), b.step_many(5e-14, 3)
            np.testing.assert_array_equal(a.copy_m(), b.copy_m())
    # the fields after the last step equal a fresh evaluation at that m
    hd = a.copy_h_demag()
    c = backend(fa, cells, prec, 23)
    c.upload_magnetization(a.copy_m())
    c.refresh_observables()
    np.testing.assert_array_equal(hd, c.copy_h_demag())
    for x in (a, b, c):
        x.close()
