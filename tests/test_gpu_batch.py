"""Graph-captured batches (fullmag_fdm_ext_batch_*) and the launch-schedule
autotuner.  A batch replays captured Heun steps; every member must end
bit-identical to the same backend stepped eagerly, over several replays with
eager work in between, and the autotuned schedule must still match the
oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEMBERS = [((32, 32, 1), 5e-14, "mixed"), ((64, 32, 2), 1e-13, "mixed"), ((24, 20, 3), 8e-14, "double"),
           ((16, 16, 8), 1e-13, "mixed")]


def member(fa, cells, prec, seed):
    n = cells[0] * cells[1] * cells[2]
    m0 = fa.generate_random_unit_vectors(seed, n)
    return fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=(5e-9, 5e-9, 3e-9), initial_magnetization=m0,
                                          saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1,
                                          external_field=(1e4, 2e3, 0.0), precision=prec, demag_kernel="device"))


def test_batch_graph_matches_eager_stepping(fa):
    batch_members = [member(fa, c, p, 40 + i) for i, (c, _, p) in enumerate(MEMBERS)]
    eager = [member(fa, c, p, 40 + i) for i, (c, _, p) in enumerate(MEMBERS)]
    dts = [dt for _, dt, _ in MEMBERS]
    run = fa.BatchRun(batch_members, dts, steps_per_graph=4)
    total = 0
    for n_steps in (9, 4, 13):  # graph replays plus an eager remainder, three times
        stats = run.run(n_steps)
        total += n_steps
        for b, e, dt, s in zip(batch_members, eager, dts, stats):
            for _ in range(n_steps):
                r = e.step(dt)
            np.testing.assert_array_equal(b.copy_m(), e.copy_m())
            assert s.step == total and s.e_total == r.e_total
            assert abs(s.time - r.time) <= 1e-12 * r.time
        # eager work on the members between replays
        for b in batch_members:
            b.copy_h_demag()
    run.close()
    for b in batch_members + eager:
        b.close()


def test_batch_rejects_bad_members(fa):
    b = member(fa, (8, 8, 1), "mixed", 3)
    with pytest.raises(fa.RunError):
        fa.BatchRun([b, b], [1e-13, 1e-13])
    with pytest.raises(fa.RunError):
        fa.BatchRun([b], [0.0])
    b.close()


@pytest.mark.parametrize("cells", [(64, 32, 8), (48, 40, 1)])
def test_autotuned_schedule_matches_oracle(fa, oracle_mod, cells):
    nx, ny, nz = cells
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(7 + n, n)
    b = fa.NativeFdmBackend(fa.FdmPlan(cells=cells, cell_size=(4e-9, 4e-9, 2e-9), initial_magnetization=m0,
                                       saturation_magnetisation=8.6e5, exchange_stiffness=1.3e-11, damping=0.05,
                                       precision="double", demag_kernel="device"))
    chosen = b.autotune(reps=2)
    assert "y_pipe=" in chosen and "fixed=" in chosen
    cfg = oracle_mod.OracleConfig(nx=nx, ny=ny, nz=nz, dx=4e-9, dy=4e-9, dz=2e-9, ms=8.6e5, a=1.3e-11, alpha=0.05)
    cpu = oracle_mod.OracleProblem(cfg, m0)
    hd = cpu.observe()["h_demag"]
    # other FFT factorisations round differently: 1e-11 of max|H| (the
    # trajectory bar is max|dm| <= 1e-10)
    np.testing.assert_allclose(b.copy_h_demag(), hd, rtol=0, atol=1e-11 * np.abs(hd).max(), err_msg=chosen)
    for _ in range(3):
        b.step(5e-14)
        cpu.step(5e-14)
    assert float(np.abs(b.copy_m() - cpu.m).max()) <= 1e-12, chosen
    b.close()


def test_batch_follows_member_changes_between_runs(fa):
    """An upload of m and a new external field on a member between two runs:
    the batch recomputes H_demag at the uploaded m and recaptures the graph
    with the new field, so members stay bit-identical to eager stepping."""
    specs = MEMBERS[:3]
    batch_members = [member(fa, c, p, 60 + i) for i, (c, _, p) in enumerate(specs)]
    eager = [member(fa, c, p, 60 + i) for i, (c, _, p) in enumerate(specs)]
    dts = [dt for _, dt, _ in specs]
    run = fa.BatchRun(batch_members, dts, steps_per_graph=3)

    def advance(n):
        run.run(n)
        for e, dt in zip(eager, dts):
            for _ in range(n):
                e.step(dt)
        for b, e in zip(batch_members, eager):
            np.testing.assert_array_equal(b.copy_m(), e.copy_m())

    advance(6)
    n0 = specs[0][0][0] * specs[0][0][1] * specs[0][0][2]
    m_new = fa.generate_random_unit_vectors(999, n0)
    for x in (batch_members[0], eager[0]):
        x.upload_magnetization(m_new)
    for x in (batch_members[1], eager[1]):
        x.set_external_field((-3e4, 1e4, 5e3))
    advance(7)
    np.testing.assert_array_equal(batch_members[1].copy_h_ext(), eager[1].copy_h_ext())
    run.close()
    for b in batch_members + eager:
        b.close()


def test_schedule_replan_knobs_match_oracle(fa, oracle_mod):
    """set_schedule with plan-shaping knobs (radices, rows per x workgroup,
    y lines per workgroup) re-plans the transforms in place; the field stays
    the oracle's, and the autotuner reports those knobs."""
    nx, ny, nz = 48, 40, 4
    n = nx * ny * nz
    m0 = fa.generate_random_unit_vectors(12, n)
    b = fa.NativeFdmBackend(fa.FdmPlan(cells=(nx, ny, nz), cell_size=(4e-9, 4e-9, 2e-9), initial_magnetization=m0,
                                       saturation_magnetisation=8.6e5, exchange_stiffness=1.3e-11, damping=0.05,
                                       precision="double", demag_kernel="device"))
    cfg = oracle_mod.OracleConfig(nx=nx, ny=ny, nz=nz, dx=4e-9, dy=4e-9, dz=2e-9, ms=8.6e5, a=1.3e-11, alpha=0.05)
    hd = oracle_mod.OracleProblem(cfg, m0).observe()["h_demag"]
    for sched in ("x_radix=4,y_radix=8", "x_rows=2,y_lpw=1", "x_radix=8,fixed=0", "x_radix=16,y_radix=16"):
        b.set_schedule(sched)
        b.refresh_observables()
        np.testing.assert_allclose(b.copy_h_demag(), hd, rtol=0, atol=1e-11 * np.abs(hd).max(), err_msg=sched)
    chosen = b.autotune(reps=1)
    for key in ("x_radix=", "y_radix=", "x_rows=", "y_lpw="):
        assert key in chosen
    b.refresh_observables()
    np.testing.assert_allclose(b.copy_h_demag(), hd, rtol=0, atol=1e-11 * np.abs(hd).max(), err_msg=chosen)
    b.close()


def test_batch_classes_of_equal_grids_match_eager(fa):
    """Members sharing a grid, precision and schedule run each pass as one
    multi-instance launch (arguments per member from a device table): three
    40x24x2 mixed members with different Ms, alpha, dt and fields, two
    32x32x1 FP64 films, two 16x16x8 mixed blocks, and a masked 40x24x2
    member that keeps its own launches.  Every member ends bit-identical to
    eager stepping."""
    rng = np.random.default_rng(9)

    def make(cells, prec, seed, dt_scale, mask=False):
        n = cells[0] * cells[1] * cells[2]
        kw = {}
        if mask:
            am = np.ones(n, np.uint8)
            am[::7] = 0
            kw["active_mask"] = am
        plan = fa.FdmPlan(cells=cells, cell_size=(5e-9, 4e-9, 3e-9),
                          initial_magnetization=fa.generate_random_unit_vectors(seed, n),
                          saturation_magnetisation=float(rng.uniform(4e5, 1.4e6)),
                          exchange_stiffness=float(rng.uniform(5e-12, 2e-11)), damping=float(rng.uniform(0.01, 0.5)),
                          external_field=tuple(rng.uniform(-2e4, 2e4, 3)), precision=prec, demag_kernel="device",
                          **kw)
        return plan, 1e-13 * dt_scale

    specs = [((40, 24, 2), "mixed", 1.0, False), ((40, 24, 2), "mixed", 0.7, False), ((40, 24, 2), "mixed", 0.4, False),
             ((32, 32, 1), "double", 1.0, False), ((32, 32, 1), "double", 0.5, False),
             ((16, 16, 8), "mixed", 0.8, False), ((16, 16, 8), "mixed", 1.0, False), ((40, 24, 2), "mixed", 1.0, True)]
    plans = [make(c, p, 100 + i, s, m) for i, (c, p, s, m) in enumerate(specs)]
    members = [fa.NativeFdmBackend(p) for p, _ in plans]
    eager = [fa.NativeFdmBackend(p) for p, _ in plans]
    dts = [dt for _, dt in plans]
    run = fa.BatchRun(members, dts, steps_per_graph=3)
    total = 0
    for n_steps in (7, 6):
        stats = run.run(n_steps)
        total += n_steps
        for b, e, dt, s in zip(members, eager, dts, stats):
            for _ in range(n_steps):
                r = e.step(dt)
            np.testing.assert_array_equal(b.copy_m(), e.copy_m())
            assert s.step == total and s.e_total == r.e_total
    run.close()
    for b in members + eager:
        b.close()
