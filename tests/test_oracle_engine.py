"""Oracle engine (fields, energies, integrators, m0) against the reference's
unit tests (crates/fullmag-engine/src/lib.rs:2874-3290) and closed forms."""

import json
import math
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

MU0 = 4.0 * math.pi * 1e-7


def cfg(O, **kw):
    base = dict(nx=3, ny=1, nz=1, dx=1.0, dy=1.0, dz=1.0, ms=1.0, a=0.5 * MU0, alpha=0.1,
                gamma=1.0, exchange=True, demag=False)
    base.update(kw)
    return O.OracleConfig(**base)


def test_uniform_state_has_zero_exchange_field_and_rhs(oracle_mod):
    # lib.rs:2984-3004
    p = oracle_mod.OracleProblem(cfg(oracle_mod), np.tile([1.0, 0, 0], (3, 1)))
    o = p.observe()
    assert np.abs(o["h_ex"]).max() <= 1e-12
    assert np.abs(o["rhs"]).max() <= 1e-12
    assert o["e_ex"] <= 1e-12


def test_center_exchange_field_matches_second_difference(oracle_mod, kats):
    # lib.rs:3006-3017: A = mu0/2, Ms = 1, dx = 1 -> H = Laplacian(m)
    k = kats["exchange"]["stencil_center"]
    p = oracle_mod.OracleProblem(cfg(oracle_mod, alpha=0.0), np.array(k["m"], float))
    np.testing.assert_allclose(p.observe()["h_ex"][1], k["expect"], atol=1e-12)


def test_masked_exchange_inactive_neighbour_is_free_surface(oracle_mod, kats):
    # lib.rs:3019-3034
    k = kats["exchange"]["masked"]
    c = cfg(oracle_mod, active_mask=np.array(k["mask"], np.uint8))
    p = oracle_mod.OracleProblem(c, np.array(k["m"], float))
    o = p.observe()
    np.testing.assert_allclose(o["h_ex"][1], k["expect1"], atol=1e-12)
    np.testing.assert_allclose(o["h_ex"][2], k["expect2"], atol=1e-12)
    np.testing.assert_allclose(p.m[2], [0, 0, 0], atol=1e-12)


def test_masked_demag_and_external_zero_outside_domain(oracle_mod):
    # lib.rs:3036-3049
    c = cfg(oracle_mod, exchange=False, demag=True, h_ext=(0.0, 0.0, 1.0),
            active_mask=np.array([1, 1, 0], np.uint8))
    p = oracle_mod.OracleProblem(c, np.array([[1, 0, 0], [1, 0, 0], [0, 1, 0]], float))
    o = p.observe()
    for key in ("h_ext", "h_demag", "h_eff"):
        np.testing.assert_allclose(o[key][2], 0.0, atol=1e-12)


def test_heun_preserves_unit_norm_and_damping_lowers_energy(oracle_mod):
    # lib.rs:3051-3080
    p = oracle_mod.OracleProblem(cfg(oracle_mod, alpha=0.5), np.array([[1, 0, 0], [0, 1, 0], [1, 0, 0]], float))
    e0 = p.observe()["e_ex"]
    p.step(1e-3)
    np.testing.assert_allclose(np.linalg.norm(p.m, axis=1), 1.0, atol=1e-12)
    for _ in range(9):
        p.step(1e-3)
    assert p.observe()["e_total"] < e0


def test_relax_mode_has_no_precession(oracle_mod):
    # lib.rs:3110-3132
    c = cfg(oracle_mod, nx=2, exchange=False, alpha=0.5, h_ext=(0.0, 0.0, 1.0), precession=False)
    p = oracle_mod.OracleProblem(c, np.array([[1, 0, 0], [1, 0, 0]], float))
    p.step(1e-3)
    assert abs(p.m[0][1]) <= 1e-12
    assert p.m[0][2] > 0.0


def test_thin_film_demag_ordering_and_nonnegative_energy(oracle_mod):
    # lib.rs:3134-3185 and 3262-3289
    c = cfg(oracle_mod, nx=8, ny=8, nz=1, dz=0.2, exchange=False, demag=True)
    e_in = oracle_mod.OracleProblem(c, np.tile([1.0, 0, 0], (64, 1))).observe()["e_demag"]
    e_out = oracle_mod.OracleProblem(c, np.tile([0, 0, 1.0], (64, 1))).observe()["e_demag"]
    assert e_out > e_in and e_in < 0.5 * e_out
    c2 = cfg(oracle_mod, nx=4, ny=4, nz=2, dz=0.2, exchange=False, demag=True)
    m = oracle_mod.random_unit_vectors(42, 32)
    e = oracle_mod.OracleProblem(c2, m).observe()["e_demag"]
    assert e >= 0.0 and np.isfinite(e)


def test_single_cube_demag_is_minus_m_over_three(oracle_mod):
    # Closed form: a uniformly magnetised cube has H = -M/3 at its centre
    # (self term N = I/3); the padded FFT path must reproduce it.
    ms = 8e5
    c = cfg(oracle_mod, nx=1, ny=1, nz=1, dx=2e-9, dy=2e-9, dz=2e-9, ms=ms, a=1e-11,
            exchange=False, demag=True)
    m = np.array([[0.6, 0.0, 0.8]])
    h = oracle_mod.OracleProblem(c, m).observe()["h_demag"][0]
    np.testing.assert_allclose(h, -ms * m[0] / 3.0, rtol=1e-9, atol=1e-6)


def test_extended_film_out_of_plane_field_approaches_minus_m(oracle_mod):
    # Closed form: an infinite film has H_z = -M_z; a 48x48x1 film of 10:1
    # aspect cells is within a few percent of it at the centre.
    ms = 1e6
    c = cfg(oracle_mod, nx=48, ny=48, nz=1, dx=10e-9, dy=10e-9, dz=1e-9, ms=ms, a=1e-11,
            exchange=False, demag=True)
    p = oracle_mod.OracleProblem(c, np.tile([0, 0, 1.0], (48 * 48, 1)))
    hz = p.observe()["h_demag"].reshape(48, 48, 3)[24, 24, 2]
    assert -1.0 < hz / ms < -0.95


def test_exchange_energy_summation_by_parts(oracle_mod):
    # E_ex = -mu0 Ms V / 2 sum m.H_ex equals the forward-pair sum
    # sum A V |m_j - m_i|^2 / d^2 (the reference GPU formula,
    # native/backends/fdm/src/reductions_fp64.cu:123-214), mask included.
    nx, ny, nz = 5, 4, 3
    rng = np.random.default_rng(3)
    mask = (rng.random(nx * ny * nz) > 0.2).astype(np.uint8)
    c = cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=2e-9, dy=3e-9, dz=4e-9, ms=8e5, a=1.3e-11,
            active_mask=mask)
    m = oracle_mod.random_unit_vectors(9, nx * ny * nz)
    p = oracle_mod.OracleProblem(c, m)
    e_field = p.observe()["e_ex"]
    mm = p.m.reshape(nz, ny, nx, 3)
    act = mask.reshape(nz, ny, nx).astype(bool)
    v = 2e-9 * 3e-9 * 4e-9
    e_pair = 0.0
    for axis, d in ((2, 2e-9), (1, 3e-9), (0, 4e-9)):
        a = np.take(mm, range(0, mm.shape[axis] - 1), axis=axis)
        b = np.take(mm, range(1, mm.shape[axis]), axis=axis)
        ok = np.take(act, range(0, act.shape[axis] - 1), axis=axis) & np.take(act, range(1, act.shape[axis]), axis=axis)
        e_pair += 1.3e-11 * v * np.sum(np.sum((b - a) ** 2, axis=-1)[ok]) / d ** 2
    assert e_field == pytest.approx(e_pair, rel=1e-12)


def analytic_spin(theta0, phi0, gamma, alpha, h, t):
    """Closed-form LLG for one spin in a static field H z (precession +
    Gilbert damping): tan(theta/2) decays as exp(-gb a H t), phi = phi0 + gb H t."""
    gb = gamma / (1.0 + alpha * alpha)
    th = 2.0 * math.atan(math.tan(theta0 / 2.0) * math.exp(-gb * alpha * h * t))
    ph = phi0 + gb * h * t
    return np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])


def spin_error(O, integrator, dt, steps, **extra):
    h, gamma, alpha = 1.0, 1.0, 0.1
    th0, ph0 = 1.0, 0.3
    m0 = np.array([[math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]])
    c = cfg(O, nx=1, ny=1, nz=1, exchange=False, alpha=alpha, gamma=gamma, h_ext=(0.0, 0.0, h),
            integrator=integrator, **extra)
    p = O.OracleProblem(c, m0)
    t = 0.0
    for _ in range(steps):
        r = p.step(dt)
        t = r["time"]
    return np.abs(p.m[0] - analytic_spin(th0, ph0, gamma, alpha, h, t)).max()


@pytest.mark.parametrize("integrator,order", [("heun", 2), ("rk4", 4), ("abm3", 3)])
def test_fixed_step_integrators_converge_at_formal_order(oracle_mod, integrator, order):
    e1 = spin_error(oracle_mod, integrator, 0.02, 100)
    e2 = spin_error(oracle_mod, integrator, 0.01, 200)
    rate = math.log2(e1 / e2)
    assert rate == pytest.approx(order, abs=0.35)


@pytest.mark.parametrize("integrator", ["dp45", "rk23"])
def test_adaptive_integrators_track_analytic_solution(oracle_mod, integrator):
    err = spin_error(oracle_mod, integrator, 0.05, 40, max_error=1e-9, dt_min=1e-9, dt_max=0.05, headroom=0.8)
    assert err < 1e-6


def test_dp45_reports_suggested_dt_and_fsal(oracle_mod):
    c = cfg(oracle_mod, nx=1, ny=1, nz=1, exchange=False, alpha=0.1, h_ext=(0.0, 0.0, 1.0),
            integrator="dp45", max_error=1e-7, dt_min=1e-9, dt_max=1.0, headroom=0.8)
    p = oracle_mod.OracleProblem(c, np.array([[1.0, 0.0, 0.0]]))
    r = p.step(0.1)
    assert r["suggested_next_dt"] > 0.0
    assert r["dt_used"] <= 0.1


def test_random_unit_vectors_bit_exact(oracle_mod):
    # crates/fullmag-plan/src/lib.rs:1927-1957 against an independent
    # arbitrary-precision restatement of the xorshift64 recurrence.
    with open(os.path.join(GOLDEN, "xorshift_states.json")) as f:
        fx = json.load(f)
    v = oracle_mod.random_unit_vectors(fx["seed"], 4)
    np.testing.assert_array_equal(v, np.array(fx["unit_vectors"]))
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
