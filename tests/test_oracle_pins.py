"""Independent pins of the oracle parts that the reference's own known-answer
tests leave open: the multi-cell demag convolution (real-space brute force
and the analytic demagnetising factors of a rectangular prism), the energy
prefactors, the RK23/DP45 error estimators and step-size controller, and the
per-pair exchange table of native/backends/fdm/src/exchange_fp64.cu:80-114.

Each check is chosen so that a plausible mistake fails it: a swapped
xz/yz entry or an x/y index swap in the padded convolution changes the local
fields of a random state; a dropped or doubled energy factor misses the
closed forms; a wrong error-estimator coefficient drops the observed order of
the estimate; a wrong controller exponent misses the formula.  Two tests
plant such mistakes in the compared data and confirm the checks reject them.
"""

import itertools
import math

import numpy as np
import pytest

MU0 = 4.0 * math.pi * 1e-7


def cfg(O, **kw):
    base = dict(nx=1, ny=1, nz=1, dx=1e-9, dy=1e-9, dz=1e-9, ms=8e5, a=1.3e-11, alpha=0.1,
                exchange=False, demag=True)
    base.update(kw)
    return O.OracleConfig(**base)


# ---------------------------------------------------------------------------
# Demag: real-space brute force
# ---------------------------------------------------------------------------

def octant_tensor(O, nx, ny, nz, dx, dy, dz):
    """N_c(|d|) for non-negative displacements, from the KAT-pinned kernel
    builder (only its non-negative octant is used)."""
    k = O.newell_kernels(nx, ny, nz, dx, dy, dz)  # (6, 2nz, 2ny, 2nx)
    return k[:, :nz, :ny, :nx]


def brute_force_demag(O, nx, ny, nz, cell, ms, m, mask=None):
    """H_i = -sum_j N(r_i - r_j) . (Ms m_j), with N at negative displacements
    from the octant by the tensor's parity: xx, yy, zz even in every axis,
    xy odd in x and y, xz odd in x and z, yz odd in y and z."""
    oc = octant_tensor(O, nx, ny, nz, *cell)
    n = nx * ny * nz
    coords = [(x, y, z) for z in range(nz) for y in range(ny) for x in range(nx)]
    act = np.ones(n, bool) if mask is None else mask.astype(bool)
    mm = np.where(act[:, None], ms * m, 0.0)
    h = np.zeros((n, 3))
    for i, (xi, yi, zi) in enumerate(coords):
        if not act[i]:
            continue
        acc = np.zeros(3)
        for j, (xj, yj, zj) in enumerate(coords):
            ddx, ddy, ddz = xi - xj, yi - yj, zi - zj
            sx = -1.0 if ddx < 0 else 1.0
            sy = -1.0 if ddy < 0 else 1.0
            sz = -1.0 if ddz < 0 else 1.0
            e = oc[:, abs(ddz), abs(ddy), abs(ddx)]
            xx, yy, zz = e[0], e[1], e[2]
            xy, xz, yz = sx * sy * e[3], sx * sz * e[4], sy * sz * e[5]
            t = np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])
            acc -= t @ mm[j]
        h[i] = acc
    return h


GRIDS = [((3, 2, 2), (2e-9, 3e-9, 5e-9)), ((4, 3, 1), (5e-9, 3e-9, 2e-9)), ((2, 2, 3), (3e-9, 2e-9, 4e-9)),
         ((5, 1, 2), (2e-9, 7e-9, 3e-9))]


@pytest.mark.parametrize("grid,cell", GRIDS)
def test_demag_convolution_equals_real_space_sum(oracle_mod, grid, cell):
    nx, ny, nz = grid
    n = nx * ny * nz
    ms = 8.6e5
    m = oracle_mod.random_unit_vectors(101 + n, n)
    p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], ms=ms), m)
    h_fft = p.observe()["h_demag"]
    h_bf = brute_force_demag(oracle_mod, nx, ny, nz, cell, ms, m)
    assert np.abs(h_fft - h_bf).max() <= 1e-12 * np.abs(h_bf).max()


def test_demag_convolution_equals_real_space_sum_masked(oracle_mod):
    nx, ny, nz = 4, 3, 2
    cell = (3e-9, 2e-9, 5e-9)
    n = nx * ny * nz
    mask = np.ones(n, np.uint8)
    mask[[1, 6, 7, 17, 22]] = 0
    m = oracle_mod.random_unit_vectors(77, n)
    p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2],
                                     active_mask=mask), m)
    h_fft = p.observe()["h_demag"]
    h_bf = brute_force_demag(oracle_mod, nx, ny, nz, cell, 8e5, m, mask)
    assert np.all(h_fft[mask == 0] == 0.0)
    assert np.abs(h_fft - h_bf).max() <= 1e-12 * np.abs(h_bf).max()


def test_brute_force_check_rejects_planted_index_mistakes(oracle_mod):
    """The brute-force comparison notices an xz/yz swap and an x/y swap of the
    field (the mistakes a pack/unpack or tensor-product slip would make)."""
    nx, ny, nz = 3, 2, 2
    cell = (2e-9, 3e-9, 5e-9)
    m = oracle_mod.random_unit_vectors(113, nx * ny * nz)
    p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2]), m)
    h = p.observe()["h_demag"]
    h_bf = brute_force_demag(oracle_mod, nx, ny, nz, cell, 8e5, m)
    tol = 1e-12 * np.abs(h_bf).max()
    # xz and yz entries swapped in the tensor product
    oc = octant_tensor(oracle_mod, nx, ny, nz, *cell)
    swapped = oc.copy()
    swapped[[4, 5]] = swapped[[5, 4]]
    assert np.abs(swapped[4] - oc[4]).max() > 0.0
    h_sw = h.copy()
    h_sw[:, [0, 1]] = h_sw[:, [1, 0]]  # x/y components exchanged on output
    assert np.abs(h_sw - h_bf).max() > tol
    # the tensor with xz/yz swapped gives a different field on a random state
    coords = [(x, y, z) for z in range(nz) for y in range(ny) for x in range(nx)]
    hw = np.zeros_like(h)
    for i, (xi, yi, zi) in enumerate(coords):
        for j, (xj, yj, zj) in enumerate(coords):
            d = (xi - xj, yi - yj, zi - zj)
            s = [(-1.0 if v < 0 else 1.0) for v in d]
            e = swapped[:, abs(d[2]), abs(d[1]), abs(d[0])]
            t = np.array([[e[0], s[0] * s[1] * e[3], s[0] * s[2] * e[4]],
                          [s[0] * s[1] * e[3], e[1], s[1] * s[2] * e[5]],
                          [s[0] * s[2] * e[4], s[1] * s[2] * e[5], e[2]]])
            hw[i] -= t @ (8e5 * m[j])
    assert np.abs(hw - h).max() > tol


# ---------------------------------------------------------------------------
# Demag: closed forms
# ---------------------------------------------------------------------------

def aharoni_dz(a, b, c):
    """Magnetometric demagnetising factor along z of a rectangular prism of
    edges 2a x 2b x 2c (A. Aharoni, J. Appl. Phys. 83, 3432 (1998), eq. 1)."""
    r = math.sqrt(a * a + b * b + c * c)
    rab, rbc, rac = math.sqrt(a * a + b * b), math.sqrt(b * b + c * c), math.sqrt(a * a + c * c)
    t = ((b * b - c * c) / (2 * b * c) * math.log((r - a) / (r + a))
         + (a * a - c * c) / (2 * a * c) * math.log((r - b) / (r + b))
         + b / (2 * c) * math.log((rab + a) / (rab - a))
         + a / (2 * c) * math.log((rab + b) / (rab - b))
         + c / (2 * a) * math.log((rbc - b) / (rbc + b))
         + c / (2 * b) * math.log((rac - a) / (rac + a))
         + 2 * math.atan(a * b / (c * r))
         + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c)
         + (a * a + b * b - 2 * c * c) / (3 * a * b * c) * r
         + c / (a * b) * (rac + rbc)
         - (rab ** 3 + rbc ** 3 + rac ** 3) / (3 * a * b * c))
    return t / math.pi


def prism_factors(lx, ly, lz):
    a, b, c = lx / 2, ly / 2, lz / 2
    return aharoni_dz(b, c, a), aharoni_dz(c, a, b), aharoni_dz(a, b, c)


def test_aharoni_factors_closed_form_checks():
    assert prism_factors(1, 1, 1) == pytest.approx((1 / 3,) * 3, abs=1e-14)
    for dims in [(1, 2, 3), (7, 2, 0.5), (100, 1, 1), (3, 3, 0.01)]:
        d = prism_factors(*dims)
        assert sum(d) == pytest.approx(1.0, abs=1e-12)
        assert all(0 < v < 1 for v in d)
    # thin film limit: D_z -> 1; long rod: D_long -> 0, D_perp -> 1/2
    assert prism_factors(1000, 1000, 1)[2] > 0.995
    dx, dy, dz = prism_factors(1000, 1, 1)
    assert dx < 0.001 and dy == pytest.approx(0.5, abs=1e-3)


@pytest.mark.parametrize("grid,cell", [((6, 3, 2), (2e-9, 3e-9, 5e-9)), ((12, 2, 2), (4e-9, 3e-9, 3e-9)),
                                       ((1, 1, 1), (2e-9, 2e-9, 2e-9))])
def test_uniform_prism_demag_energy_equals_aharoni(oracle_mod, grid, cell):
    """E_d of a uniformly magnetised prism = (1/2) mu0 Ms^2 V D_axis: the
    Newell tensor is the exact cell-averaged interaction, so the sum over all
    cell pairs reproduces the prism's magnetometric factor (all distances
    below the 40-cell asymptotic switch).  The 1x1x1 cube gives mu0 Ms^2 V / 6."""
    nx, ny, nz = grid
    n = nx * ny * nz
    ms = 8e5
    vol = nx * ny * nz * cell[0] * cell[1] * cell[2]
    d = prism_factors(nx * cell[0], ny * cell[1], nz * cell[2])
    for axis in range(3):
        m = np.zeros((n, 3))
        m[:, axis] = 1.0
        p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2],
                                         ms=ms), m)
        e = p.observe()["e_demag"]
        assert e == pytest.approx(0.5 * MU0 * ms * ms * vol * d[axis], rel=1e-9)
    if n == 1:
        assert e == pytest.approx(MU0 * ms * ms * vol / 6.0, rel=1e-12)


def test_external_energy_closed_form(oracle_mod):
    """Uniform m in a uniform field: E_ext = -mu0 Ms V N (m . H)."""
    nx, ny, nz = 4, 3, 2
    n = nx * ny * nz
    cell = (2e-9, 3e-9, 4e-9)
    ms = 8e5
    h = (1.5e4, -2e4, 7e3)
    m = np.tile([0.6, 0.0, 0.8], (n, 1))
    p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], ms=ms,
                                     demag=False, h_ext=h), m)
    o = p.observe()
    v = cell[0] * cell[1] * cell[2]
    assert o["e_ext"] == pytest.approx(-MU0 * ms * v * n * (0.6 * h[0] + 0.8 * h[2]), rel=1e-13)
    assert o["e_total"] == pytest.approx(o["e_ext"], rel=1e-13)
    # masked cells carry no Zeeman energy
    mask = np.ones(n, np.uint8)
    mask[:5] = 0
    q = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], ms=ms,
                                     demag=False, h_ext=h, active_mask=mask), m)
    assert q.observe()["e_ext"] == pytest.approx(o["e_ext"] * (n - 5) / n, rel=1e-13)


def test_exchange_energy_closed_form_spiral(oracle_mod):
    """A planar spiral with angle step q between x neighbours: every pair
    contributes A V |2 sin(q/2)|^2 / dx^2, so E_ex = A V (nx-1) ny nz
    (2 sin(q/2) / dx)^2."""
    nx, ny, nz = 7, 2, 3
    cell = (2e-9, 3e-9, 4e-9)
    q = 0.3
    a = 1.3e-11
    x = np.arange(nx)
    ang = np.broadcast_to(q * x[None, None, :], (nz, ny, nx)).reshape(-1)
    m = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], -1)
    p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], a=a,
                                     exchange=True, demag=False), m)
    v = cell[0] * cell[1] * cell[2]
    expect = a * v * (nx - 1) * ny * nz * (2 * math.sin(q / 2) / cell[0]) ** 2
    assert p.observe()["e_ex"] == pytest.approx(expect, rel=1e-12)


# ---------------------------------------------------------------------------
# Adaptive integrators: error estimator and controller (lib.rs:1332-1568,
# 1653-1674)
# ---------------------------------------------------------------------------

def spin_problem(O, integrator, **kw):
    th0, ph0 = 1.0, 0.3
    m0 = np.array([[math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]])
    c = O.OracleConfig(nx=1, ny=1, nz=1, dx=1.0, dy=1.0, dz=1.0, ms=1.0, a=0.0, alpha=0.1, gamma=1.0,
                       exchange=False, demag=False, h_ext=(0.0, 0.0, 1.0), integrator=integrator, **kw)
    return O.OracleProblem(c, m0)


@pytest.mark.parametrize("integrator,order", [("dp45", 5), ("rk23", 3)])
def test_error_estimate_has_the_embedded_order(oracle_mod, integrator, order):
    """The estimate dt |sum_j E_j k_j| is the local error of the embedded
    lower-order solution, O(dt^(p+1)) with p = 4 (DP45) or 2 (RK23): halving
    dt divides it by 2^5 or 2^3.  A wrong E_j breaks the cancellation of the
    lower-order terms and the ratio falls to 2 or 4."""
    p = spin_problem(oracle_mod, integrator, max_error=1.0, dt_min=1e-12, dt_max=10.0, headroom=0.8)
    errs = [p.trial_error(dt) for dt in (0.04, 0.02, 0.01, 0.005)]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(3)]
    assert rates[-1] == pytest.approx(order, abs=0.15), rates
    np.testing.assert_allclose(p.m[0], spin_problem(oracle_mod, integrator).m[0], atol=0)  # state untouched


@pytest.mark.parametrize("integrator,expo", [("dp45", 0.2), ("rk23", 1.0 / 3.0)])
def test_controller_suggests_headroom_scaled_dt(oracle_mod, integrator, expo):
    """Accepted step: dt_next = clamp(headroom dt (tol / max(err, 1e-30))^expo,
    dt_min, dt_max) (lib.rs:1395-1404, 1550-1566)."""
    tol, hr = 1e-6, 0.8
    p = spin_problem(oracle_mod, integrator, max_error=tol, dt_min=1e-12, dt_max=10.0, headroom=hr)
    dt = 0.01
    err = p.trial_error(dt)
    assert err <= tol
    r = p.step(dt)
    assert r["dt_used"] == dt
    assert r["suggested_next_dt"] == pytest.approx(hr * dt * (tol / err) ** expo, rel=1e-12)
    # the upper clamp
    q = spin_problem(oracle_mod, integrator, max_error=tol, dt_min=1e-12, dt_max=0.011, headroom=hr)
    assert q.step(dt)["suggested_next_dt"] == 0.011


@pytest.mark.parametrize("integrator,expo", [("dp45", 0.2), ("rk23", 1.0 / 3.0)])
def test_controller_rejects_and_retries_with_shrunk_dt(oracle_mod, integrator, expo):
    """A rejected attempt (err > tol) retries at headroom dt (tol/err)^expo;
    the accepted dt is that retry when it passes."""
    tol, hr = 1e-8, 0.8
    dt0 = 0.2
    p = spin_problem(oracle_mod, integrator, max_error=tol, dt_min=1e-12, dt_max=10.0, headroom=hr)
    e0 = p.trial_error(dt0)
    assert e0 > tol
    dt1 = hr * dt0 * (tol / e0) ** expo
    e1 = p.trial_error(dt1)
    r = p.step(dt0)
    if e1 <= tol:
        assert r["dt_used"] == pytest.approx(dt1, rel=1e-14)
    else:
        assert r["dt_used"] < dt1
    assert r["dt_used"] < dt0


def test_dp45_fsal_slope_reused(oracle_mod):
    """After an accepted DP45 step the next step's first slope is the last
    stage slope of the previous one (FSAL, lib.rs:1412-1568): stepping on
    equals a fresh problem started from the new state, to rounding."""
    p = spin_problem(oracle_mod, "dp45", max_error=1e-7, dt_min=1e-12, dt_max=10.0, headroom=0.8)
    p.step(0.01)
    q = spin_problem(oracle_mod, "dp45", max_error=1e-7, dt_min=1e-12, dt_max=10.0, headroom=0.8)
    q.m = p.m
    a = p.step(0.02)
    b = q.step(0.02)
    assert a["dt_used"] == b["dt_used"]
    np.testing.assert_allclose(p.m, q.m, atol=1e-15)


# ---------------------------------------------------------------------------
# Per-pair exchange table (native/backends/fdm/src/exchange_fp64.cu:80-114)
# ---------------------------------------------------------------------------

def lut_with(a_same, pairs):
    lut = np.zeros((256, 256))
    for r in range(256):
        lut[r, r] = a_same
    for (i, j), v in pairs.items():
        lut[i, j] = lut[j, i] = v
    return lut


def test_region_table_three_cell_chain_known_answer(oracle_mod):
    """Regions (0, 1, 1), A_00 = A_11 = A = mu0/2, A_01 = 0.3 A, Ms = 1,
    d = 1: H_f = sum_faces (A_f,j / A) (m_j - m_f), worked by hand:
    H_0 = 0.3 (m_1 - m_0), H_1 = 0.3 (m_0 - m_1) + (m_2 - m_1),
    H_2 = m_1 - m_2."""
    a = 0.5 * MU0
    m = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    c = cfg(oracle_mod, nx=3, dx=1.0, dy=1.0, dz=1.0, ms=1.0, a=a, exchange=True, demag=False,
            region_mask=np.array([0, 1, 1]), exchange_lut=lut_with(a, {(0, 1): 0.3 * a}))
    h = oracle_mod.OracleProblem(c, m).observe()["h_ex"]
    np.testing.assert_allclose(h, [[-0.3, 0.3, 0.0], [0.3, -1.3, 1.0], [0.0, 1.0, -1.0]], atol=1e-15)


def test_region_table_default_and_masked_neighbour(oracle_mod):
    """Without a caller table the default decouples regions (A on the
    diagonal only); an inactive neighbour contributes nothing whatever its
    table entry; a uniform region id equals plain exchange."""
    nx, ny, nz = 5, 3, 2
    n = nx * ny * nz
    cell = (2e-9, 3e-9, 4e-9)
    m = oracle_mod.random_unit_vectors(5, n)
    region = np.zeros(n, np.uint32)
    region.reshape(nz, ny, nx)[:, :, 3:] = 1
    common = dict(nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], exchange=True, demag=False)
    h_lut = oracle_mod.OracleProblem(cfg(oracle_mod, region_mask=region, **common), m).observe()["h_ex"]
    h_mask = oracle_mod.OracleProblem(cfg(oracle_mod, active_mask=(region == 0).astype(np.uint8), **common),
                                      m).observe()["h_ex"]
    sel = region == 0
    np.testing.assert_allclose(h_lut[sel], h_mask[sel], rtol=1e-13, atol=1e-6)
    h_uni = oracle_mod.OracleProblem(cfg(oracle_mod, region_mask=np.zeros(n, np.uint32), **common),
                                     m).observe()["h_ex"]
    h_plain = oracle_mod.OracleProblem(cfg(oracle_mod, **common), m).observe()["h_ex"]
    np.testing.assert_allclose(h_uni, h_plain, rtol=1e-12, atol=1e-6 * np.abs(h_plain).max() * 1e-6)


def test_region_table_energy_is_the_pair_sum(oracle_mod):
    """With a symmetric table, -mu0 Ms V / 2 sum m . H_ex equals the GPU
    reference's forward-pair energy sum A_ij V |m_j - m_i|^2 / d^2
    (native/backends/fdm/src/reductions_fp64.cu:123-214)."""
    nx, ny, nz = 4, 3, 2
    n = nx * ny * nz
    cell = (2e-9, 3e-9, 4e-9)
    a = 1.3e-11
    region = (np.arange(n) % 3).astype(np.uint32)
    lut = lut_with(a, {(0, 1): 0.4 * a, (1, 2): 2.0 * a, (0, 2): 0.0})
    m = oracle_mod.random_unit_vectors(8, n)
    p = oracle_mod.OracleProblem(cfg(oracle_mod, nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], a=a,
                                     exchange=True, demag=False, region_mask=region, exchange_lut=lut), m)
    e = p.observe()["e_ex"]
    mm = m.reshape(nz, ny, nx, 3)
    rr = region.reshape(nz, ny, nx)
    v = cell[0] * cell[1] * cell[2]
    e_pair = 0.0
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        for (ox, oy, oz), d in (((1, 0, 0), cell[0]), ((0, 1, 0), cell[1]), ((0, 0, 1), cell[2])):
            X, Y, Z = x + ox, y + oy, z + oz
            if X < nx and Y < ny and Z < nz:
                aij = lut[rr[z, y, x], rr[Z, Y, X]]
                e_pair += aij * v * np.sum((mm[Z, Y, X] - mm[z, y, x]) ** 2) / d ** 2
    assert e == pytest.approx(e_pair, rel=1e-12)
