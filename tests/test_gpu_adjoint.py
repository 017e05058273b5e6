"""The HIP Heun adjoint (fullmag_fdm_ext_heun_adjoint) against the oracle's
adjoint, which tests/test_oracle_adjoint.py pins with central finite
differences: FP64 to 1e-10, mixed precision to 1e-5 of the FP64 adjoint."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1e-13


@pytest.fixture(scope="module")
def torch_cuda(fa):
    import torch

    assert torch.cuda.is_available(), (f"torch {torch.__version__} sees no GPU (hip {torch.version.hip}, "
                                       f"device_count {torch.cuda.device_count()})")
    return torch


def soa(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda().contiguous()


def gpu_adjoint(fa, torch, plan_kw, m0, lam1):
    n = m0.shape[0]
    b = fa.NativeFdmBackend(fa.FdmPlan(initial_magnetization=m0, **plan_kw))
    m0_t, l1_t = soa(torch, m0), soa(torch, lam1)
    l0_t = torch.zeros_like(m0_t)
    gms_t = torch.zeros(n, dtype=torch.float64, device="cuda")
    gh = np.zeros(3)
    torch.cuda.synchronize()
    b.heun_adjoint(DT, m0_t.data_ptr(), l1_t.data_ptr(), l0_t.data_ptr(), gms_t.data_ptr(), gh)
    out = l0_t.cpu().numpy().T.copy(), gms_t.cpu().numpy(), gh
    b.close()
    return out


CASES = [((12, 10, 3), False), ((16, 12, 1), False), ((9, 7, 2), True), ((32, 16, 4), False)]


@pytest.mark.parametrize("cells,masked", CASES)
def test_heun_adjoint_fp64_matches_oracle(fa, oracle_mod, torch_cuda, cells, masked):
    torch = torch_cuda
    nx, ny, nz = cells
    n = nx * ny * nz
    rng = np.random.default_rng(n + masked)
    m0 = fa.generate_random_unit_vectors(21 + n, n)
    ms_cell = rng.uniform(6e5, 9e5, n)
    mask = (rng.random(n) > 0.2).astype(np.uint8) if masked else None
    if mask is not None:
        m0 = m0 * mask[:, None]
    lam1 = rng.standard_normal((n, 3))
    h_ext = (2e4, 5e3, -1e4)
    cell = (5e-9, 5e-9, 3e-9)
    plan_kw = dict(cells=cells, cell_size=cell, saturation_magnetisation=8e5, exchange_stiffness=1.3e-11,
                   damping=0.1, external_field=h_ext, precision="double", demag_kernel="device", ms_cell=ms_cell,
                   active_mask=mask)
    l0, gms, gh = gpu_adjoint(fa, torch, plan_kw, m0, lam1)
    cfg = oracle_mod.OracleConfig(nx=nx, ny=ny, nz=nz, dx=cell[0], dy=cell[1], dz=cell[2], ms=8e5, a=1.3e-11,
                                  alpha=0.1, h_ext=h_ext, ms_cell=ms_cell, active_mask=mask)
    p = oracle_mod.OracleProblem(cfg, m0)
    r0, rms, rh = p.heun_adjoint(DT, m0, lam1)
    np.testing.assert_allclose(l0, r0, rtol=0, atol=1e-10 * np.abs(r0).max())
    np.testing.assert_allclose(gms, rms, rtol=0, atol=1e-10 * np.abs(rms).max())
    np.testing.assert_allclose(gh, rh, rtol=1e-10, atol=0)


def test_heun_adjoint_mixed_tracks_fp64(fa, torch_cuda):
    torch = torch_cuda
    cells = (64, 64, 4)
    n = 64 * 64 * 4
    rng = np.random.default_rng(4)
    m0 = fa.generate_random_unit_vectors(8, n)
    ms_cell = rng.uniform(6e5, 9e5, n)
    lam1 = rng.standard_normal((n, 3))
    base = dict(cells=cells, cell_size=(5e-9, 5e-9, 3e-9), saturation_magnetisation=8e5,
                exchange_stiffness=1.3e-11, damping=0.1, external_field=(1.6e4, 4e3, 0.0), demag_kernel="device",
                ms_cell=ms_cell)
    d0, dms, dh = gpu_adjoint(fa, torch, dict(base, precision="double"), m0, lam1)
    x0, xms, xh = gpu_adjoint(fa, torch, dict(base, precision="mixed"), m0, lam1)
    assert float(np.abs(x0 - d0).max()) <= 1e-5 * float(np.abs(d0).max())
    assert float(np.abs(xms - dms).max()) <= 1e-4 * float(np.abs(dms).max())
    np.testing.assert_allclose(xh, dh, rtol=1e-4)


def test_device_state_round_trip(fa, torch_cuda):
    torch = torch_cuda
    n = 16 * 8 * 2
    m0 = fa.generate_random_unit_vectors(5, n)
    b = fa.NativeFdmBackend(fa.FdmPlan(cells=(16, 8, 2), cell_size=(4e-9, 4e-9, 2e-9), initial_magnetization=m0,
                                       saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1,
                                       precision="double", demag_kernel="device"))
    t = torch.zeros(3 * n, dtype=torch.float64, device="cuda")
    b.copy_state_device(t.data_ptr())
    np.testing.assert_array_equal(t.cpu().numpy().reshape(3, n).T, b.copy_m())
    s0 = b.step(1e-13)
    m1 = b.copy_m()
    b.set_state_device(t.data_ptr())
    np.testing.assert_array_equal(b.copy_m(), m0)
    s1 = b.step(1e-13)
    np.testing.assert_array_equal(b.copy_m(), m1)
    assert s1.e_total == s0.e_total
    b.close()


def rollout_loss(fa, torch, plan, m0, ms, h, steps, every, window):
    m = fa.heun_rollout(m0, ms, h, plan, DT, steps, every)
    nx, ny, nz = plan.cells
    x0, y0, w = window
    mx = m[:, 0].reshape(nz, ny, nx)[:, y0:y0 + w, x0:x0 + w]
    return mx.mean()


def test_heun_rollout_gradients_match_finite_differences(fa, torch_cuda):
    """Film 32x32x2 of 5x5x3 nm, per-cell Ms ~ U[6e5, 9e5], alpha 0.1,
    B = (20, 5, 0) mT, 200 Heun steps of 1e-13 s with checkpoints every 20;
    loss = mean m_x over a centred 8x8 window of all layers.  Adjoint
    gradients vs central differences (h = 1e-4 relative) at 6 cells and the
    field components, FP64."""
    torch = torch_cuda
    cells = (32, 32, 2)
    n = 32 * 32 * 2
    rng = np.random.default_rng(2024)
    m0 = fa.generate_random_unit_vectors(99, n)
    ms_np = rng.uniform(6e5, 9e5, n)
    mu0 = 4e-7 * np.pi
    h_np = np.array([20e-3, 5e-3, 0.0]) / mu0
    plan = fa.FdmPlan(cells=cells, cell_size=(5e-9, 5e-9, 3e-9), initial_magnetization=m0,
                      saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1, precision="double",
                      demag_kernel="device")
    steps, every, window = 200, 20, (12, 12, 8)
    ms = torch.tensor(ms_np, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(h_np, dtype=torch.float64, requires_grad=True)
    loss = rollout_loss(fa, torch, plan, m0, ms, h, steps, every, window)
    loss.backward()
    g_ms, g_h = ms.grad.numpy(), h.grad.numpy()

    def L(ms_v, h_v):
        with torch.no_grad():
            return float(rollout_loss(fa, torch, plan, m0, torch.tensor(ms_v), torch.tensor(h_v), steps, every,
                                      window))

    # cells inside and next to the window, where the gradient is largest
    idx = [(0 * 32 + 16) * 32 + 16, (1 * 32 + 13) * 32 + 18, (0 * 32 + 11) * 32 + 12, (1 * 32 + 20) * 32 + 19,
           (0 * 32 + 2) * 32 + 30, (1 * 32 + 16) * 32 + 8]
    scale = np.abs(g_ms).max()
    for i in idx:
        d = 1e-4 * ms_np[i]
        p, q = ms_np.copy(), ms_np.copy()
        p[i] += d
        q[i] -= d
        fd = (L(p, h_np) - L(q, h_np)) / (2 * d)
        assert abs(fd - g_ms[i]) <= 1e-4 * max(abs(fd), 1e-3 * scale), (i, fd, g_ms[i])
    for k in range(2):
        d = 1e-4 * np.abs(h_np).max()
        p, q = h_np.copy(), h_np.copy()
        p[k] += d
        q[k] -= d
        fd = (L(ms_np, p) - L(ms_np, q)) / (2 * d)
        assert abs(fd - g_h[k]) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(g_h).max()), (k, fd, g_h[k])


def smooth_film_m0(nx, ny, nz):
    """In-plane state with gentle x and y modulations (no domain walls)."""
    x = np.arange(nx)[None, None, :] * np.ones((nz, ny, 1))
    y = np.arange(ny)[None, :, None] * np.ones((nz, 1, nx))
    th = 0.4 * np.sin(2 * np.pi * x / nx) + 0.2 * np.cos(2 * np.pi * y / ny)
    m = np.stack([np.cos(th), np.sin(th), 0.1 * np.ones_like(th)], -1).reshape(-1, 3)
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def test_differentiable_film_128x128x4_full_size(fa, torch_cuda):
    """The differentiable workload at its configured size: 128x128x4 film of
    5x5x3 nm, per-cell Ms ~ U[6e5, 9e5] (seed 4), A 1.3e-11, alpha 0.1,
    B = (20, 5, 0) mT, 2000 Heun steps of 1e-13 s with checkpoints every 100
    from a smooth in-plane m0; loss = mean m_x over the centred 32x32 window
    of all layers.  FP64 adjoint gradients w.r.t. Ms at 32 seeded cells and
    the in-plane B components against central differences (h = 1e-4
    relative) to 1e-4 relative; the mixed-precision gradient within 1e-4 of
    the largest FP64 entry."""
    torch = torch_cuda
    nx, ny, nz = 128, 128, 4
    n = nx * ny * nz
    steps, every = 2000, 100
    rng = np.random.default_rng(4)
    ms_np = rng.uniform(6e5, 9e5, n)
    m0 = smooth_film_m0(nx, ny, nz)
    mu0 = 4e-7 * np.pi
    b_mt = np.array([20.0, 5.0, 0.0])
    grads = {}
    for prec in ("double", "mixed"):
        plan = fa.FdmPlan(cells=(nx, ny, nz), cell_size=(5e-9, 5e-9, 3e-9), initial_magnetization=m0,
                          saturation_magnetisation=8e5, exchange_stiffness=1.3e-11, damping=0.1, precision=prec,
                          demag_kernel="device")
        ms = torch.tensor(ms_np, dtype=torch.float64, requires_grad=True)
        bt = torch.tensor(b_mt, dtype=torch.float64, requires_grad=True)
        loss = rollout_loss(fa, torch, plan, m0, ms, bt * 1e-3 / mu0, steps, every, (48, 48, 32))
        loss.backward()
        grads[prec] = (ms.grad.numpy().copy(), bt.grad.numpy().copy(), plan)
    g_ms, g_b, plan = grads["double"]

    def L(ms_v, b_v):
        with torch.no_grad():
            return float(rollout_loss(fa, torch, plan, m0, torch.tensor(ms_v), torch.tensor(b_v * 1e-3 / mu0),
                                      steps, steps, (48, 48, 32)))

    cells = rng.choice(n, size=32, replace=False)
    fds, errs = [], []
    for i in cells:
        d = 1e-4 * ms_np[i]
        p, q = ms_np.copy(), ms_np.copy()
        p[i] += d
        q[i] -= d
        fds.append((L(p, b_mt) - L(q, b_mt)) / (2 * d))
        errs.append(abs(fds[-1] - g_ms[i]))
    scale = max(abs(f) for f in fds)
    rel = [e / max(abs(f), 1e-3 * scale) for e, f in zip(errs, fds)]
    print(f"128x128x4 adjoint vs FD: max rel err (Ms, 32 cells) {max(rel):.3e}")
    assert max(rel) <= 1e-4, (max(rel), int(cells[int(np.argmax(rel))]))
    for k in range(2):
        d = 1e-4 * np.abs(b_mt).max()
        p, q = b_mt.copy(), b_mt.copy()
        p[k] += d
        q[k] -= d
        fd = (L(ms_np, p) - L(ms_np, q)) / (2 * d)
        print(f"  B[{k}]: adjoint {g_b[k]:.6e} fd {fd:.6e}")
        assert abs(fd - g_b[k]) <= 1e-4 * abs(fd), (k, fd, g_b[k])
    dev = float(np.abs(grads["mixed"][0] - g_ms).max() / np.abs(g_ms).max())
    print(f"  mixed vs FP64 gradient: {dev:.3e} of the largest entry")
    assert dev <= 1e-4
