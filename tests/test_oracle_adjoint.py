"""The oracle's discrete Heun adjoint against central finite differences of
the oracle's own Heun step: state directions (the dot-product identity
<lam1, J dm> = <J^T lam1, dm>), per-cell Ms and H_ext, with exchange, demag,
Zeeman, masks and thin films."""

import numpy as np
import pytest

from oracle import oracle as O

DT = 1e-13


def problem(cells, m0, ms_cell, h_ext, mask=None, alpha=0.3, precession=True):
    nx, ny, nz = cells
    cfg = O.OracleConfig(nx=nx, ny=ny, nz=nz, dx=4e-9, dy=3e-9, dz=2e-9, ms=8e5, a=1.3e-11, alpha=alpha,
                         h_ext=h_ext, active_mask=mask, ms_cell=ms_cell, precession=precession)
    return O.OracleProblem(cfg, m0)


def objective(cells, m0_raw, ms_cell, h_ext, lam1, mask=None, **kw):
    p = problem(cells, np.ones_like(m0_raw), ms_cell, h_ext, mask, **kw)
    p.set_m_raw(m0_raw)
    p.step(DT)
    return float(np.sum(lam1 * p.m))


CASES = [
    ((6, 5, 3), None, {}),
    ((7, 4, 1), None, {}),
    ((5, 5, 2), "mask", {}),
    ((4, 6, 2), None, {"precession": False, "alpha": 1.0}),
]


@pytest.mark.parametrize("cells,maskkind,kw", CASES)
def test_heun_adjoint_matches_finite_differences(cells, maskkind, kw):
    nx, ny, nz = cells
    n = nx * ny * nz
    rng = np.random.default_rng(n)
    m0 = O.random_unit_vectors(3 + n, n)
    ms_cell = rng.uniform(6e5, 9e5, n)
    h_ext = (2e4, -1.5e4, 8e3)
    mask = None
    if maskkind == "mask":
        mask = (rng.random(n) > 0.25).astype(np.uint8)
        m0 = m0 * mask[:, None]
    lam1 = rng.standard_normal((n, 3))
    p = problem(cells, m0, ms_cell, h_ext, mask, **kw)
    lam0, gms, gh = p.heun_adjoint(DT, m0, lam1)

    # state directions: <lam1, J dm> = <lam0, dm>
    for trial in range(3):
        dm = rng.standard_normal((n, 3))
        if mask is not None:
            dm *= mask[:, None]
        h = 1e-5
        fp = objective(cells, m0 + h * dm, ms_cell, h_ext, lam1, mask, **kw)
        fm = objective(cells, m0 - h * dm, ms_cell, h_ext, lam1, mask, **kw)
        fd = (fp - fm) / (2 * h)
        ad = float(np.sum(lam0 * dm))
        assert abs(fd - ad) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(lam1).sum()), (trial, fd, ad)

    # per-cell Ms
    active = np.arange(n) if mask is None else np.nonzero(mask)[0]
    for i in rng.choice(active, size=4, replace=False):
        h = 1e-4 * ms_cell[i]
        msp, msm = ms_cell.copy(), ms_cell.copy()
        msp[i] += h
        msm[i] -= h
        fd = (objective(cells, m0, msp, h_ext, lam1, mask, **kw) -
              objective(cells, m0, msm, h_ext, lam1, mask, **kw)) / (2 * h)
        assert abs(fd - gms[i]) <= 1e-5 * max(abs(fd), 1e-9 * np.abs(gms).max()), (i, fd, gms[i])

    # H_ext
    for k in range(3):
        h = 1.0  # A/m, against fields of 1e4-1e6 A/m
        hp, hm = list(h_ext), list(h_ext)
        hp[k] += h
        hm[k] -= h
        fd = (objective(cells, m0, ms_cell, tuple(hp), lam1, mask, **kw) -
              objective(cells, m0, ms_cell, tuple(hm), lam1, mask, **kw)) / (2 * h)
        assert abs(fd - gh[k]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gh).max()), (k, fd, gh[k])


def test_heun_adjoint_planted_mistake_is_caught():
    """A wrong sign in one adjoint term would fail the identity above: the
    identity itself distinguishes J^T from -J^T and from J (J is not
    symmetric here)."""
    cells = (5, 4, 2)
    n = 40
    rng = np.random.default_rng(9)
    m0 = O.random_unit_vectors(77, n)
    ms_cell = rng.uniform(6e5, 9e5, n)
    lam1 = rng.standard_normal((n, 3))
    p = problem(cells, m0, ms_cell, (1e4, 0.0, 0.0))
    lam0, _, _ = p.heun_adjoint(DT, m0, lam1)
    dm = rng.standard_normal((n, 3))
    h = 1e-5
    fd = (objective(cells, m0 + h * dm, ms_cell, (1e4, 0.0, 0.0), lam1) -
          objective(cells, m0 - h * dm, ms_cell, (1e4, 0.0, 0.0), lam1)) / (2 * h)
    assert abs(fd - np.sum(lam0 * dm)) <= 1e-6 * abs(fd)
    assert abs(fd + np.sum(lam0 * dm)) > 1e-3 * abs(fd)
