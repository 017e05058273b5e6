"""ctypes wrapper around the CPU FP64 oracle (oracle/fdm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the fullmag_amd package.

The oracle restates fullmag-engine's CPU reference (crates/fullmag-engine/
src/lib.rs), fullmag-fdm-demag's Newell tensor (crates/fullmag-fdm-demag/src/
newell.rs) and the planner's m0 generator (crates/fullmag-plan/src/lib.rs:
1927-1957).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

MU0 = 4.0 * np.pi * 1e-7
INTEGRATORS = {"heun": 1, "dp45": 2, "rk45": 2, "abm3": 3, "rk4": 4, "rk23": 5}

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, d, i = ctypes.c_uint64, ctypes.c_double, ctypes.c_int
        L.oracle_newell_f.restype = d
        L.oracle_newell_f.argtypes = [d, d, d]
        L.oracle_newell_g.restype = d
        L.oracle_newell_g.argtypes = [d, d, d]
        L.oracle_newell_kernels.argtypes = [u64, u64, u64, d, d, d, _dp]
        L.oracle_newell_spectra.restype = i
        L.oracle_newell_spectra.argtypes = [u64, u64, u64, d, d, d, i, _dp]
        L.oracle_fft.argtypes = [_dp, u64, i]
        L.oracle_fft3.argtypes = [_dp, u64, u64, u64, i]
        L.oracle_random_unit_vectors.argtypes = [u64, u64, _dp]
        L.oracle_problem_create.restype = ctypes.c_void_p
        L.oracle_problem_create.argtypes = [u64, u64, u64, _dp, ctypes.POINTER(i),
                                            ctypes.c_void_p, ctypes.c_void_p, _dp]
        L.oracle_problem_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_problem_step.restype = i
        L.oracle_problem_step.argtypes = [ctypes.c_void_p, d, _dp]
        L.oracle_problem_get_m.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_problem_set_m.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_problem_set_m_raw.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_problem_heun_adjoint.restype = i
        L.oracle_problem_heun_adjoint.argtypes = [ctypes.c_void_p, d, _dp, _dp, _dp, ctypes.c_void_p,
                                                  ctypes.c_void_p]
        L.oracle_problem_time.restype = d
        L.oracle_problem_time.argtypes = [ctypes.c_void_p]
        L.oracle_problem_observe.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_problem_set_regions.restype = i
        L.oracle_problem_set_regions.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_problem_set_tiers.restype = i
        L.oracle_problem_set_tiers.argtypes = [ctypes.c_void_p, i, d, d, _dp, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]
        L.oracle_newell_kernels_shifted.argtypes = [u64, u64, u64, d, d, d, d, _dp]
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_push_m.argtypes = [_dp, _u64p, _dp, _u64p, _dp, _dp]
        L.oracle_pull_h.argtypes = [_dp, _u64p, _dp, _u64p, _dp, _dp]
        L.oracle_stack_create.restype = ctypes.c_void_p
        L.oracle_stack_create.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p), _dp, _u64p, _dp, i]
        L.oracle_stack_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_stack_step.restype = i
        L.oracle_stack_step.argtypes = [ctypes.c_void_p, d, _dp]
        L.oracle_stack_observe.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_problem_trial_error.restype = d
        L.oracle_problem_trial_error.argtypes = [ctypes.c_void_p, d]
        L.oracle_num_threads.restype = i
        L.oracle_set_num_threads.argtypes = [i]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def set_num_threads(n: int) -> None:
    lib().oracle_set_num_threads(int(n))


def num_threads() -> int:
    return lib().oracle_num_threads()


def newell_f(x, y, z) -> float:
    return lib().oracle_newell_f(x, y, z)


def newell_g(x, y, z) -> float:
    return lib().oracle_newell_g(x, y, z)


def newell_kernels(nx, ny, nz, dx, dy, dz) -> np.ndarray:
    """Real-space kernels, shape (6, 2nz, 2ny, 2nx), order xx,yy,zz,xy,xz,yz."""
    out = np.zeros((6, 2 * nz, 2 * ny, 2 * nx))
    lib().oracle_newell_kernels(nx, ny, nz, dx, dy, dz, _ptr(out))
    return out


def newell_spectra(nx, ny, nz, dx, dy, dz, thin_film_2d: bool = False) -> np.ndarray:
    """Interleaved complex spectra as handed to the C ABI: shape (6, 2*Npad).

    thin_film_2d follows compute_newell_kernel_spectra_thin_film_2d
    (crates/fullmag-engine/src/lib.rs:553-610): nz must be 1, Npad = 4 nx ny.
    Otherwise compute_newell_kernel_spectra (lib.rs:511-551), Npad = 8 nx ny nz.
    """
    npad = 4 * nx * ny if thin_film_2d else 8 * nx * ny * nz
    out = np.zeros((6, 2 * npad))
    rc = lib().oracle_newell_spectra(nx, ny, nz, dx, dy, dz, 1 if thin_film_2d else 0, _ptr(out))
    if rc != 0:
        raise ValueError("thin-film 2-D spectra need nz == 1")
    return out


def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().oracle_fft(a.ctypes.data_as(_dp), a.size, 1 if inverse else 0)
    return a


def fft3(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """x has shape (pz, py, px) (x fastest)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    pz, py, px = a.shape
    lib().oracle_fft3(a.ctypes.data_as(_dp), px, py, pz, 1 if inverse else 0)
    return a


def random_unit_vectors(seed: int, count: int) -> np.ndarray:
    out = np.zeros((count, 3))
    lib().oracle_random_unit_vectors(ctypes.c_uint64(seed), count, _ptr(out))
    return out


@dataclass
class OracleConfig:
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    ms: float
    a: float
    alpha: float
    gamma: float = 2.211e5
    integrator: str = "heun"
    precession: bool = True
    exchange: bool = True
    demag: bool = True
    h_ext: Optional[tuple] = None  # A/m
    active_mask: Optional[np.ndarray] = None
    ms_cell: Optional[np.ndarray] = None  # extension: per-cell Ms
    region_mask: Optional[np.ndarray] = None  # region id per cell (< 256)
    exchange_lut: Optional[np.ndarray] = None  # 256 x 256 A_ij; None: A on the diagonal
    # sub-cell boundary tiers: 1 = T0 face links, >= 2 = T1 distances; phi
    # (n,), face_links / deltas (6, n) in the order xp, xm, yp, ym, zp, zm,
    # correction rows: target (rows,), source (rows, stencil), tensor
    # (rows, stencil, 6) as Nxx, Nxy, Nxz, Nyy, Nyz, Nzz
    tier: int = 0
    phi: Optional[np.ndarray] = None
    face_links: Optional[np.ndarray] = None
    deltas: Optional[np.ndarray] = None
    phi_floor: float = 0.0
    delta_min: float = 0.0
    corr_target: Optional[np.ndarray] = None
    corr_source: Optional[np.ndarray] = None
    corr_tensor: Optional[np.ndarray] = None
    max_error: float = 0.0
    dt_min: float = 0.0
    dt_max: float = 0.0
    headroom: float = 0.0


REPORT_FIELDS = ("time", "dt_used", "suggested_next_dt", "e_ex", "e_demag", "e_ext",
                 "e_total", "max_h_eff", "max_h_demag", "max_dm_dt")


class OracleProblem:
    """ExchangeLlgProblem + ExchangeLlgState + workspace (lib.rs:869-1147)."""

    def __init__(self, cfg: OracleConfig, m0: np.ndarray):
        self.cfg = cfg
        self.n = cfg.nx * cfg.ny * cfg.nz
        h = cfg.h_ext if cfg.h_ext is not None else (0.0, 0.0, 0.0)
        params = np.array([cfg.dx, cfg.dy, cfg.dz, cfg.ms, cfg.a, cfg.alpha, cfg.gamma,
                           h[0], h[1], h[2], cfg.max_error, cfg.dt_min, cfg.dt_max,
                           cfg.headroom], dtype=np.float64)
        flags = (ctypes.c_int * 5)(INTEGRATORS[cfg.integrator], int(cfg.precession),
                                   int(cfg.exchange), int(cfg.demag), int(cfg.h_ext is not None))
        m0 = np.ascontiguousarray(m0, dtype=np.float64).reshape(self.n, 3)
        self._mask = None
        if cfg.active_mask is not None:
            self._mask = np.ascontiguousarray(cfg.active_mask, dtype=np.uint8).ravel()
        self._ms = None
        if cfg.ms_cell is not None:
            self._ms = np.ascontiguousarray(cfg.ms_cell, dtype=np.float64).ravel()
        self._h = lib().oracle_problem_create(
            cfg.nx, cfg.ny, cfg.nz, _ptr(params), flags,
            None if self._mask is None else self._mask.ctypes.data_as(ctypes.c_void_p),
            None if self._ms is None else self._ms.ctypes.data_as(ctypes.c_void_p),
            _ptr(m0))
        self._region = self._lut = None
        if cfg.region_mask is not None:
            self._region = np.ascontiguousarray(cfg.region_mask, dtype=np.uint32).ravel()
            if cfg.exchange_lut is not None:
                self._lut = np.ascontiguousarray(cfg.exchange_lut, dtype=np.float64).reshape(256 * 256)
            rc = lib().oracle_problem_set_regions(
                self._h, self._region.ctypes.data_as(ctypes.c_void_p),
                None if self._lut is None else self._lut.ctypes.data_as(ctypes.c_void_p))
            if rc != 0:
                raise ValueError("region ids must be below 256")
        if cfg.tier and cfg.phi is not None:
            self._tiers = [np.ascontiguousarray(cfg.phi, dtype=np.float64).reshape(self.n)]
            vp = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
            face = delta = None
            if cfg.face_links is not None:
                face = np.ascontiguousarray(cfg.face_links, dtype=np.float64).reshape(6 * self.n)
            if cfg.deltas is not None:
                delta = np.ascontiguousarray(cfg.deltas, dtype=np.float64).reshape(6 * self.n)
            rows = stencil = 0
            tgt = src = ten = None
            if cfg.corr_target is not None:
                tgt = np.ascontiguousarray(cfg.corr_target, dtype=np.int32).reshape(-1)
                rows = tgt.size
                src = np.ascontiguousarray(cfg.corr_source, dtype=np.int32).reshape(rows, -1)
                stencil = src.shape[1]
                ten = np.ascontiguousarray(cfg.corr_tensor, dtype=np.float64).reshape(rows * stencil * 6)
            self._tiers += [face, delta, tgt, src, ten]
            lib().oracle_problem_set_tiers(self._h, int(cfg.tier), cfg.phi_floor, cfg.delta_min,
                                           _ptr(self._tiers[0]), vp(face), vp(delta), rows, stencil, vp(tgt),
                                           vp(src), vp(ten))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_problem_destroy(h)
            self._h = None

    def step(self, dt: float) -> dict:
        r = np.zeros(10)
        rc = lib().oracle_problem_step(self._h, dt, _ptr(r))
        if rc != 0:
            raise ValueError("dt must be positive")
        return dict(zip(REPORT_FIELDS, r.tolist()))

    @property
    def m(self) -> np.ndarray:
        out = np.zeros((self.n, 3))
        lib().oracle_problem_get_m(self._h, _ptr(out))
        return out

    @m.setter
    def m(self, value: np.ndarray) -> None:
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(self.n, 3)
        lib().oracle_problem_set_m(self._h, _ptr(v))

    def set_m_raw(self, value: np.ndarray) -> None:
        """State without normalisation (inactive cells zero)."""
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(self.n, 3)
        lib().oracle_problem_set_m_raw(self._h, _ptr(v))

    def heun_adjoint(self, dt: float, m0: np.ndarray, lam1: np.ndarray):
        """Adjoint of one Heun step from m0: returns (lam0, grad_ms (N),
        grad_hext (3)) with lam0 = (dm1/dm0)^T lam1 and the gradients of
        <lam1, m1> with respect to per-cell Ms and H_ext."""
        m0 = np.ascontiguousarray(m0, dtype=np.float64).reshape(self.n, 3)
        lam1 = np.ascontiguousarray(lam1, dtype=np.float64).reshape(self.n, 3)
        lam0 = np.zeros((self.n, 3))
        gms = np.zeros(self.n)
        gh = np.zeros(3)
        lib().oracle_problem_heun_adjoint(self._h, dt, _ptr(m0), _ptr(lam1), _ptr(lam0),
                                          gms.ctypes.data_as(ctypes.c_void_p), gh.ctypes.data_as(ctypes.c_void_p))
        return lam0, gms, gh

    def trial_error(self, dt: float) -> float:
        """Error norm of one attempted RK23 / DP45 step of size dt from the
        current state (state unchanged)."""
        return lib().oracle_problem_trial_error(self._h, dt)

    @property
    def time(self) -> float:
        return lib().oracle_problem_time(self._h)

    def observe(self) -> dict:
        """Fields (AoS, shape (N,3)) and scalars at the current state."""
        f = np.zeros((5, self.n, 3))
        s = np.zeros(10)
        lib().oracle_problem_observe(self._h, f.ctypes.data_as(ctypes.c_void_p),
                                     s.ctypes.data_as(ctypes.c_void_p))
        out = dict(zip(REPORT_FIELDS, s.tolist()))
        out.update(h_ex=f[0], h_demag=f[1], h_ext=f[2], h_eff=f[3], rhs=f[4])
        return out


def newell_kernels_shifted(nx, ny, nz, dx, dy, dz, z_offset) -> np.ndarray:
    """Shifted cross-layer kernel (newell.rs:528-651), shape (6, 2nz, 2ny, 2nx)."""
    out = np.zeros((6, 2 * nz, 2 * ny, 2 * nx))
    lib().oracle_newell_kernels_shifted(nx, ny, nz, dx, dy, dz, z_offset, _ptr(out))
    return out


def _u64(v):
    return (ctypes.c_uint64 * 3)(*[int(x) for x in v])


def push_m(m, native_cells, native_cell, conv_cells, conv_cell) -> np.ndarray:
    """transfer.rs push_m: native (N, 3) -> convolution grid (Nc, 3)."""
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((int(np.prod(conv_cells)), 3))
    lib().oracle_push_m(_ptr(m), _u64(native_cells), _ptr(np.asarray(native_cell, dtype=np.float64)),
                        _u64(conv_cells), _ptr(np.asarray(conv_cell, dtype=np.float64)), _ptr(out))
    return out


def pull_h(h, conv_cells, conv_cell, native_cells, native_cell) -> np.ndarray:
    """transfer.rs pull_h: convolution grid (Nc, 3) -> native (N, 3)."""
    h = np.ascontiguousarray(h, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((int(np.prod(native_cells)), 3))
    lib().oracle_pull_h(_ptr(h), _u64(conv_cells), _ptr(np.asarray(conv_cell, dtype=np.float64)),
                        _u64(native_cells), _ptr(np.asarray(native_cell, dtype=np.float64)), _ptr(out))
    return out


class OracleStack:
    """Multilayer stack (crates/fullmag-engine/src/multilayer.rs, runner
    multilayer_reference.rs): `layers` are OracleProblem objects built with
    demag=False (their exchange, Zeeman, mask and material), coupled by the
    stack's demag on the common convolution grid; steps are Heun."""

    def __init__(self, layers, origins, common_cells, conv_cell, demag=True):
        self.layers = list(layers)
        self._org = np.ascontiguousarray(origins, dtype=np.float64).reshape(len(self.layers), 3)
        arr = (ctypes.c_void_p * len(self.layers))(*[p._h for p in self.layers])
        self._h = lib().oracle_stack_create(len(self.layers), arr, _ptr(self._org), _u64(common_cells),
                                            _ptr(np.asarray(conv_cell, dtype=np.float64)), 1 if demag else 0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_stack_destroy(h)
            self._h = None

    def step(self, dt: float) -> dict:
        r = np.zeros(10)
        if lib().oracle_stack_step(self._h, dt, _ptr(r)) != 0:
            raise ValueError("dt must be positive")
        return dict(zip(REPORT_FIELDS, r.tolist()))

    def observe(self) -> dict:
        ns = [p.n for p in self.layers]
        f = np.zeros(15 * sum(ns))
        sc = np.zeros(10)
        lib().oracle_stack_observe(self._h, f.ctypes.data_as(ctypes.c_void_p), sc.ctypes.data_as(ctypes.c_void_p))
        out = dict(zip(REPORT_FIELDS, sc.tolist()))
        off = 0
        for key in ("h_ex", "h_demag", "h_ext", "h_eff", "rhs"):
            out[key] = []
        for n in ns:
            for key in ("h_ex", "h_demag", "h_ext", "h_eff", "rhs"):
                out[key].append(f[off:off + 3 * n].reshape(n, 3).copy())
                off += 3 * n
        return out
