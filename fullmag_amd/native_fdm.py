"""Host-side mirror of fullmag's native FDM backend wrapper.

`NativeFdmBackend` follows crates/fullmag-runner/src/native_fdm.rs:44-800
(create / step / copy_* / upload / refresh / device_info / snapshots) and
drives libfullmag_fdm.so.0 through the C ABI only; `FdmPlan` carries the
FdmPlanIR fields that the runner turns into a fullmag_fdm_plan_desc
(crates/fullmag-ir/src/lib.rs:662-695, native_fdm.rs:90-348).

Errors surface as `RunError` with the backend's last_error text, as the
runner's `last_error_or` does (native_fdm.rs:696-706).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi as A

MU0 = 4.0 * np.pi * 1e-7

_PRECISIONS = {"single": A.PRECISION_SINGLE, "double": A.PRECISION_DOUBLE, "mixed": A.PRECISION_MIXED}
_INTEGRATORS = {"heun": A.INTEGRATOR_HEUN, "rk45": A.INTEGRATOR_DP45, "dp45": A.INTEGRATOR_DP45,
                "abm3": A.INTEGRATOR_ABM3, "rk4": A.INTEGRATOR_RK4, "rk23": A.INTEGRATOR_RK23}
OBSERVABLES = {"m": A.OBSERVABLE_M, "H_ex": A.OBSERVABLE_H_EX, "H_demag": A.OBSERVABLE_H_DEMAG,
               "H_ext": A.OBSERVABLE_H_EXT, "H_eff": A.OBSERVABLE_H_EFF}


class RunError(RuntimeError):
    pass


def is_available() -> bool:
    """fullmag_fdm_is_available() == 1 (native_fdm.rs:31-40)."""
    return A.lib().fullmag_fdm_is_available() == 1


def generate_random_unit_vectors(seed: int, count: int) -> np.ndarray:
    """The planner's seeded m0 (crates/fullmag-plan/src/lib.rs:1927-1957), (count, 3)."""
    out = np.empty((count, 3), dtype=np.float64)
    A.lib().fullmag_fdm_ext_random_unit_vectors(ctypes.c_uint64(seed), count,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def compute_newell_kernel_spectra(nx: int, ny: int, nz: int, dx: float, dy: float, dz: float,
                                  thin_film_2d: Optional[bool] = None) -> np.ndarray:
    """Newell spectra in the plan-descriptor layout, shape (6, 2*Npad).

    Same contract as fullmag_engine::compute_newell_kernel_spectra and
    compute_newell_kernel_spectra_thin_film_2d (crates/fullmag-engine/src/
    lib.rs:511-610), chosen as the runner does for nz == 1
    (native_fdm.rs:150-171); computed on the device.
    """
    if thin_film_2d is None:
        thin_film_2d = nz == 1
    npad = 4 * nx * ny if thin_film_2d else 8 * nx * ny * nz
    out = np.empty((6, 2 * npad), dtype=np.float64)
    rc = A.lib().fullmag_fdm_ext_newell_spectra(nx, ny, nz, dx, dy, dz, 1 if thin_film_2d else 0,
                                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size)
    if rc != A.FULLMAG_FDM_OK:
        raise RunError(f"Newell spectra computation failed (rc={rc})")
    return out


@dataclass
class BoundaryGeometry:
    """Sub-cell boundary geometry (crates/fullmag-ir/src/lib.rs:699-724,
    BoundaryGeometryIR): per-cell volume fraction, face-link fractions and
    intersection distances, each of cell_count values (an empty array is
    passed as NULL), and the sparse demag correction rows: target cell per
    row, stencil_size source cells per row (-1: none) and a symmetric tensor
    per (row, source) as Nxx, Nxy, Nxz, Nyy, Nyz, Nzz."""

    volume_fraction: np.ndarray
    face_link_xp: Optional[np.ndarray] = None
    face_link_xm: Optional[np.ndarray] = None
    face_link_yp: Optional[np.ndarray] = None
    face_link_ym: Optional[np.ndarray] = None
    face_link_zp: Optional[np.ndarray] = None
    face_link_zm: Optional[np.ndarray] = None
    delta_xp: Optional[np.ndarray] = None
    delta_xm: Optional[np.ndarray] = None
    delta_yp: Optional[np.ndarray] = None
    delta_ym: Optional[np.ndarray] = None
    delta_zp: Optional[np.ndarray] = None
    delta_zm: Optional[np.ndarray] = None
    demag_corr_target_idx: Optional[np.ndarray] = None
    demag_corr_source_idx: Optional[np.ndarray] = None
    demag_corr_tensor: Optional[np.ndarray] = None
    demag_corr_stencil_size: int = 0


# FdmPlanIR.boundary_correction strings (native_fdm.rs:250-254)
_BOUNDARY = {"volume": A.BOUNDARY_VOLUME, "full": A.BOUNDARY_FULL}


@dataclass
class FdmPlan:
    """FdmPlanIR subset consumed by the FDM backend."""

    cells: Sequence[int]
    cell_size: Sequence[float]
    initial_magnetization: np.ndarray
    saturation_magnetisation: float
    exchange_stiffness: float
    damping: float
    gyromagnetic_ratio: float = 2.211e5
    precision: str = "double"  # "single" | "double" | "mixed"
    integrator: str = "heun"
    enable_exchange: bool = True
    enable_demag: bool = True
    external_field: Optional[Sequence[float]] = None  # A/m
    active_mask: Optional[np.ndarray] = None
    region_mask: Optional[np.ndarray] = None
    inter_region_exchange: list = field(default_factory=list)  # (ri, rj, A_ij)
    disable_precession: bool = False  # LLG-overdamped relaxation
    adaptive_max_error: float = 0.0
    adaptive_dt_min: float = 0.0
    adaptive_dt_max: float = 0.0
    adaptive_headroom: float = 0.0
    # "spectra": pass Newell spectra through the plan descriptor as the runner
    # does; "device": let the library build the tensor; an explicit (6, 2*Npad)
    # array is passed through as given.
    demag_kernel: object = "device"
    ms_cell: Optional[np.ndarray] = None  # extension: per-cell Ms [A/m]
    # sub-cell boundary tiers: "volume" (T0 face links) or "full" (T1
    # intersection distances), with the planner's geometry
    boundary_correction: Optional[str] = None
    boundary_geometry: Optional[BoundaryGeometry] = None
    boundary_phi_floor: float = 0.0  # 0: the library default 0.05
    boundary_delta_min: float = 0.0  # 0: the library default 0.1 min(dx, dy, dz)

    @property
    def cell_count(self) -> int:
        return int(self.cells[0]) * int(self.cells[1]) * int(self.cells[2])


@dataclass
class StepStats:
    """Per-step diagnostics (crates/fullmag-runner/src/native_fdm.rs:388-407)."""

    step: int
    time: float
    dt: float
    e_ex: float
    e_demag: float
    e_ext: float
    e_total: float
    max_h_eff: float
    max_h_demag: float
    max_dm_dt: float
    wall_time_ns: int
    dt_suggested: Optional[float]

    @staticmethod
    def from_raw(s) -> "StepStats":
        return StepStats(s.step, s.time_seconds, s.dt_seconds, s.exchange_energy_joules, s.demag_energy_joules,
                         s.external_energy_joules, s.total_energy_joules, s.max_effective_field_amplitude,
                         s.max_demag_field_amplitude, s.max_rhs_amplitude, s.wall_time_ns,
                         s.suggested_next_dt if s.suggested_next_dt > 0.0 else None)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _describe(plan: FdmPlan):
    """fullmag_fdm_plan_desc and extension options of a plan, as the runner
    fills the descriptor (native_fdm.rs:90-348), with the host arrays that
    must outlive the create call."""
    n = plan.cell_count
    nx, ny, nz = (int(v) for v in plan.cells)
    dx, dy, dz = (float(v) for v in plan.cell_size)
    keep = []  # host arrays that must outlive the create() call

    d = A.PlanDesc()
    d.grid = A.GridDesc(nx, ny, nz, dx, dy, dz)
    d.material = A.MaterialDesc(plan.saturation_magnetisation, plan.exchange_stiffness, plan.damping,
                                plan.gyromagnetic_ratio)
    prec = _PRECISIONS[plan.precision]
    d.precision = A.PRECISION_SINGLE if prec == A.PRECISION_MIXED else prec
    d.integrator = _INTEGRATORS[plan.integrator]
    d.disable_precession = 1 if plan.disable_precession else 0
    d.enable_exchange = 1 if plan.enable_exchange else 0
    d.enable_demag = 1 if plan.enable_demag else 0
    d.has_external_field = 1 if plan.external_field is not None else 0
    ext = plan.external_field if plan.external_field is not None else (0.0, 0.0, 0.0)
    d.external_field_am = (ctypes.c_double * 3)(*[float(v) for v in ext])

    m0 = np.ascontiguousarray(plan.initial_magnetization, dtype=np.float64).reshape(-1)
    if m0.size != 3 * n:
        raise RunError(f"initial magnetization has {m0.size} values, expected {3 * n}")
    keep.append(m0)
    d.initial_magnetization_xyz = _dp(m0)
    d.initial_magnetization_len = m0.size

    newell_on_device = 0
    if plan.enable_demag:
        spectra = None
        if isinstance(plan.demag_kernel, np.ndarray):
            spectra = np.ascontiguousarray(plan.demag_kernel, dtype=np.float64)
        elif plan.demag_kernel == "spectra":
            spectra = compute_newell_kernel_spectra(nx, ny, nz, dx, dy, dz)
        elif plan.demag_kernel == "device":
            newell_on_device = 1
        elif plan.demag_kernel not in (None, "projection"):
            raise RunError(f"unknown demag_kernel option {plan.demag_kernel!r}")
        if spectra is not None:
            keep.append(spectra)
            ptrs = [spectra[c].ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for c in range(6)]
            (d.demag_kernel_xx_spectrum, d.demag_kernel_yy_spectrum, d.demag_kernel_zz_spectrum,
             d.demag_kernel_xy_spectrum, d.demag_kernel_xz_spectrum, d.demag_kernel_yz_spectrum) = ptrs
            d.demag_kernel_spectrum_len = spectra.shape[1]

    if plan.active_mask is not None:
        mask = np.ascontiguousarray(np.asarray(plan.active_mask).reshape(-1).astype(np.uint8))
        keep.append(mask)
        d.active_mask = mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        d.active_mask_len = mask.size
    if plan.region_mask is not None:
        reg = np.ascontiguousarray(np.asarray(plan.region_mask).reshape(-1).astype(np.uint32))
        keep.append(reg)
        d.region_mask = reg.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        d.region_mask_len = reg.size
        # native_fdm.rs:173-196: diagonal = material A, caller overrides symmetric
        R = A.MAX_EXCHANGE_REGIONS
        lut = np.zeros(R * R, dtype=np.float64)
        lut[np.arange(R) * R + np.arange(R)] = plan.exchange_stiffness
        for ri, rj, aij in plan.inter_region_exchange:
            if ri < R and rj < R:
                lut[ri * R + rj] = aij
                lut[rj * R + ri] = aij
        keep.append(lut)
        d.exchange_lut = _dp(lut)
        d.exchange_lut_len = lut.size
    d.boundary_correction = _BOUNDARY.get(plan.boundary_correction, A.BOUNDARY_NONE)
    d.boundary_phi_floor = plan.boundary_phi_floor
    d.boundary_delta_min = plan.boundary_delta_min
    bg = plan.boundary_geometry
    if bg is not None:
        def f64(name):
            v = getattr(bg, name)
            if v is None or np.asarray(v).size == 0:
                return None
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
            keep.append(a)
            return a

        vf = f64("volume_fraction")
        if vf is not None:
            d.volume_fraction = _dp(vf)
            d.volume_fraction_len = vf.size
        for name in ("face_link_xp", "face_link_xm", "face_link_yp", "face_link_ym", "face_link_zp",
                     "face_link_zm", "delta_xp", "delta_xm", "delta_yp", "delta_ym", "delta_zp", "delta_zm"):
            a = f64(name)
            if a is not None:
                if a.size != n:
                    raise RunError(f"boundary_geometry.{name} has {a.size} values, expected {n}")
                setattr(d, name, _dp(a))
        tgt = bg.demag_corr_target_idx
        if tgt is not None and np.asarray(tgt).size > 0:
            # native_fdm.rs:313-338: the flag follows a non-empty target list
            rows = np.asarray(tgt).size
            st = int(bg.demag_corr_stencil_size)
            t = np.ascontiguousarray(np.asarray(tgt, dtype=np.int32).reshape(-1))
            src = np.ascontiguousarray(np.asarray(bg.demag_corr_source_idx, dtype=np.int32).reshape(-1))
            ten = np.ascontiguousarray(np.asarray(bg.demag_corr_tensor, dtype=np.float64).reshape(-1))
            if src.size != rows * st or ten.size != rows * st * 6:
                raise RunError(f"demag correction arrays do not match {rows} rows x stencil {st}")
            keep += [t, src, ten]
            d.has_demag_boundary_corr = 1
            d.demag_corr_target_idx = t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            d.demag_corr_source_idx = src.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            d.demag_corr_tensor = _dp(ten)
            d.demag_corr_target_count = rows
            d.demag_corr_stencil_size = st
    d.adaptive_max_error = plan.adaptive_max_error
    d.adaptive_dt_min = plan.adaptive_dt_min
    d.adaptive_dt_max = plan.adaptive_dt_max
    d.adaptive_headroom = plan.adaptive_headroom

    opt = A.ExtOptions()
    opt.struct_size = ctypes.sizeof(A.ExtOptions)
    opt.precision_override = prec
    opt.newell_on_device = newell_on_device
    if plan.ms_cell is not None:
        ms = np.ascontiguousarray(plan.ms_cell, dtype=np.float64).reshape(-1)
        keep.append(ms)
        opt.ms_cell = _dp(ms)
        opt.ms_cell_len = ms.size

    return d, opt, keep


class NativeFdmBackend:
    """RAII owner of one backend handle (native_fdm.rs:44-47, 708-715)."""

    def __init__(self, plan: FdmPlan):
        L = A.lib()
        self.plan = plan
        self.cell_count = plan.cell_count
        d, opt, keep = _describe(plan)
        self._handle = L.fullmag_fdm_ext_backend_create(ctypes.byref(d), ctypes.byref(opt))
        if not self._handle:
            raise RunError("fullmag_fdm_backend_create returned NULL")
        err = L.fullmag_fdm_backend_last_error(self._handle)
        if err is not None:
            msg = err.decode()
            L.fullmag_fdm_backend_destroy(self._handle)
            self._handle = None
            raise RunError(msg)
        self.precision = plan.precision

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        h = getattr(self, "_handle", None)
        if h:
            A.lib().fullmag_fdm_backend_destroy(h)
            self._handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._handle

    def last_error_or(self, fallback: str) -> RunError:
        err = A.lib().fullmag_fdm_backend_last_error(self._handle)
        return RunError(err.decode() if err else fallback)

    # -- stepping ----------------------------------------------------------
    def step(self, dt: float) -> StepStats:
        s = A.StepStats()
        rc = A.lib().fullmag_fdm_backend_step(self._handle, dt, ctypes.byref(s))
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("step failed")
        return StepStats.from_raw(s)

    def step_many(self, dt: float, n_steps: int) -> StepStats:
        s = A.StepStats()
        rc = A.lib().fullmag_fdm_ext_step_many(self._handle, dt, n_steps, ctypes.byref(s))
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("step_many failed")
        return StepStats.from_raw(s)

    # -- fields ------------------------------------------------------------
    def copy_field(self, observable: str) -> np.ndarray:
        out = np.empty((self.cell_count, 3), dtype=np.float64)
        rc = A.lib().fullmag_fdm_backend_copy_field_f64(self._handle, OBSERVABLES[observable], _dp(out), out.size)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or(f"copy_field({observable}) failed")
        return out

    def copy_field_f32(self, observable: str) -> np.ndarray:
        out = np.empty((self.cell_count, 3), dtype=np.float32)
        rc = A.lib().fullmag_fdm_backend_copy_field_f32(
            self._handle, OBSERVABLES[observable], out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or(f"copy_field_f32({observable}) failed")
        return out

    def copy_m(self):
        return self.copy_field("m")

    def copy_h_ex(self):
        return self.copy_field("H_ex")

    def copy_h_demag(self):
        return self.copy_field("H_demag")

    def copy_h_ext(self):
        return self.copy_field("H_ext")

    def copy_h_eff(self):
        return self.copy_field("H_eff")

    def copy_live_preview_field(self, observable: str, preview_cells: Sequence[int], z_origin: int = 0,
                                z_stride: int = 1, f32: bool = False) -> np.ndarray:
        pnx, pny, pnz = (int(v) for v in preview_cells)
        dt = np.float32 if f32 else np.float64
        out = np.empty((pnx * pny * pnz, 3), dtype=dt)
        L = A.lib()
        if f32:
            rc = L.fullmag_fdm_backend_copy_field_preview_f32(
                self._handle, OBSERVABLES[observable], pnx, pny, pnz, z_origin, z_stride,
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size)
        else:
            rc = L.fullmag_fdm_backend_copy_field_preview_f64(
                self._handle, OBSERVABLES[observable], pnx, pny, pnz, z_origin, z_stride, _dp(out), out.size)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("preview copy failed")
        return out

    def begin_field_snapshot(self, observable: str) -> "FieldSnapshot":
        h = A.lib().fullmag_fdm_backend_begin_field_snapshot(self._handle, OBSERVABLES[observable])
        if not h:
            raise self.last_error_or("snapshot failed")
        return FieldSnapshot(h)

    def upload_magnetization(self, m: np.ndarray) -> None:
        a = np.ascontiguousarray(m, dtype=np.float64).reshape(-1)
        rc = A.lib().fullmag_fdm_backend_upload_magnetization_f64(self._handle, _dp(a), a.size)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("upload failed")

    def upload_magnetization_f32(self, m: np.ndarray) -> None:
        a = np.ascontiguousarray(m, dtype=np.float32).reshape(-1)
        rc = A.lib().fullmag_fdm_backend_upload_magnetization_f32(
            self._handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("upload failed")

    def refresh_observables(self) -> None:
        rc = A.lib().fullmag_fdm_backend_refresh_observables(self._handle)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("refresh failed")

    def set_external_field(self, h_am: Sequence[float]) -> None:
        h = (ctypes.c_double * 3)(*[float(v) for v in h_am])
        rc = A.lib().fullmag_fdm_ext_set_external_field(self._handle, h)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("set_external_field failed")

    def device_info(self) -> dict:
        info = A.DeviceInfo()
        rc = A.lib().fullmag_fdm_backend_get_device_info(self._handle, ctypes.byref(info))
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("device info failed")
        return {"name": info.name.decode(errors="replace"), "compute_capability": (
            info.compute_capability_major, info.compute_capability_minor),
            "driver_version": info.driver_version, "runtime_version": info.runtime_version}

    def precision_mode(self) -> int:
        return A.lib().fullmag_fdm_ext_precision(self._handle)

    def autotune(self, reps: int = 5) -> str:
        """Pick the fastest launch schedule for this backend's grid (see
        fullmag_fdm_ext_autotune); returns it as "key=value,..."."""
        buf = ctypes.create_string_buffer(256)
        rc = A.lib().fullmag_fdm_ext_autotune(self._handle, int(reps), buf, 256)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("autotune failed")
        return buf.value.decode()

    def set_schedule(self, schedule: str) -> None:
        """Apply a schedule string as autotune() returns it."""
        rc = A.lib().fullmag_fdm_ext_set_schedule(self._handle, schedule.encode())
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("set_schedule failed")

    # -- device-resident state and the Heun adjoint (extension) -------------
    # Device arrays are passed as raw pointers (e.g. torch.Tensor.data_ptr()
    # of contiguous float64 tensors of 3 * cells in SoA order, or cells for
    # grad_ms); the caller orders its own device work before these calls.
    def copy_state_device(self, dst_ptr: int) -> None:
        rc = A.lib().fullmag_fdm_ext_copy_state_device(self._handle, ctypes.c_void_p(dst_ptr), 3 * self.cell_count)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("copy_state_device failed")

    def set_state_device(self, src_ptr: int) -> None:
        rc = A.lib().fullmag_fdm_ext_set_state_device(self._handle, ctypes.c_void_p(src_ptr), 3 * self.cell_count)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("set_state_device failed")

    def heun_adjoint(self, dt: float, m0_ptr: int, lam1_ptr: int, lam0_ptr: int, grad_ms_ptr: int = 0,
                     grad_hext=None) -> None:
        """lam0 = (dm1/dm0)^T lam1 for one Heun step from m0; accumulates
        the per-cell Ms gradient (device) and the H_ext gradient (numpy (3,),
        A/m) of <lam1, m1>."""
        gh = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
        rc = A.lib().fullmag_fdm_ext_heun_adjoint(self._handle, float(dt), ctypes.c_void_p(m0_ptr),
                                                  ctypes.c_void_p(lam1_ptr), ctypes.c_void_p(lam0_ptr),
                                                  ctypes.c_void_p(grad_ms_ptr or None), gh)
        if rc != A.FULLMAG_FDM_OK:
            raise self.last_error_or("heun_adjoint failed")
        if grad_hext is not None:
            grad_hext += np.array(gh[:])

    def device_bytes(self) -> int:
        return int(A.lib().fullmag_fdm_ext_device_bytes(self._handle))

    # -- kernel timing (extension) -------------------------------------------
    def set_timing(self, enable: bool) -> None:
        A.lib().fullmag_fdm_ext_set_timing(self._handle, 1 if enable else 0)

    def timings(self) -> dict:
        L = A.lib()
        out = {}
        i = 0
        name = ctypes.c_char_p()
        ms = ctypes.c_double()
        cnt = ctypes.c_uint64()
        while L.fullmag_fdm_ext_get_timing(self._handle, i, ctypes.byref(name), ctypes.byref(ms),
                                           ctypes.byref(cnt)) == A.FULLMAG_FDM_OK:
            out[name.value.decode()] = (ms.value, cnt.value)
            i += 1
        return out


class BatchRun:
    """Independent Heun backends advanced together through one hipGraph
    (fullmag_fdm_ext_batch_create / _run).  The members stay owned by the
    caller and must not be stepped outside the batch while it exists."""

    def __init__(self, backends: Sequence["NativeFdmBackend"], dts: Sequence[float], steps_per_graph: int = 10):
        self.backends = list(backends)
        n = len(self.backends)
        handles = (ctypes.c_void_p * n)(*[b.handle for b in self.backends])
        d = np.ascontiguousarray(dts, dtype=np.float64)
        if d.size != n:
            raise ValueError("one dt per backend")
        self._h = A.lib().fullmag_fdm_ext_batch_create(handles, n, _dp(d), int(steps_per_graph))
        if not self._h:
            raise RunError("batch_create returned NULL")
        err = A.lib().fullmag_fdm_ext_batch_last_error(self._h)
        if err:
            msg = err.decode(errors="replace")
            self.close()
            raise RunError(msg)

    def run(self, n_steps: int) -> list:
        stats = (A.StepStats * len(self.backends))()
        rc = A.lib().fullmag_fdm_ext_batch_run(self._h, int(n_steps), stats)
        if rc != A.FULLMAG_FDM_OK:
            err = A.lib().fullmag_fdm_ext_batch_last_error(self._h)
            raise RunError(err.decode(errors="replace") if err else f"batch_run failed ({rc})")
        return [StepStats.from_raw(s) for s in stats]

    def close(self) -> None:
        if getattr(self, "_h", None):
            A.lib().fullmag_fdm_ext_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


@dataclass
class FdmLayerPlan:
    """FdmLayerPlanIR (crates/fullmag-ir/src/lib.rs:414-427): one magnet of a
    multilayer plan on its own grid."""

    magnet_name: str
    native_grid: Sequence[int]
    native_cell_size: Sequence[float]
    native_origin: Sequence[float]
    initial_magnetization: np.ndarray
    saturation_magnetisation: float
    exchange_stiffness: float
    damping: float
    native_active_mask: Optional[np.ndarray] = None


@dataclass
class FdmMultilayerPlan:
    """FdmMultilayerPlanIR (crates/fullmag-ir/src/lib.rs:395-411): layers
    coupled by demag on the common convolution grid `common_cells` of cell
    size `convolution_cell_size`, stepped together with Heun."""

    common_cells: Sequence[int]
    convolution_cell_size: Sequence[float]
    layers: list
    enable_exchange: bool = True
    enable_demag: bool = True
    external_field: Optional[Sequence[float]] = None  # A/m, every layer
    gyromagnetic_ratio: float = 2.211e5
    precision: str = "double"
    disable_precession: bool = False


class _StackLayer(NativeFdmBackend):
    """A layer of a MultilayerStack: the backend handle belongs to the stack."""

    def __init__(self, handle, plan: FdmPlan, stack):
        self._handle = handle
        self.plan = plan
        self.cell_count = plan.cell_count
        self.precision = plan.precision
        self._stack = stack  # keeps the stack, and so the handle, alive

    def close(self) -> None:
        self._handle = None


class MultilayerStack:
    """A multilayer plan on the device (fullmag_fdm_ext_stack_*): one backend
    per layer for exchange, Zeeman, torque and stats, the demag coupling of
    all layers (self and shifted kernels, push/pull transfers) and Heun steps
    of the whole stack, as crates/fullmag-runner/src/multilayer_reference.rs
    advances a FdmMultilayerPlanIR."""

    def __init__(self, plan: FdmMultilayerPlan):
        L = A.lib()
        self.plan = plan
        self.layer_plans = [
            FdmPlan(cells=tuple(int(v) for v in lp.native_grid), cell_size=tuple(lp.native_cell_size),
                    initial_magnetization=lp.initial_magnetization,
                    saturation_magnetisation=lp.saturation_magnetisation,
                    exchange_stiffness=lp.exchange_stiffness, damping=lp.damping,
                    gyromagnetic_ratio=plan.gyromagnetic_ratio, precision=plan.precision, integrator="heun",
                    enable_exchange=plan.enable_exchange, enable_demag=False,
                    external_field=plan.external_field, active_mask=lp.native_active_mask,
                    disable_precession=plan.disable_precession)
            for lp in plan.layers]
        described = [_describe(p) for p in self.layer_plans]
        n = len(described)
        ptrs = (ctypes.POINTER(A.PlanDesc) * n)(*[ctypes.pointer(d) for d, _, _ in described])
        origins = np.ascontiguousarray([list(lp.native_origin) for lp in plan.layers], dtype=np.float64)
        sd = A.StackDesc()
        sd.struct_size = ctypes.sizeof(A.StackDesc)
        sd.layer_count = n
        sd.layers = ptrs
        sd.layer_origins = _dp(origins)
        sd.common_cells = (ctypes.c_uint32 * 3)(*[int(v) for v in plan.common_cells])
        sd.convolution_cell_size = (ctypes.c_double * 3)(*[float(v) for v in plan.convolution_cell_size])
        sd.enable_demag = 1 if plan.enable_demag else 0
        sd.precision_override = _PRECISIONS[plan.precision]
        self._h = L.fullmag_fdm_ext_stack_create(ctypes.byref(sd))
        if not self._h:
            raise RunError("fullmag_fdm_ext_stack_create returned NULL")
        err = L.fullmag_fdm_ext_stack_last_error(self._h)
        if err is not None:
            msg = err.decode(errors="replace")
            self.close()
            raise RunError(msg)
        self.layers = [_StackLayer(L.fullmag_fdm_ext_stack_layer(self._h, i), p, self)
                       for i, p in enumerate(self.layer_plans)]

    def step(self, dt: float, n_steps: int = 1) -> StepStats:
        s = A.StepStats()
        rc = A.lib().fullmag_fdm_ext_stack_step(self._h, float(dt), int(n_steps), ctypes.byref(s))
        if rc != A.FULLMAG_FDM_OK:
            err = A.lib().fullmag_fdm_ext_stack_last_error(self._h)
            raise RunError(err.decode(errors="replace") if err else f"stack_step failed ({rc})")
        return StepStats.from_raw(s)

    def refresh(self) -> None:
        if A.lib().fullmag_fdm_ext_stack_refresh(self._h) != A.FULLMAG_FDM_OK:
            err = A.lib().fullmag_fdm_ext_stack_last_error(self._h)
            raise RunError(err.decode(errors="replace") if err else "stack_refresh failed")

    def close(self) -> None:
        if getattr(self, "_h", None):
            for layer in getattr(self, "layers", []):
                layer._handle = None
            A.lib().fullmag_fdm_ext_stack_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class FieldSnapshot:
    """Async SoA snapshot (native_fdm.rs:718-800)."""

    def __init__(self, handle):
        self._h = handle

    def wait(self) -> np.ndarray:
        data = ctypes.c_void_p()
        nbytes = ctypes.c_uint64()
        desc = A.SnapshotDesc()
        rc = A.lib().fullmag_fdm_field_snapshot_wait(self._h, ctypes.byref(data), ctypes.byref(nbytes),
                                                    ctypes.byref(desc))
        if rc != A.FULLMAG_FDM_OK:
            raise RunError("snapshot wait failed")
        dt = np.float32 if desc.scalar_type == A.SNAPSHOT_F32 else np.float64
        buf = (ctypes.c_char * nbytes.value).from_address(data.value)
        return np.frombuffer(buf, dtype=dt).reshape(3, desc.cell_count).copy()

    def close(self):
        if self._h:
            A.lib().fullmag_fdm_field_snapshot_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
