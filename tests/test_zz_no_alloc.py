"""The allocation-free preview entry point (fullmag_fdm_ext_copy_field_preview_noalloc)
against the allocating one it mirrors (fullmag_fdm_backend_copy_field_preview_f64
/ _f32), through the C ABI: identical box averages on two calls in a row, the
second smaller, in both output precisions; the size limit and the unallocated
H_ex / H_eff buffers rejected with an error."""

import ctypes

import numpy as np
import pytest

from fullmag_amd import _abi as A
from fullmag_amd.native_fdm import OBSERVABLES

pytestmark = pytest.mark.gpu


def noalloc(b, obs, cells, z0=0, zs=1, f64=True):
    pnx, pny, pnz = cells
    out = np.full((pnx * pny * pnz, 3), np.nan, dtype=np.float64 if f64 else np.float32)
    rc = A.lib().fullmag_fdm_ext_copy_field_preview_noalloc(
        b._handle, OBSERVABLES[obs], pnx, pny, pnz, z0, zs, 1 if f64 else 0, out.ctypes.data_as(ctypes.c_void_p),
        out.size)
    return rc, out


def backend(fa, cells, prec="mixed"):
    nx, ny, nz = cells
    n = nx * ny * nz
    plan = fa.FdmPlan(cells=cells, cell_size=(4e-9, 5e-9, 3e-9),
                      initial_magnetization=fa.generate_random_unit_vectors(11 + n, n), saturation_magnetisation=8e5,
                      exchange_stiffness=1.3e-11, damping=0.1, external_field=(3e3, -1e3, 2e3), precision=prec,
                      demag_kernel="device")
    return fa.NativeFdmBackend(plan)


@pytest.mark.parametrize("prec", ["mixed", "double"])
def test_noalloc_preview_matches_allocating_preview(fa, prec):
    b = backend(fa, (48, 40, 4), prec)
    for _ in range(3):
        b.step(1e-13)
    b.copy_h_ex()  # allocates the H_ex / H_eff observable buffers
    for f64 in (True, False):
        for obs in ("m", "H_demag", "H_ex", "H_eff", "H_ext"):
            # two calls in a row, the second smaller (and strided in z)
            for cells, z0, zs in (((24, 20, 4), 0, 1), ((6, 5, 1), 1, 2)):
                rc, got = noalloc(b, obs, cells, z0, zs, f64)
                assert rc == A.FULLMAG_FDM_OK, b.last_error_or("noalloc")
                ref = b.copy_live_preview_field(obs, cells, z0, zs, f32=not f64)
                np.testing.assert_array_equal(got, ref)
        # a full-resolution request goes through the same staging
        rc, got = noalloc(b, "m", (48, 40, 4), f64=f64)
        assert rc == A.FULLMAG_FDM_OK
        ref = b.copy_field("m") if f64 else b.copy_live_preview_field("m", (48, 40, 4), f32=True)
        np.testing.assert_array_equal(got, ref)
    b.close()


def test_noalloc_preview_limits(fa):
    b = backend(fa, (256, 256, 4))
    # H_ex / H_eff buffers are allocated by the first copy of either; before
    # that the allocation-free entry point refuses
    rc, _ = noalloc(b, "H_eff", (16, 16, 1))
    assert rc == A.FULLMAG_FDM_ERR_INVALID
    assert "not allocated" in b.last_error_or("").args[0]
    # 3 x 262144 doubles exceed the 4 MiB array; the same preview in FP32 fits
    assert 3 * 256 * 256 * 4 * 8 > A.NOALLOC_SCRATCH_BYTES >= 3 * 256 * 256 * 4 * 4
    rc, _ = noalloc(b, "m", (256, 256, 4), f64=True)
    assert rc == A.FULLMAG_FDM_ERR_INVALID
    assert "scratch" in b.last_error_or("").args[0]
    rc, got = noalloc(b, "m", (256, 256, 4), f64=False)
    assert rc == A.FULLMAG_FDM_OK
    np.testing.assert_array_equal(got, b.copy_live_preview_field("m", (256, 256, 4), f32=True))
    # after the refusals a normal call works and matches
    rc, got = noalloc(b, "H_demag", (32, 32, 2), f64=True)
    assert rc == A.FULLMAG_FDM_OK
    np.testing.assert_array_equal(got, b.copy_live_preview_field("H_demag", (32, 32, 2)))
    b.close()
