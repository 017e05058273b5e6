"""C-ABI library checks that need no GPU: it loads, exports every entry point
declared in include/*.h, and its struct layouts match the reference layout
(crates/fullmag-fdm-sys/src/lib.rs:67-184)."""

import ctypes
import os
import re

import pytest

from tests.conftest import ROOT
from fullmag_amd import _abi as A


def declared_functions():
    names = []
    for h in ("fullmag_fdm.h", "fullmag_fdm_ext.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"\b(fullmag_fdm_\w+)\s*\(", src)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    L = A.lib()
    names = declared_functions()
    assert len(names) >= 16
    for name in names:
        assert hasattr(L, name), name
        assert name in A.SIGNATURES, f"{name} has no ctypes signature"


def test_reference_abi_has_exactly_sixteen_entry_points():
    src = open(os.path.join(ROOT, "include", "fullmag_fdm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    assert len(set(re.findall(r"\b(fullmag_fdm_\w+)\s*\(", src))) == 16


@pytest.mark.parametrize("cls,size", [(A.GridDesc, 40), (A.MaterialDesc, 32), (A.PlanDesc, 448),
                                      (A.StepStats, 96), (A.DeviceInfo, 144), (A.SnapshotDesc, 24)])
def test_struct_sizes_match_header(cls, size):
    assert ctypes.sizeof(cls) == size


def test_plan_desc_offsets():
    assert A.PlanDesc.external_field_am.offset == 96
    assert A.PlanDesc.demag_kernel_xx_spectrum.offset == 120
    assert A.PlanDesc.active_mask.offset == 176
    assert A.PlanDesc.boundary_correction.offset == 224
    assert A.PlanDesc.initial_magnetization_xyz.offset == 400
    assert A.PlanDesc.adaptive_max_error.offset == 416


def test_null_handle_conventions():
    # fullmag_fdm.h: create(NULL) -> NULL, last_error(NULL) == "null handle",
    # destroy/snapshot_destroy are NULL-safe, calls on NULL give ERR_INVALID.
    L = A.lib()
    assert L.fullmag_fdm_backend_create(None) is None
    assert L.fullmag_fdm_backend_last_error(None) == b"null handle"
    L.fullmag_fdm_backend_destroy(None)
    L.fullmag_fdm_field_snapshot_destroy(None)
    s = A.StepStats()
    assert L.fullmag_fdm_backend_step(None, 1e-13, ctypes.byref(s)) == A.FULLMAG_FDM_ERR_INVALID
    assert L.fullmag_fdm_backend_refresh_observables(None) == A.FULLMAG_FDM_ERR_INVALID


def test_random_unit_vectors_match_oracle_bit_for_bit(oracle_mod):
    import numpy as np

    from fullmag_amd import generate_random_unit_vectors

    for seed in (1, 42, 2 ** 63 + 5):
        np.testing.assert_array_equal(generate_random_unit_vectors(seed, 500),
                                      oracle_mod.random_unit_vectors(seed, 500))


def test_extension_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of fullmag_fdm_ext_options and
    fullmag_fdm_ext_stack_desc against the C compiler's layout of the header."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fullmag_fdm_ext.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(fullmag_fdm_ext_options),\n'
        "         sizeof(fullmag_fdm_ext_stack_desc), offsetof(fullmag_fdm_ext_stack_desc, layers),\n"
        "         offsetof(fullmag_fdm_ext_stack_desc, layer_origins), offsetof(fullmag_fdm_ext_stack_desc, common_cells),\n"
        "         offsetof(fullmag_fdm_ext_stack_desc, convolution_cell_size),\n"
        "         offsetof(fullmag_fdm_ext_stack_desc, enable_demag),\n"
        "         offsetof(fullmag_fdm_ext_stack_desc, precision_override));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    S = A.StackDesc
    want = [ctypes.sizeof(A.ExtOptions), ctypes.sizeof(S), S.layers.offset, S.layer_origins.offset,
            S.common_cells.offset, S.convolution_cell_size.offset, S.enable_demag.offset,
            S.precision_override.offset]
    assert got == want
