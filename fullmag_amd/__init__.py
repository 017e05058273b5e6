"""fullmag_amd -- MI355X-native FDM LLG time step for fullmag.

The compute path is libfullmag_fdm.so.0 (HIP, gfx950), built in-tree from
fullmag_amd/csrc and reached through its C ABI (include/fullmag_fdm.h).
This package is the host-side mirror of fullmag's runner wrapper around that
ABI (crates/fullmag-runner/src/native_fdm.rs).
"""

from .native_fdm import (  # noqa: F401
    BatchRun,
    BoundaryGeometry,
    FdmLayerPlan,
    FdmMultilayerPlan,
    FdmPlan,
    FieldSnapshot,
    NativeFdmBackend,
    RunError,
    StepStats,
    compute_newell_kernel_spectra,
    MultilayerStack,
    generate_random_unit_vectors,
    is_available,
)

from .adjoint import heun_rollout  # noqa: F401,E402

MU0 = 4.0 * 3.141592653589793 * 1e-7
