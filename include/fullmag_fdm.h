/*
 * fullmag_fdm.h -- C ABI of the MI355X-native FDM LLG backend.
 *
 * This header is the drop-in boundary.  It declares the same 16 entry points
 * and the same struct layouts as fullmag's native/include/fullmag_fdm.h, so
 * crates/fullmag-fdm-sys (src/lib.rs:12-288) binds libfullmag_fdm.so.0 built
 * from this repository without any change on the Rust side.  Field order,
 * widths and enum values are fixed by that binding; the comments here are
 * this implementation's own.
 *
 * Layout contract (checked by static_asserts in fullmag_amd/csrc/api.cpp and
 * by tests/test_abi_cpu.py):
 *   fullmag_fdm_grid_desc       40 bytes
 *   fullmag_fdm_material_desc   32 bytes
 *   fullmag_fdm_plan_desc      448 bytes
 *   fullmag_fdm_step_stats      96 bytes
 *   fullmag_fdm_device_info    144 bytes
 *   fullmag_fdm_snapshot_desc   24 bytes
 */

#ifndef FULLMAG_FDM_H
#define FULLMAG_FDM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (reference fullmag_fdm.h:37-40).  ERR_CUDA keeps its name for
 * ABI compatibility; here it reports a HIP runtime / device failure. */
#define FULLMAG_FDM_OK            0
#define FULLMAG_FDM_ERR_INVALID  -1
#define FULLMAG_FDM_ERR_CUDA     -2
#define FULLMAG_FDM_ERR_INTERNAL -3

/* Size of the square inter-region exchange table (reference :43). */
#define FULLMAG_FDM_MAX_EXCHANGE_REGIONS 256

typedef enum {
    FULLMAG_FDM_PRECISION_SINGLE = 1,
    FULLMAG_FDM_PRECISION_DOUBLE = 2,
} fullmag_fdm_precision;

typedef enum {
    FULLMAG_FDM_INTEGRATOR_HEUN = 1,
    FULLMAG_FDM_INTEGRATOR_DP45 = 2,
    FULLMAG_FDM_INTEGRATOR_ABM3 = 3,
    FULLMAG_FDM_INTEGRATOR_RK4  = 4,
    FULLMAG_FDM_INTEGRATOR_RK23 = 5,
} fullmag_fdm_integrator;

typedef enum {
    FULLMAG_FDM_OBSERVABLE_M        = 1,
    FULLMAG_FDM_OBSERVABLE_H_EX     = 2,
    FULLMAG_FDM_OBSERVABLE_H_DEMAG  = 3,
    FULLMAG_FDM_OBSERVABLE_H_EXT    = 4,
    FULLMAG_FDM_OBSERVABLE_H_EFF    = 5,
} fullmag_fdm_observable;

typedef enum {
    FULLMAG_FDM_SNAPSHOT_SCALAR_F32 = 1,
    FULLMAG_FDM_SNAPSHOT_SCALAR_F64 = 2,
} fullmag_fdm_snapshot_scalar_type;

typedef enum {
    FULLMAG_FDM_BOUNDARY_NONE   = 0,
    FULLMAG_FDM_BOUNDARY_VOLUME = 1,
    FULLMAG_FDM_BOUNDARY_FULL   = 2,
} fullmag_fdm_boundary_correction;

/* Grid: cell counts and cell edge lengths in metres. */
typedef struct {
    uint32_t nx;
    uint32_t ny;
    uint32_t nz;
    double   dx;
    double   dy;
    double   dz;
} fullmag_fdm_grid_desc;

/* Material (SI): Ms [A/m], A [J/m], alpha, gamma [m/(A s)]. */
typedef struct {
    double saturation_magnetisation;
    double exchange_stiffness;
    double damping;
    double gyromagnetic_ratio;
} fullmag_fdm_material_desc;

/*
 * Executable plan (reference fullmag_fdm.h:97-198).  Every pointer is
 * borrowed for the duration of fullmag_fdm_backend_create only.
 */
typedef struct {
    fullmag_fdm_grid_desc      grid;
    fullmag_fdm_material_desc  material;
    fullmag_fdm_precision      precision;
    fullmag_fdm_integrator     integrator;
    int                        disable_precession; /* 1: damping-only RHS */
    int                        enable_exchange;
    int                        enable_demag;
    int                        has_external_field;
    double                     external_field_am[3]; /* A/m */

    /* Newell tensor spectra, interleaved complex f64, each of
     * demag_kernel_spectrum_len doubles = 2 * padded cell count, where the
     * padded grid is (2nx, 2ny, 2nz), or (2nx, 2ny) for nz == 1 thin films.
     * All six NULL with length 0: k-space projection fallback. */
    const double              *demag_kernel_xx_spectrum;
    const double              *demag_kernel_yy_spectrum;
    const double              *demag_kernel_zz_spectrum;
    const double              *demag_kernel_xy_spectrum;
    const double              *demag_kernel_xz_spectrum;
    const double              *demag_kernel_yz_spectrum;
    uint64_t                   demag_kernel_spectrum_len;

    /* 1 = active, 0 = inactive; length = cell count. */
    const uint8_t             *active_mask;
    uint64_t                   active_mask_len;

    /* Region id per cell (< FULLMAG_FDM_MAX_EXCHANGE_REGIONS). */
    const uint32_t            *region_mask;
    uint64_t                   region_mask_len;

    /* Row-major MAX_EXCHANGE_REGIONS^2 table of A_ij [J/m].  NULL with a
     * region mask present: A_ii = exchange_stiffness, A_ij = 0 otherwise. */
    const double              *exchange_lut;
    uint64_t                   exchange_lut_len;

    /* Sub-cell boundary tiers (accepted for layout compatibility). */
    fullmag_fdm_boundary_correction boundary_correction;
    double                     boundary_phi_floor;
    double                     boundary_delta_min;
    const double              *volume_fraction;
    uint64_t                   volume_fraction_len;
    const double              *face_link_xp;
    const double              *face_link_xm;
    const double              *face_link_yp;
    const double              *face_link_ym;
    const double              *face_link_zp;
    const double              *face_link_zm;
    const double              *delta_xp;
    const double              *delta_xm;
    const double              *delta_yp;
    const double              *delta_ym;
    const double              *delta_zp;
    const double              *delta_zm;
    int                        has_demag_boundary_corr;
    const int32_t             *demag_corr_target_idx;
    const int32_t             *demag_corr_source_idx;
    const double              *demag_corr_tensor;
    uint32_t                   demag_corr_target_count;
    uint32_t                   demag_corr_stencil_size;

    /* Initial magnetisation, AoS [m0x, m0y, m0z, m1x, ...], 3 * cell count. */
    const double              *initial_magnetization_xyz;
    uint64_t                   initial_magnetization_len;

    /* Adaptive stepping (DP45, RK23); 0 selects 1e-5 / 1e-18 / 1e-10 / 0.8. */
    double                     adaptive_max_error;
    double                     adaptive_dt_min;
    double                     adaptive_dt_max;
    double                     adaptive_headroom;
} fullmag_fdm_plan_desc;

/* Diagnostics of the state reached by the step (reference :202-215). */
typedef struct {
    uint64_t step;
    double   time_seconds;
    double   dt_seconds;
    double   exchange_energy_joules;
    double   demag_energy_joules;
    double   external_energy_joules;
    double   total_energy_joules;
    double   max_effective_field_amplitude;
    double   max_demag_field_amplitude;
    double   max_rhs_amplitude;
    double   suggested_next_dt;   /* adaptive integrators only, else 0 */
    uint64_t wall_time_ns;
} fullmag_fdm_step_stats;

typedef struct {
    char name[128];
    int  compute_capability_major;  /* gfx major, e.g. 9 for gfx950 */
    int  compute_capability_minor;  /* gfx minor/stepping, e.g. 50 */
    int  driver_version;
    int  runtime_version;
} fullmag_fdm_device_info;

typedef struct {
    uint64_t cell_count;
    uint32_t component_count;
    uint32_t scalar_bytes;
    fullmag_fdm_snapshot_scalar_type scalar_type;
} fullmag_fdm_snapshot_desc;

typedef struct fullmag_fdm_backend fullmag_fdm_backend;
typedef struct fullmag_fdm_field_snapshot fullmag_fdm_field_snapshot;

/* 1 when a usable gfx950 device is visible (honours FULLMAG_FDM_GPU_INDEX /
 * FULLMAG_CUDA_DEVICE_INDEX), else 0.  Replaces fullmag_fdm.h:245. */
int fullmag_fdm_is_available(void);

/* Allocates device state, ingests the plan, evaluates the fields at m0.
 * NULL only for a NULL plan or host OOM; any other failure returns a handle
 * whose last_error is set.  Replaces fullmag_fdm.h:252-253. */
fullmag_fdm_backend *fullmag_fdm_backend_create(const fullmag_fdm_plan_desc *plan);

/* One step of the configured integrator; stats describe the new state.
 * Replaces fullmag_fdm.h:260-263. */
int fullmag_fdm_backend_step(fullmag_fdm_backend *handle, double dt_seconds,
                             fullmag_fdm_step_stats *out_stats);

/* Field download, AoS, out_len = 3 * cell count.  Replaces :269-282. */
int fullmag_fdm_backend_copy_field_f64(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                       double *out_xyz, uint64_t out_len);
int fullmag_fdm_backend_copy_field_f32(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                       float *out_xyz, uint64_t out_len);

/* Box-averaged preview on a (pnx, pny, pnz) grid; z bins start at z_origin
 * and span z_stride layers.  Replaces :290-315. */
int fullmag_fdm_backend_copy_field_preview_f64(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                               uint32_t z_origin, uint32_t z_stride,
                                               double *out_xyz, uint64_t out_len);
int fullmag_fdm_backend_copy_field_preview_f32(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                               uint32_t z_origin, uint32_t z_stride,
                                               float *out_xyz, uint64_t out_len);

/* Asynchronous SoA snapshot [x..][y..][z..] in the backend's storage
 * precision, staged on the device and copied to pinned memory on a private
 * stream; wait may run on another host thread.  Replaces :331-359. */
fullmag_fdm_field_snapshot *fullmag_fdm_backend_begin_field_snapshot(fullmag_fdm_backend *handle,
                                                                     fullmag_fdm_observable observable);
int fullmag_fdm_field_snapshot_wait(fullmag_fdm_field_snapshot *snapshot, const void **out_data,
                                    uint64_t *out_len_bytes, fullmag_fdm_snapshot_desc *out_desc);
void fullmag_fdm_field_snapshot_destroy(fullmag_fdm_field_snapshot *snapshot);

/* Replace m (AoS); inactive cells forced to zero; time does not advance and
 * field observables keep their previous values until refresh or the next
 * step.  Replaces :366-384. */
int fullmag_fdm_backend_upload_magnetization_f64(fullmag_fdm_backend *handle, const double *m_xyz, uint64_t len);
int fullmag_fdm_backend_upload_magnetization_f32(fullmag_fdm_backend *handle, const float *m_xyz, uint64_t len);

/* Re-evaluate the field observables at the current m.  Replaces :390-391. */
int fullmag_fdm_backend_refresh_observables(fullmag_fdm_backend *handle);

int fullmag_fdm_backend_get_device_info(fullmag_fdm_backend *handle, fullmag_fdm_device_info *out_info);

/* Sticky error text, NULL when none, "null handle" for NULL.  Replaces :403. */
const char *fullmag_fdm_backend_last_error(fullmag_fdm_backend *handle);

/* Frees all device memory; NULL-safe.  Replaces :409. */
void fullmag_fdm_backend_destroy(fullmag_fdm_backend *handle);

#ifdef __cplusplus
}
#endif

#endif /* FULLMAG_FDM_H */
