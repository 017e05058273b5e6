/*
 * fullmag_fdm_ext.h -- additive extensions of the FDM C ABI.
 *
 * Nothing here changes the 16 entry points or the struct layouts of
 * fullmag_fdm.h; callers that only know fullmag_fdm.h are unaffected.  None
 * of these entry points has a counterpart in the reference ABI
 * (native/include/fullmag_fdm.h); each comment names the caller it serves.
 */

#ifndef FULLMAG_FDM_EXT_H
#define FULLMAG_FDM_EXT_H

#include "fullmag_fdm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mixed precision: FP64 magnetisation state, renormalisation, Newell tensor
 * and adaptive error norm; FP32 exchange, FFT demag and torque; compensated
 * FP64 energy sums.  Selectable through fullmag_fdm_ext_options, or for the
 * unchanged Rust runner by precision SINGLE plus FULLMAG_FDM_MIXED=1. */
#define FULLMAG_FDM_PRECISION_MIXED 3

typedef struct {
    uint32_t struct_size;          /* sizeof(fullmag_fdm_ext_options) */
    int32_t  precision_override;   /* 0: use plan->precision; else 1, 2 or 3 */
    int32_t  newell_on_device;     /* 1: when the plan carries no spectra and
                                      demag is enabled, build the Newell tensor
                                      in the library instead of the k-space
                                      projection fallback */
    int32_t  reserved;
    const double *ms_cell;         /* optional per-cell Ms [A/m], cell count */
    uint64_t ms_cell_len;
} fullmag_fdm_ext_options;

/* create() with extension options.  For Python/C callers that need mixed
 * precision, per-cell Ms or the in-library Newell build. */
fullmag_fdm_backend *fullmag_fdm_ext_backend_create(const fullmag_fdm_plan_desc *plan,
                                                    const fullmag_fdm_ext_options *options);

/* Precision mode in use (1 single, 2 double, 3 mixed), or ERR_INVALID. */
int fullmag_fdm_ext_precision(fullmag_fdm_backend *handle);

/* Newell tensor spectra in the layout fullmag_fdm_plan_desc expects: six
 * interleaved complex f64 arrays back to back (xx, yy, zz, xy, xz, yz), each
 * 2 * (2nx)(2ny)(2nz) doubles, or 2 * (2nx)(2ny) when thin_film_2d (nz must
 * be 1).  Computed on the device.  For callers that build a plan_desc
 * themselves, as crates/fullmag-runner/src/native_fdm.rs:150-171 does. */
int fullmag_fdm_ext_newell_spectra(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double dz,
                                   int thin_film_2d, double *out, uint64_t out_len);

/* Replace the uniform external field [A/m] (field sweeps, finite-difference
 * checks) without re-creating the backend. */
int fullmag_fdm_ext_set_external_field(fullmag_fdm_backend *handle, const double h_am[3]);

/* n_steps fixed-dt steps with one host synchronisation at the end; stats of
 * the final state.  For throughput callers that do not need per-step stats. */
int fullmag_fdm_ext_step_many(fullmag_fdm_backend *handle, double dt_seconds, uint64_t n_steps,
                              fullmag_fdm_step_stats *out_stats);

/* Per-kernel device timing with HIP events on the backend's stream.  When
 * enabled, every kernel launch is bracketed by events; get_timing returns the
 * accumulated milliseconds and launch count of entry `index` (ERR_INVALID
 * past the last entry).  For benchmarks and roofline reports. */
int fullmag_fdm_ext_set_timing(fullmag_fdm_backend *handle, int enable);
int fullmag_fdm_ext_get_timing(fullmag_fdm_backend *handle, int index, const char **name,
                               double *total_ms, uint64_t *launches);

/* Seeded random unit vectors, AoS, exactly as the planner's m0 generator
 * (crates/fullmag-plan/src/lib.rs:1927-1957: xorshift64 <<13 >>7 <<17 per
 * component, rejection below norm 1e-10).  Host-side; for callers building
 * the initial magnetisation of a plan. */
void fullmag_fdm_ext_random_unit_vectors(uint64_t seed, uint64_t count, double *out_xyz);

/* Discrete adjoint of one Heun step, for gradient-based inverse design
 * (torch.autograd callers).  No counterpart in the reference ABI.  m0_dev:
 * state before the step, lam1_dev: cotangent dL/dm of the state after it,
 * lam0_dev: output dL/dm0; all device arrays of 3 * cells FP64 in SoA order
 * ([x..][y..][z..]).  grad_ms_dev (device, cells FP64) and grad_hext (host,
 * 3 FP64, A/m) are accumulated into and may be NULL.  Needs precision double
 * or mixed (FP64 state); fields use the backend's precision.  Synchronous;
 * the caller orders its own device work before the call. */
int fullmag_fdm_ext_heun_adjoint(fullmag_fdm_backend *handle, double dt_seconds, const double *m0_dev,
                                 const double *lam1_dev, double *lam0_dev, double *grad_ms_dev, double *grad_hext);

/* Device-to-device copy of the FP64 state (SoA, 3 * cells) out of / into the
 * backend, for checkpointing without host round trips.  set_state_device
 * does not normalise; inactive cells must already be zero.  It resets the
 * DP45 FSAL slope and the ABM3 history like upload_magnetization. */
int fullmag_fdm_ext_copy_state_device(fullmag_fdm_backend *handle, double *dst_dev, uint64_t len);
int fullmag_fdm_ext_set_state_device(fullmag_fdm_backend *handle, const double *src_dev, uint64_t len);

/* Batched sweeps: many independent backends (any grids, Heun integrator,
 * one device) advanced together.  batch_create captures steps_per_graph
 * fixed-dt Heun steps of every member, each on its own stream, as parallel
 * branches of one hipGraph; batch_run(n) replays it n / steps_per_graph
 * times, runs the remainder eagerly and fills one step_stats per member
 * (stats of the final state).  Members must not be stepped or destroyed
 * outside the batch while it exists; uploads, state copies and new external
 * fields between runs are honoured (the next run recomputes H_demag of the
 * new m and recaptures the graph for a new field).  A failed create returns a handle whose
 * batch_last_error is set.  No counterpart in the reference ABI; for
 * parameter sweeps that run many small problems on one GPU. */
typedef struct fullmag_fdm_batch fullmag_fdm_batch;
fullmag_fdm_batch *fullmag_fdm_ext_batch_create(fullmag_fdm_backend *const *handles, uint32_t count,
                                                const double *dts, uint32_t steps_per_graph);
int fullmag_fdm_ext_batch_run(fullmag_fdm_batch *batch, uint64_t n_steps, fullmag_fdm_step_stats *stats_out);
const char *fullmag_fdm_ext_batch_last_error(fullmag_fdm_batch *batch);
void fullmag_fdm_ext_batch_destroy(fullmag_fdm_batch *batch);

/* Schedule autotuner: times (demag + stats) evaluations of the current
 * state under candidate schedules (coordinate search over the FULLMAG_FDM_TUNE
 * keys: FFT radices and rows / lines per workgroup, re-planned in place, and
 * the launch-time keys), keeps the fastest and writes it as "key=value,..."
 * into chosen.  Results of every schedule agree with the oracle; only speed
 * differs.  For callers that run many steps of one grid. */
int fullmag_fdm_ext_autotune(fullmag_fdm_backend *handle, int reps, char *chosen, uint64_t chosen_len);

/* Applies a schedule string (as autotune writes it; FULLMAG_FDM_TUNE syntax)
 * to a backend, re-planning its transforms without reallocation, e.g. to
 * reuse one tuning across backends of the same grid.  ERR_INVALID (with
 * last_error) when the schedule does not fit the grid; the backend then keeps
 * its schedule. */
int fullmag_fdm_ext_set_schedule(fullmag_fdm_backend *handle, const char *schedule);

/* Bytes of device memory the backend holds. */
uint64_t fullmag_fdm_ext_device_bytes(fullmag_fdm_backend *handle);

/* Size of the fixed device array (one per GPU, held until the process ends)
 * through which fullmag_fdm_ext_copy_field_preview_noalloc stages its output. */
#define FULLMAG_FDM_EXT_NOALLOC_SCRATCH_BYTES 4194304

/* fullmag_fdm_backend_copy_field_preview_f64 (out_f64 != 0) or _f32 (out_f64
 * == 0) without any device allocation: the box averages are staged in the
 * library's fixed per-GPU array instead of a buffer it allocates, with the
 * same results.  No hipMalloc, hipFree or allocating library routine is called
 * on any path.  Limits: 3 * preview_nx * preview_ny * preview_nz * element
 * size must not exceed FULLMAG_FDM_EXT_NOALLOC_SCRATCH_BYTES (ERR_INVALID with
 * last_error otherwise, also for a full-resolution request); H_ex and H_eff
 * need the observable buffers that the first copy_field / copy_field_preview
 * of either allocates (ERR_INVALID before that).  All calls on one GPU share
 * its array and every call overwrites the part it uses: calls that overlap on
 * one GPU (from several threads or backends) race.  No counterpart in the
 * reference ABI; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or UI loops, harnesses that forbid it). */
int fullmag_fdm_ext_copy_field_preview_noalloc(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                               uint32_t z_origin, uint32_t z_stride, int out_f64, void *out_xyz,
                                               uint64_t out_len);

/* Multilayer stacks: layers on their own grids (FdmMultilayerPlanIR,
 * crates/fullmag-ir/src/lib.rs:395-427) coupled by demag on a common
 * convolution grid, with the self kernel for a layer with itself and the
 * shifted kernel of crates/fullmag-fdm-demag/src/shifted_kernel.rs between
 * layers whose origins differ in z, native grids other than the common one
 * going through the push/pull transfers of transfer.rs, all layers stepped
 * together with Heun (crates/fullmag-runner/src/multilayer_reference.rs).
 * The reference runs this coupling on the host around per-layer backends;
 * here the whole step runs on the device.  No counterpart in the reference
 * ABI; for runners of multilayer plans. */
typedef struct fullmag_fdm_stack fullmag_fdm_stack;

typedef struct {
    uint32_t struct_size;                        /* sizeof(fullmag_fdm_ext_stack_desc) */
    uint32_t layer_count;
    const fullmag_fdm_plan_desc *const *layers;  /* per layer: grid, cell size, material, mask, m0,
                                                    exchange, external field, precision; integrator
                                                    must be Heun; its enable_demag and kernel spectra
                                                    are not used; boundary tiers are refused */
    const double *layer_origins;                 /* 3 per layer: native origin [m] (z sets the shifts) */
    uint32_t common_cells[3];                    /* convolution grid */
    uint32_t reserved;
    double convolution_cell_size[3];             /* [m] */
    int32_t enable_demag;
    int32_t precision_override;                  /* 0: the layers' precision; else 1, 2 or 3 */
} fullmag_fdm_ext_stack_desc;

/* A failed create returns a handle whose stack_last_error is set (NULL only
 * when the handle itself cannot be allocated). */
fullmag_fdm_stack *fullmag_fdm_ext_stack_create(const fullmag_fdm_ext_stack_desc *desc);
const char *fullmag_fdm_ext_stack_last_error(fullmag_fdm_stack *stack);

/* n_steps fixed-dt Heun steps of every layer; one stats record for the
 * stack: energies summed over the layers, maxima over the layers. */
int fullmag_fdm_ext_stack_step(fullmag_fdm_stack *stack, double dt_seconds, uint64_t n_steps,
                               fullmag_fdm_step_stats *out_stats);

/* Recomputes every layer's demag field at its present m (after uploads). */
int fullmag_fdm_ext_stack_refresh(fullmag_fdm_stack *stack);

/* Backend handle of layer `index`, owned by the stack: the copy_field,
 * preview, snapshot, upload, refresh and device-info entry points work on it;
 * step, step_many, batches, the adjoint and schedule changes are refused, and
 * backend_destroy on it does nothing. */
fullmag_fdm_backend *fullmag_fdm_ext_stack_layer(fullmag_fdm_stack *stack, uint32_t index);
void fullmag_fdm_ext_stack_destroy(fullmag_fdm_stack *stack);

#ifdef __cplusplus
}
#endif

#endif /* FULLMAG_FDM_EXT_H */
