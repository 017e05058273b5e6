// context.hpp -- per-backend device state of the HIP FDM library.
//
// Layout in HBM (all vector fields are SoA, component-major, 3 * n elements,
// index x + nx (y + ny z) as crates/fullmag-engine/src/lib.rs:139-141):
//   m        State   magnetisation (FP64 in double/mixed, FP32 in single)
//   mst      State   candidate state of adaptive integrators
//   hd       Real    demag field at the most recently evaluated point
//   mt[2]    Real    stage points of multi-stage integrators
//   k[i]     Real    stage slopes dm/dt
//   spec     cplx<Real>  half spectrum of the padded convolution,
//                    [c][z][kx][ky], kx <= nx, ky < 2ny (3-D) or ny (2-D)
//   kern     Real    Newell tensor spectrum on the non-negative octant
//                    [comp][kz <= nz (3-D) or 0][kx <= nx][ky <= ny]
#pragma once

#include "common.hpp"
#include "fft.hpp"
#include "fullmag_fdm.h"
#include "fullmag_fdm_ext.h"

#include <cstring>
#include <string>
#include <vector>

namespace fm {

enum Mode { MODE_SINGLE = 1, MODE_DOUBLE = 2, MODE_MIXED = 3 };

// Launch-schedule knobs of the demag transforms (0 = automatic).  Settable
// per backend from the FULLMAG_FDM_TUNE environment variable as a comma list
// of key=value (x_radix, x_lanes, x_rows, y_radix, y_lanes, y_lpw, fixed,
// z_quad, xcd, y_pipe, y_pipe_f, z_tile, zc, y_split).  x_radix, x_lanes,
// x_rows, y_radix, y_lanes, y_lpw, z_tile and y_split shape the plan made at
// set-up; the others are read at every launch.
struct DemagTuning {
    int x_radix = 16, x_lanes = 0, x_rows = 0;
    int y_radix = 16, y_lanes = 0, y_lpw = 0;
    int fixed = 1;   // 0: runtime-planned line kernels even for fixed schedules
    int z_quad = -1;  // 1: four lanes per z column instead of two (even nz); -1: FP64 only
    int xcd = 1;     // 0: no per-XCD renumbering of workgroups
    int y_pipe = 1;    // line groups per workgroup of the y line kernel; > 1 prefetches the next (0: off)
    int y_pipe_f = 1;  // > 0: the same for the forward y pass (half-zero input)
    int z_tile = 0;    // > 0: at most this many columns per long-z (LDS line) workgroup
    int zc = 0;        // z planes per workgroup of the cell kernels (0: automatic)
    int y_split = 0;   // 1: thin-film y convolution as separate line and tensor passes
};

// Applies "key=value,..." (unknown keys ignored) on top of t.
void parse_tuning_string(DemagTuning &t, const char *text);

struct DemagPlan {
    DemagTuning tune;
    bool enabled = false;
    bool three_d = false;     // false: nz == 1 thin film, 2-D padded transform
    bool projection = false;  // no tensor: k-space projection fallback
    int nx = 0, ny = 0, nz = 0;
    int px = 0, py = 0, pz = 1;
    int hx = 0;    // nx + 1 columns of the half spectrum
    size_t lds_limit = 65536;  // LDS bytes one workgroup may allocate on this device
    bool y_split = false;      // thin film: y lines, tensor and inverse y lines as three passes
    int rows = 0;  // rows per plane of `spec`
    int kcomp = 0; // 6 (3-D) or 4 (2-D: xz = yz = 0 on the film plane)
    int kz_n = 1;  // octant planes in kz: nz + 1 (3-D) or 1
    double inv_npad = 0.0;
    LinePlan lx, ly, lz;
    // launch geometry
    int xr = 2, x_ls = 0;        // x passes: real rows per workgroup, LDS line stride
    int y_lines_per_wg = 1;      // 3-D y passes
    int yconv_kpw = 1;           // thin-film y convolution: kx columns per workgroup
    int z_tile = 1;              // long z lines: columns per workgroup
    // device buffers (element type depends on Real)
    void *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr;
    void *spec = nullptr;
    void *kern = nullptr;
    uint64_t spec_elems = 0;  // complex elements
    uint64_t kern_elems = 0;  // real elements
    // tensor layout: one block of kblock reals per (kx, oy); the even-oz
    // entries first, the odd-oz ones from kodd (see kern_at)
    int kblock = 0, kodd = 0;
};

// One kernel launch as recorded for a multi-instance batch: the kernel's
// multi-instance twin (null when the launch has none), the launch geometry
// of one instance and its argument pack.  The twin has the signature
// (const Pack *table, int zdiv) and serves instance blockIdx.z / zdiv.
struct LaunchRec {
    const void *multi = nullptr;
    dim3 grid, block;
    size_t lds = 0;
    int zdiv = 1;
    std::vector<unsigned char> pack;
};
struct LaunchRecorder {
    std::vector<LaunchRec> recs;
};

struct TimingEntry {
    std::string name;
    double total_ms = 0.0;
    uint64_t launches = 0;
};

struct PendingEvent {
    int entry;
    hipEvent_t a, b;
};

struct Context {
    // grid and material
    uint32_t nx = 0, ny = 0, nz = 0;
    uint64_t n = 0;
    double dx = 0, dy = 0, dz = 0;
    double ms = 0, a = 0, alpha = 0, gamma = 0;
    int mode = MODE_DOUBLE;
    int integrator = FULLMAG_FDM_INTEGRATOR_HEUN;
    bool precession = true, exchange = true, demag = false, has_ext = false;
    double hext[3] = {0, 0, 0};

    bool has_mask = false;
    std::vector<uint8_t> mask_host;
    uint8_t *d_mask = nullptr;
    bool use_lut = false;        // per-pair A_ij from the region table
    double a_exchange = 0;       // uniform A used when !use_lut
    uint32_t *d_region = nullptr;
    double *d_lut = nullptr;
    bool has_ms_cell = false;
    double *d_ms_cell = nullptr;

    // sub-cell boundary tiers (double and mixed precision; single precision
    // ignores them as the reference's FP32 kernels do)
    int tier = 0;               // plan boundary_correction as an 8-bit value; 0: none
    int tier_ex = 0;            // exchange stencil: 0 standard, 1 T0 face links, 2 T1 distances
    double phi_floor = 0.05, delta_min = 0.0;
    double *d_phi = nullptr;    // volume fraction; non-null: demag packs phi Ms m
    float *d_phi32 = nullptr;   // the same in FP32, in its place under mixed precision
    double *d_tier_w = nullptr; // 6 n face links (T0) or distances (T1): xp, xm, yp, ym, zp, zm
    float *d_tier_w32 = nullptr; // the same in FP32, in their place under mixed precision
    // sparse demag correction, rows grouped by target cell: group g adds rows
    // d_corr_row[d_corr_off[g] .. d_corr_off[g + 1]) (plan order) to cell d_corr_cell[g]
    uint32_t corr_groups = 0, corr_stencil = 0;
    int32_t *d_corr_cell = nullptr;
    uint32_t *d_corr_off = nullptr, *d_corr_row = nullptr;
    int32_t *d_corr_src = nullptr;  // rows x stencil source cells (< 0: unused)
    void *d_corr_n = nullptr;       // rows x stencil x 6: Nxx, Nxy, Nxz, Nyy, Nyz, Nzz; FP64, FP32 when mixed

    double max_err = 1e-5, dt_min = 1e-18, dt_max = 1e-10, headroom = 0.8;

    // time stepping
    uint64_t step_count = 0;
    double time = 0.0;
    bool fsal_valid = false;
    int abm_count = 0, abm_startup = 0;
    double abm_last_dt = 0.0;
    int abm_slot[3] = {4, 5, 6};  // k[] indices of f_n, f_{n-1}, f_{n-2}

    // device
    int device = 0;
    hipStream_t stream = nullptr;
    void *m = nullptr, *mst = nullptr, *hd = nullptr;
    void *mt[2] = {nullptr, nullptr};
    void *k[7] = {};
    void *h_ex_obs = nullptr, *h_eff_obs = nullptr;
    bool hd_at_m = false;     // hd holds H_demag(m)
    // Heun: k[0] = f(m, H(m)) and mt[0] = n(m + pre_dt k[0]) were written by
    // the stats pass of the previous step (its torque is the next step's
    // first slope, evaluated on the same inputs)
    bool pre_valid = false;
    // the spectrum holds the convolved predictor mt[0] left with pre_valid
    // (enqueued after a step's stats while the host returns); any other
    // demag evaluation clears it
    bool spec_pre = false;
    hipEvent_t ev_stats = nullptr;  // recorded after a step's stats copy
    double pre_dt = 0.0;
    bool obs_frozen = false;  // observables kept from before an upload
    bool obs_valid = false;   // h_ex_obs / h_eff_obs hold the fields at m
    DemagPlan dp;

    // reductions
    double *d_partials = nullptr;
    int max_partial_blocks = 0;
    double *d_result = nullptr;
    double *h_result = nullptr;  // pinned: the merged stats of the last stats pass
    double *h_partials = nullptr;  // pinned, coherent: workgroup partials of the stats / error passes
    unsigned stats_blocks = 0;     // workgroups of the last stats pass

    // adjoint work buffers (allocated by the first adjoint call)
    bool adj_ready = false;
    void *adj_hd0 = nullptr, *adj_k1 = nullptr, *adj_mt = nullptr, *adj_hdt = nullptr;  // Real
    void *adj_g = nullptr, *adj_q = nullptr, *adj_w = nullptr;                          // Real
    double *adj_lu = nullptr, *adj_lmt = nullptr;

    // scratch for previews
    void *d_preview = nullptr;
    uint64_t preview_bytes = 0;

    // timing instrumentation
    bool timing = false;
    std::vector<TimingEntry> timings;
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> event_pool;

    // non-null: kernel launches are recorded here instead of being issued
    // (multi-instance batch set-up)
    LaunchRecorder *rec = nullptr;

    // non-null: this backend is a layer of a multilayer stack, whose coupled
    // demag fills `hd` (the layer's own demag plan is unused)
    struct Stack *stack = nullptr;

    std::string last_error;
    fullmag_fdm_device_info info{};
    bool info_valid = false;
    uint64_t device_bytes = 0;
    std::vector<void *> allocations;
    bool ready = false;  // create() finished successfully

    size_t state_bytes() const { return mode == MODE_SINGLE ? 4 : 8; }
    size_t real_bytes() const { return mode == MODE_DOUBLE ? 8 : 4; }
};

// A multilayer stack: layer backends (one Context each, all on the stack's
// stream) coupled by demag on a common convolution grid
// (crates/fullmag-engine/src/multilayer.rs).  Spectra are laid out
// [layer][c][plane][kx][ky] with ky < py; the pair kernels
// [kernel][6][plane][kx][ky] (xx, yy, zz, xy, xz, yz), one real each (the
// transform of a component is purely real or purely imaginary).
struct Stack {
    std::vector<Context *> layers;
    std::vector<double> oz;   // native origin z per layer [m]
    Context conv;             // the common grid: transform plan, twiddles, buffers
    bool demag = true;
    int planes = 1;           // pz (3-D) or 1 (thin layers, 2-D transform)
    uint64_t layer_spec = 0;  // complex elements per layer spectrum
    void *spec_m = nullptr, *spec_h = nullptr;
    void *kern = nullptr;      // per kernel 6 components of one real per k-point
    uint64_t kern_stride = 0;  // reals per pair kernel
    int nkern = 0;
    std::vector<int> pair;    // kernel of (dst, src) at dst * L + src
    int *d_pair = nullptr;
    uint32_t *d_kimag = nullptr;  // per kernel: bit c set when component c is imaginary
    // layers on another grid than the common one (push/pull transfer):
    // conv-grid m (FP64) and H (Real); null where the grids agree
    std::vector<void *> conv_m, conv_h;
    uint64_t step_count = 0;
    double time = 0.0;
    bool ready = false;
    std::string last_error;
};

// Issues a launch through `plain`, or, while c.rec is set, records it with
// its multi-instance twin `multi` (null: none) and argument pack instead.
template <class Pack, class F>
inline void launch_rec(Context &c, const void *multi, dim3 grid, dim3 block, size_t lds, int zdiv, const Pack &pack,
                       F &&plain) {
    if (c.rec) {
        LaunchRec r;
        r.multi = multi;
        r.grid = grid;
        r.block = block;
        r.lds = lds;
        r.zdiv = zdiv;
        r.pack.resize(sizeof(Pack));
        std::memcpy(r.pack.data(), &pack, sizeof(Pack));
        c.rec->recs.push_back(std::move(r));
        return;
    }
    plain();
}

// context.cpp
bool ctx_alloc(Context &c, void **ptr, size_t bytes, const char *what);
bool ctx_upload(Context &c, void *dst, const void *src, size_t bytes, const char *what, std::string &err);
void ctx_free_all(Context &c);
bool ctx_upload_m(Context &c, const void *host, bool host_is_f64);
bool ctx_query_device_info(Context &c);
int timing_entry(Context &c, const char *name);
void timing_begin(Context &c, int entry, hipEvent_t *a);
void timing_end(Context &c, int entry, hipEvent_t a);
void timing_resolve(Context &c);

// demag.hip
bool demag_setup(Context &c, const fullmag_fdm_plan_desc &plan, bool newell_on_device);
// Evaluates H_demag of the field `src` (3n elements, FP64 when src_f64) into
// `out` (Real): -N * (Ms src), or the bare convolution -N * src when
// weight_ms is false.  Inactive cells get zero.  With boundary tiers (and
// weight_ms) the source is phi Ms src in every cell and the sparse
// correction rows are added afterwards, also at inactive targets.
// phase: kDemagFull; kDemagSpectrum stops before the x C2R, leaving the
// convolved spectrum of src; kDemagFinish runs only the x C2R (and the
// boundary correction rows) from such a spectrum.
constexpr int kDemagFull = 0, kDemagSpectrum = 1, kDemagFinish = 2;
void demag_eval(Context &c, const void *src, bool src_f64, void *out, bool weight_ms = true, int phase = kDemagFull);
// Multilayer stack demag (demag.hip).  stack_demag_setup plans the common
// grid's transforms and builds the pair kernels (self kernel for a layer
// with itself, the shifted kernel of newell.rs:528-651 otherwise);
// stack_demag_eval writes every layer's H_demag into its `hd` from the
// sources src[l] (the layers' State type when src_state, else Real).
bool stack_demag_setup(Stack &s, std::string &err);
void stack_demag_eval(Stack &s, const void *const *src, bool src_state);

// Re-plans the demag lines and launch geometry under the schedule `t`
// (radices, lanes, rows per workgroup, tiles) without reallocating; on
// failure the backend keeps its plan and `err` says why.
bool demag_replan(Context &c, const DemagTuning &t, std::string &err);
bool newell_spectra_full(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double dz, bool thin_film_2d,
                         double *out, std::string &err);

// newell.cpp: Newell tensor on the non-negative displacement octant,
// out[c][k][j][i] for i < nx, j < ny, k < nz, c = xx, yy, zz, xy, xz, yz, at
// displacements (i dx, j dy, k dz + z_offset).
void newell_octant(size_t nx, size_t ny, size_t nz, double dx, double dy, double dz, std::vector<double> &out,
                   double z_offset = 0.0);

// llg.hip
int integrator_step(Context &c, double dt, fullmag_fdm_step_stats *st);
bool refresh_observables(Context &c);
bool materialize_observables(Context &c);  // fills h_ex_obs / h_eff_obs at m
// Heun steps of a whole stack (n_steps, stats of the final state summed over
// the layers; energies add, maxima combine) and its demag at the layers' m.
int stack_step(Stack &s, double dt, uint64_t n_steps, fullmag_fdm_step_stats *st);
bool stack_refresh(Stack &s);
bool preview_field(Context &c, const void *field, bool field_f64, uint32_t pnx, uint32_t pny, uint32_t pnz,
                   uint32_t z0, uint32_t zs, void *dev_out, bool out_f64);
bool alloc_integrator_buffers(Context &c);
int integrator_step_many(Context &c, double dt, uint64_t n, fullmag_fdm_step_stats *st);
// Graph-captured batches: prepare (H_demag(m) valid, timing off), enqueue
// n Heun steps on c.stream without host synchronisation (capturable),
// finish (stats at the final state, counters advanced by nsteps).
bool integrator_prepare(Context &c);
void integrator_enqueue_heun(Context &c, double dt, int nsteps);
int integrator_finish(Context &c, double dt, uint64_t nsteps, fullmag_fdm_step_stats *st);
double integrator_probe_ms(Context &c, int reps);

// adjoint.hip: discrete adjoint of one Heun step from the state m0 (device,
// FP64 SoA 3n): lam0 = (dm1/dm0)^T lam1; grad_ms (device, n, may be null)
// and grad_hext (host, 3, may be null) accumulate the gradients of
// <lam1, m1> with respect to per-cell Ms and H_ext.  State must be FP64.
int heun_adjoint(Context &c, double dt, const double *m0, const double *lam1, double *lam0, double *grad_ms,
                 double *grad_hext);

}  // namespace fm
