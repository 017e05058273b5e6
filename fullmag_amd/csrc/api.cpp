// api.cpp -- the C ABI of include/fullmag_fdm.h and include/fullmag_fdm_ext.h.
//
// Entry-point semantics follow native/backends/fdm/src/api.cpp:68-718 (the
// reference implementation of the same header): create() returns a handle
// even when it fails and reports through last_error; length mismatches give
// ERR_INVALID with a message; device failures give ERR_CUDA.
#include "context.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <optional>

using namespace fm;

static_assert(sizeof(fullmag_fdm_grid_desc) == 40, "grid_desc layout");
static_assert(sizeof(fullmag_fdm_material_desc) == 32, "material_desc layout");
static_assert(sizeof(fullmag_fdm_plan_desc) == 448, "plan_desc layout");
static_assert(offsetof(fullmag_fdm_plan_desc, demag_kernel_xx_spectrum) == 120, "plan_desc layout");
static_assert(offsetof(fullmag_fdm_plan_desc, boundary_correction) == 224, "plan_desc layout");
static_assert(offsetof(fullmag_fdm_plan_desc, initial_magnetization_xyz) == 400, "plan_desc layout");
static_assert(sizeof(fullmag_fdm_step_stats) == 96, "step_stats layout");
static_assert(sizeof(fullmag_fdm_device_info) == 144, "device_info layout");
static_assert(sizeof(fullmag_fdm_snapshot_desc) == 24, "snapshot_desc layout");

namespace {

// native/backends/fdm/src/api.cpp:36-62
std::optional<int> env_device() {
    const char *a = std::getenv("FULLMAG_FDM_GPU_INDEX");
    const char *b = std::getenv("FULLMAG_CUDA_DEVICE_INDEX");
    const char *raw = a ? a : b;
    if (!raw || !*raw) return std::nullopt;
    char *end = nullptr;
    long v = std::strtol(raw, &end, 10);
    if (end == raw || *end != '\0' || v < 0) return std::nullopt;
    return (int)v;
}

bool is_gfx950(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return false;
    return std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
}

Context *as_ctx(fullmag_fdm_backend *h) { return reinterpret_cast<Context *>(h); }
fullmag_fdm_backend *as_handle(Context *c) { return reinterpret_cast<fullmag_fdm_backend *>(c); }

bool env_mixed() {
    const char *v = std::getenv("FULLMAG_FDM_MIXED");
    return v && (std::strcmp(v, "1") == 0 || std::strcmp(v, "true") == 0);
}

Context *create_impl(const fullmag_fdm_plan_desc *plan, const fullmag_fdm_ext_options *opt) {
    if (!plan) return nullptr;
    Context *c = new (std::nothrow) Context();
    if (!c) return nullptr;
    try {
        auto dev = env_device();
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
            c->last_error = "no HIP device available";
            return c;
        }
        if (dev) {
            if (*dev >= count) {
                c->last_error = "requested GPU index " + std::to_string(*dev) + " but only " + std::to_string(count) +
                                " device(s) are visible";
                return c;
            }
            if (!hip_ok(hipSetDevice(*dev), "hipSetDevice", c->last_error)) return c;
        }
        hipGetDevice(&c->device);
        if (!is_gfx950(c->device)) {
            c->last_error = "this library is built for gfx950 (MI355X) only";
            return c;
        }

        const fullmag_fdm_plan_desc &p = *plan;
        c->nx = p.grid.nx;
        c->ny = p.grid.ny;
        c->nz = p.grid.nz;
        c->n = (uint64_t)c->nx * c->ny * c->nz;
        c->dx = p.grid.dx;
        c->dy = p.grid.dy;
        c->dz = p.grid.dz;
        c->ms = p.material.saturation_magnetisation;
        c->a = p.material.exchange_stiffness;
        c->alpha = p.material.damping;
        c->gamma = p.material.gyromagnetic_ratio;
        int prec = (int)p.precision;
        if (opt && opt->precision_override) prec = opt->precision_override;
        else if (prec == FULLMAG_FDM_PRECISION_SINGLE && env_mixed()) prec = MODE_MIXED;
        if (prec != MODE_SINGLE && prec != MODE_DOUBLE && prec != MODE_MIXED) {
            c->last_error = "unsupported precision " + std::to_string(prec);
            return c;
        }
        c->mode = prec;
        c->integrator = (int)p.integrator;
        if (c->integrator < FULLMAG_FDM_INTEGRATOR_HEUN || c->integrator > FULLMAG_FDM_INTEGRATOR_RK23)
            c->integrator = FULLMAG_FDM_INTEGRATOR_HEUN;  // reference default branch
        c->precession = p.disable_precession == 0;
        c->exchange = p.enable_exchange != 0;
        c->demag = p.enable_demag != 0;
        c->has_ext = p.has_external_field != 0;
        for (int i = 0; i < 3; ++i) c->hext[i] = c->has_ext ? p.external_field_am[i] : 0.0;
        // api.cpp:131-134 defaults
        c->max_err = p.adaptive_max_error > 0 ? p.adaptive_max_error : 1e-5;
        c->dt_min = p.adaptive_dt_min > 0 ? p.adaptive_dt_min : 1e-18;
        c->dt_max = p.adaptive_dt_max > 0 ? p.adaptive_dt_max : 1e-10;
        c->headroom = p.adaptive_headroom > 0 ? p.adaptive_headroom : 0.8;

        // validation (api.cpp:137-198)
        if (c->n == 0) {
            c->last_error = "grid has zero cells";
            return c;
        }
        const uint64_t n = c->n;
        if (n >= (uint64_t(1) << 32)) {
            // cell coordinates are derived with 32-bit index arithmetic
            c->last_error = "grid has " + std::to_string(n) + " cells; at most 2^32 - 1 are supported";
            return c;
        }
        if (p.initial_magnetization_len != 3 * n || !p.initial_magnetization_xyz) {
            c->last_error = "initial_magnetization_len mismatch: expected " + std::to_string(3 * n) + ", got " +
                            std::to_string(p.initial_magnetization_len);
            return c;
        }
        if (p.active_mask && p.active_mask_len != n) {
            c->last_error = "active_mask_len mismatch: expected " + std::to_string(n) + ", got " +
                            std::to_string(p.active_mask_len);
            return c;
        }
        if (p.region_mask && p.region_mask_len != n) {
            c->last_error = "region_mask_len mismatch: expected " + std::to_string(n) + ", got " +
                            std::to_string(p.region_mask_len);
            return c;
        }
        // sub-cell boundary tiers (reference api.cpp:246-290, context.cu:627-730):
        // the tier is the plan value as an 8-bit number; volume fractions of
        // the wrong length switch the tiers off; face links (tier 1) and
        // distances (tier >= 2) are used only when all six arrays are given.
        // Single precision ignores the tiers, as the reference's FP32 kernels do.
        const uint8_t tier = (uint8_t)p.boundary_correction;
        std::vector<int32_t> corr_cell;
        std::vector<uint32_t> corr_off, corr_row;
        if (tier != 0 && p.volume_fraction && p.volume_fraction_len == n && prec != MODE_SINGLE) {
            c->tier = tier;
            c->phi_floor = p.boundary_phi_floor > 0.0 ? p.boundary_phi_floor : 0.05;
            c->delta_min =
                p.boundary_delta_min > 0.0 ? p.boundary_delta_min : 0.1 * std::min({c->dx, c->dy, c->dz});
            const bool faces = p.face_link_xp && p.face_link_xm && p.face_link_yp && p.face_link_ym &&
                               p.face_link_zp && p.face_link_zm;
            const bool deltas = p.delta_xp && p.delta_xm && p.delta_yp && p.delta_ym && p.delta_zp && p.delta_zm;
            c->tier_ex = (tier == 1 && faces) ? 1 : (tier >= 2 && deltas) ? 2 : 0;
            if (p.has_demag_boundary_corr && p.demag_corr_target_idx && p.demag_corr_target_count > 0 &&
                p.demag_corr_stencil_size > 0 && p.demag_corr_source_idx && p.demag_corr_tensor) {
                const uint64_t rows = p.demag_corr_target_count, st = p.demag_corr_stencil_size;
                // an index past the grid would be an out-of-bounds access
                for (uint64_t t = 0; t < rows; ++t)
                    if (p.demag_corr_target_idx[t] >= 0 && (uint64_t)p.demag_corr_target_idx[t] >= n) {
                        c->last_error = "demag_corr_target_idx[" + std::to_string(t) + "] = " +
                                        std::to_string(p.demag_corr_target_idx[t]) + " is not a cell index";
                        return c;
                    }
                for (uint64_t r = 0; r < rows * st; ++r)
                    if (p.demag_corr_source_idx[r] >= 0 && (uint64_t)p.demag_corr_source_idx[r] >= n) {
                        c->last_error = "demag_corr_source_idx[" + std::to_string(r) + "] = " +
                                        std::to_string(p.demag_corr_source_idx[r]) + " is not a cell index";
                        return c;
                    }
                // rows grouped by target cell, plan order within a cell
                for (uint64_t t = 0; t < rows; ++t)
                    if (p.demag_corr_target_idx[t] >= 0) corr_row.push_back((uint32_t)t);
                std::stable_sort(corr_row.begin(), corr_row.end(), [&](uint32_t a, uint32_t b) {
                    return p.demag_corr_target_idx[a] < p.demag_corr_target_idx[b];
                });
                for (size_t e = 0; e < corr_row.size(); ++e) {
                    const int32_t cell = p.demag_corr_target_idx[corr_row[e]];
                    if (e == 0 || cell != corr_cell.back()) {
                        corr_cell.push_back(cell);
                        corr_off.push_back((uint32_t)e);
                    }
                }
                corr_off.push_back((uint32_t)corr_row.size());
                c->corr_groups = (uint32_t)corr_cell.size();
                c->corr_stencil = (uint32_t)st;
            }
        }
        const uint64_t len3 = 2ull * (2ull * c->nx) * (2ull * c->ny) * (2ull * c->nz);
        const uint64_t len2 = 2ull * (2ull * c->nx) * (2ull * c->ny);
        if (p.demag_kernel_spectrum_len != 0) {
            if (!p.demag_kernel_xx_spectrum || !p.demag_kernel_yy_spectrum || !p.demag_kernel_zz_spectrum ||
                !p.demag_kernel_xy_spectrum || !p.demag_kernel_xz_spectrum || !p.demag_kernel_yz_spectrum) {
                c->last_error =
                    "demag kernel spectra pointers must all be present when demag_kernel_spectrum_len is set";
                return c;
            }
            const bool ok2 = c->nz == 1 && p.demag_kernel_spectrum_len == len2;
            if (!ok2 && p.demag_kernel_spectrum_len != len3) {
                c->last_error = "demag_kernel_spectrum_len mismatch: expected " + std::to_string(len3) + " (3D)" +
                                (c->nz == 1 ? " or " + std::to_string(len2) + " (thin-film 2D)" : std::string()) +
                                ", got " + std::to_string(p.demag_kernel_spectrum_len);
                return c;
            }
        }
        if (p.active_mask) {
            c->has_mask = true;
            c->mask_host.assign(p.active_mask, p.active_mask + n);
        }
        // exchange coupling: uniform unless the region table couples unequal ids
        c->a_exchange = c->a;
        std::vector<double> lut;
        std::vector<uint32_t> regions;
        if (p.region_mask) {
            constexpr uint64_t R = FULLMAG_FDM_MAX_EXCHANGE_REGIONS;
            regions.assign(p.region_mask, p.region_mask + n);
            for (uint32_t r : regions)
                if (r >= R) {
                    c->last_error = "region id " + std::to_string(r) + " exceeds FULLMAG_FDM_MAX_EXCHANGE_REGIONS";
                    return c;
                }
            lut.assign(R * R, 0.0);
            if (p.exchange_lut && p.exchange_lut_len == R * R) std::memcpy(lut.data(), p.exchange_lut, R * R * 8);
            else
                for (uint64_t r = 0; r < R; ++r) lut[r * R + r] = c->a;
            const uint32_t r0 = regions[0];
            bool uniform = true;
            for (uint64_t i = 0; i < n && uniform; ++i)
                if (regions[i] != r0 && !(c->has_mask && c->mask_host[i] == 0)) uniform = false;
            if (uniform) {
                c->a_exchange = lut[(uint64_t)r0 * R + r0];
            } else {
                c->use_lut = true;
            }
        }
        if (opt && opt->ms_cell) {
            if (opt->ms_cell_len != n) {
                c->last_error = "ms_cell_len mismatch: expected " + std::to_string(n);
                return c;
            }
            c->has_ms_cell = true;
        }

        if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate", c->last_error))
            return c;
        if (c->has_mask) {
            if (!ctx_alloc(*c, (void **)&c->d_mask, n, "active mask")) return c;
            if (!ctx_upload(*c, c->d_mask, c->mask_host.data(), n, "hipMemcpy(mask)", c->last_error))
                return c;
        }
        if (c->use_lut) {
            if (!ctx_alloc(*c, (void **)&c->d_region, n * 4, "region mask") ||
                !ctx_alloc(*c, (void **)&c->d_lut, lut.size() * 8, "exchange table"))
                return c;
            if (!ctx_upload(*c, c->d_region, regions.data(), n * 4, "hipMemcpy(regions)", c->last_error) ||
                !ctx_upload(*c, c->d_lut, lut.data(), lut.size() * 8, "hipMemcpy(lut)", c->last_error))
                return c;
        }
        if (c->has_ms_cell) {
            if (!ctx_alloc(*c, (void **)&c->d_ms_cell, n * 8, "per-cell Ms")) return c;
            if (!ctx_upload(*c, c->d_ms_cell, opt->ms_cell, n * 8, "hipMemcpy(Ms)", c->last_error))
                return c;
        }
        if (c->tier) {
            if (c->mode == MODE_MIXED) {
                std::vector<float> phi32(p.volume_fraction, p.volume_fraction + n);
                if (!ctx_alloc(*c, (void **)&c->d_phi32, n * 4, "volume fraction") ||
                    !ctx_upload(*c, c->d_phi32, phi32.data(), n * 4, "hipMemcpy(volume fraction)", c->last_error))
                    return c;
            } else if (!ctx_alloc(*c, (void **)&c->d_phi, n * 8, "volume fraction") ||
                       !ctx_upload(*c, c->d_phi, p.volume_fraction, n * 8, "hipMemcpy(volume fraction)",
                                   c->last_error)) {
                return c;
            }
            if (c->tier_ex) {
                const double *w[6] = {p.face_link_xp, p.face_link_xm, p.face_link_yp,
                                      p.face_link_ym, p.face_link_zp, p.face_link_zm};
                if (c->tier_ex == 2) {
                    const double *dl[6] = {p.delta_xp, p.delta_xm, p.delta_yp, p.delta_ym, p.delta_zp, p.delta_zm};
                    std::copy(dl, dl + 6, w);
                }
                if (c->mode == MODE_MIXED) {
                    // FP32 like the fields they shape (half the bytes per pass)
                    std::vector<float> w32(6 * n);
                    for (int k = 0; k < 6; ++k)
                        for (uint64_t i = 0; i < n; ++i) w32[k * n + i] = (float)w[k][i];
                    if (!ctx_alloc(*c, (void **)&c->d_tier_w32, 6 * n * 4, "boundary face links / distances") ||
                        !ctx_upload(*c, c->d_tier_w32, w32.data(), 6 * n * 4, "hipMemcpy(boundary arrays)",
                                    c->last_error))
                        return c;
                } else {
                    if (!ctx_alloc(*c, (void **)&c->d_tier_w, 6 * n * 8, "boundary face links / distances"))
                        return c;
                    for (int k = 0; k < 6; ++k)
                        if (!ctx_upload(*c, c->d_tier_w + k * n, w[k], n * 8, "hipMemcpy(boundary arrays)",
                                        c->last_error))
                            return c;
                }
            }
            if (c->corr_groups > 0) {
                const uint64_t rows = p.demag_corr_target_count, st = c->corr_stencil;
                if (!ctx_alloc(*c, (void **)&c->d_corr_cell, corr_cell.size() * 4, "demag correction cells") ||
                    !ctx_alloc(*c, (void **)&c->d_corr_off, corr_off.size() * 4, "demag correction offsets") ||
                    !ctx_alloc(*c, (void **)&c->d_corr_row, corr_row.size() * 4, "demag correction rows") ||
                    !ctx_alloc(*c, (void **)&c->d_corr_src, rows * st * 4, "demag correction sources") ||
                    !ctx_alloc(*c, &c->d_corr_n, rows * st * 6 * (c->mode == MODE_MIXED ? 4 : 8),
                               "demag correction tensors"))
                    return c;
                // under mixed precision the rows are stored in FP32, the
                // precision of the demag field they correct
                std::vector<float> ten32;
                if (c->mode == MODE_MIXED)
                    ten32.assign(p.demag_corr_tensor, p.demag_corr_tensor + rows * st * 6);
                if (!ctx_upload(*c, c->d_corr_cell, corr_cell.data(), corr_cell.size() * 4, "hipMemcpy(corr)",
                                c->last_error) ||
                    !ctx_upload(*c, c->d_corr_off, corr_off.data(), corr_off.size() * 4, "hipMemcpy(corr)",
                                c->last_error) ||
                    !ctx_upload(*c, c->d_corr_row, corr_row.data(), corr_row.size() * 4, "hipMemcpy(corr)",
                                c->last_error) ||
                    !ctx_upload(*c, c->d_corr_src, p.demag_corr_source_idx, rows * st * 4, "hipMemcpy(corr)",
                                c->last_error) ||
                    !(ten32.empty()
                          ? ctx_upload(*c, c->d_corr_n, p.demag_corr_tensor, rows * st * 6 * 8, "hipMemcpy(corr)",
                                       c->last_error)
                          : ctx_upload(*c, c->d_corr_n, ten32.data(), rows * st * 6 * 4, "hipMemcpy(corr)",
                                       c->last_error)))
                    return c;
            }
        }
        if (!alloc_integrator_buffers(*c)) return c;
        const bool newell_dev = opt && opt->newell_on_device;
        if (!demag_setup(*c, p, newell_dev)) return c;
        if (c->demag && c->dp.projection)
            std::fprintf(stderr,
                         "[fullmag] WARNING: demag enabled but no Newell tensor kernel loaded -- using spectral "
                         "projection fallback (inaccurate for finite cells)\n");
        if (!ctx_upload_m(*c, p.initial_magnetization_xyz, true)) return c;
        if (!refresh_observables(*c)) return c;
        ctx_query_device_info(*c);
        c->ready = true;
    } catch (const std::bad_alloc &) {
        c->last_error = "host allocation failed";
    } catch (const std::exception &e) {
        c->last_error = e.what();
    }
    return c;
}

// SoA device field of `elem` bytes per scalar -> AoS host array of T.
template <typename T>
bool download_soa(Context &c, const void *dev, size_t elem, T *out) {
    const uint64_t n = c.n;
    if (elem == 8) {
        std::vector<double> h(3 * n);
        if (!hip_ok(hipMemcpyAsync(h.data(), dev, h.size() * 8, hipMemcpyDeviceToHost, c.stream), "hipMemcpy(field)",
                    c.last_error) ||
            !hip_ok(hipStreamSynchronize(c.stream), "hipMemcpy(field)", c.last_error))
            return false;
        for (uint64_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) out[3 * i + k] = (T)h[k * n + i];
    } else {
        std::vector<float> h(3 * n);
        if (!hip_ok(hipMemcpyAsync(h.data(), dev, h.size() * 4, hipMemcpyDeviceToHost, c.stream), "hipMemcpy(field)",
                    c.last_error) ||
            !hip_ok(hipStreamSynchronize(c.stream), "hipMemcpy(field)", c.last_error))
            return false;
        for (uint64_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) out[3 * i + k] = (T)h[k * n + i];
    }
    return true;
}

// Resolves an observable to its device buffer; null buffer means all-zero.
// Returns false for an unknown observable.
bool field_source(Context &c, fullmag_fdm_observable obs, const void **buf, size_t *elem) {
    *buf = nullptr;
    *elem = c.real_bytes();
    switch (obs) {
        case FULLMAG_FDM_OBSERVABLE_M:
            *buf = c.m;
            *elem = c.state_bytes();
            return true;
        case FULLMAG_FDM_OBSERVABLE_H_DEMAG:
            *buf = c.demag ? c.hd : nullptr;
            return true;
        case FULLMAG_FDM_OBSERVABLE_H_EX:
            if (!materialize_observables(c)) return false;
            *buf = c.exchange ? c.h_ex_obs : nullptr;
            return true;
        case FULLMAG_FDM_OBSERVABLE_H_EFF:
            if (!materialize_observables(c)) return false;
            *buf = c.h_eff_obs;
            return true;
        default:
            return false;
    }
}

template <typename T>
int copy_field(fullmag_fdm_backend *h, fullmag_fdm_observable obs, T *out, uint64_t len) {
    if (!h || !out) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(h);
    if (!c.ready) return FULLMAG_FDM_ERR_INVALID;
    if (len != 3 * c.n) {
        c.last_error = "out_len mismatch";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        for (uint64_t i = 0; i < c.n; ++i) {
            const bool act = !c.has_mask || c.mask_host[i] != 0;
            for (int k = 0; k < 3; ++k) out[3 * i + k] = (T)((c.has_ext && act) ? c.hext[k] : 0.0);
        }
        return FULLMAG_FDM_OK;
    }
    const void *buf = nullptr;
    size_t elem = 8;
    if (!field_source(c, obs, &buf, &elem)) {
        if (c.last_error.empty()) c.last_error = "unknown observable";
        return FULLMAG_FDM_ERR_CUDA;
    }
    if (!buf) {
        std::fill(out, out + len, T(0));
        return FULLMAG_FDM_OK;
    }
    return download_soa<T>(c, buf, elem, out) ? FULLMAG_FDM_OK : FULLMAG_FDM_ERR_CUDA;
}

// Box-average preview of the device field `buf` staged through `stage`
// (3 pnx pny pnz elements of T), then copied to the host.
template <typename T>
int preview_staged(Context &c, const void *buf, size_t elem, uint32_t pnx, uint32_t pny, uint32_t pnz, uint32_t z0,
                   uint32_t zs, void *stage, T *out, uint64_t len) {
    if (!preview_field(c, buf, elem == 8, pnx, pny, pnz, z0, zs, stage, sizeof(T) == 8)) return FULLMAG_FDM_ERR_CUDA;
    if (!hip_ok(hipMemcpyAsync(out, stage, len * sizeof(T), hipMemcpyDeviceToHost, c.stream), "hipMemcpy(preview)",
                c.last_error) ||
        !hip_ok(hipStreamSynchronize(c.stream), "hipMemcpy(preview)", c.last_error))
        return FULLMAG_FDM_ERR_CUDA;
    return FULLMAG_FDM_OK;
}

// The external field's box average, computed on the host (context.cu:935-989:
// the uniform field scaled by the active fraction of each bin).
template <typename T>
void preview_h_ext(const Context &c, uint32_t pnx, uint32_t pny, uint32_t pnz, uint32_t z0, uint32_t zs, T *out) {
    for (uint32_t bz = 0; bz < pnz; ++bz) {
        uint32_t za = z0 + bz * zs;
        if (za >= c.nz) za = c.nz - 1;
        uint32_t zb = z0 + (bz + 1) * zs;
        if (zb <= za) zb = za + 1;
        if (zb > c.nz) zb = c.nz;
        for (uint32_t by = 0; by < pny; ++by) {
            uint32_t y0 = (uint32_t)((uint64_t)by * c.ny / pny), y1 = (uint32_t)((uint64_t)(by + 1) * c.ny / pny);
            if (y1 <= y0) y1 = y0 + 1;
            if (y1 > c.ny) y1 = c.ny;
            for (uint32_t bx = 0; bx < pnx; ++bx) {
                uint32_t x0 = (uint32_t)((uint64_t)bx * c.nx / pnx),
                         x1 = (uint32_t)((uint64_t)(bx + 1) * c.nx / pnx);
                if (x1 <= x0) x1 = x0 + 1;
                if (x1 > c.nx) x1 = c.nx;
                double act = 0, cnt = 0;
                for (uint32_t z = za; z < zb; ++z)
                    for (uint32_t y = y0; y < y1; ++y)
                        for (uint32_t x = x0; x < x1; ++x) {
                            const uint64_t i = ((uint64_t)z * c.ny + y) * c.nx + x;
                            act += (!c.has_mask || c.mask_host[i]) ? 1.0 : 0.0;
                            cnt += 1.0;
                        }
                const uint64_t pi = ((uint64_t)bz * pny + by) * pnx + bx;
                const double s = (c.has_ext && cnt > 0) ? act / cnt : 0.0;
                for (int k = 0; k < 3; ++k) out[pi * 3 + k] = (T)(c.hext[k] * s);
            }
        }
    }
}

template <typename T>
int copy_preview(fullmag_fdm_backend *h, fullmag_fdm_observable obs, uint32_t pnx, uint32_t pny, uint32_t pnz,
                 uint32_t z0, uint32_t zs, T *out, uint64_t len) {
    if (!h || !out || pnx == 0 || pny == 0 || pnz == 0 || zs == 0) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(h);
    if (!c.ready) return FULLMAG_FDM_ERR_INVALID;
    const uint64_t count = (uint64_t)pnx * pny * pnz;
    if (len != 3 * count) {
        c.last_error = "preview out_len mismatch";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (z0 >= c.nz) return FULLMAG_FDM_ERR_CUDA;  // reference returns false -> ERR_CUDA
    if (pnx == c.nx && pny == c.ny && pnz == c.nz && z0 == 0 && zs == 1) return copy_field<T>(h, obs, out, len);
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        preview_h_ext<T>(c, pnx, pny, pnz, z0, zs, out);
        return FULLMAG_FDM_OK;
    }
    const void *buf = nullptr;
    size_t elem = 8;
    if (!field_source(c, obs, &buf, &elem)) return FULLMAG_FDM_ERR_CUDA;
    if (!buf) {
        std::fill(out, out + len, T(0));
        return FULLMAG_FDM_OK;
    }
    const size_t bytes = len * sizeof(T);
    if (c.preview_bytes < bytes) {
        if (!ctx_alloc(c, &c.d_preview, bytes, "preview scratch")) return FULLMAG_FDM_ERR_CUDA;
        c.preview_bytes = bytes;
    }
    return preview_staged<T>(c, buf, elem, pnx, pny, pnz, z0, zs, c.d_preview, out, len);
}

struct Snapshot {
    void *staging = nullptr;
    void *host = nullptr;
    uint64_t bytes = 0;
    uint64_t cells = 0;
    uint32_t scalar_bytes = 8;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    bool needs_wait = false;
};

void destroy_snapshot(Snapshot *s) {
    if (!s) return;
    if (s->done) hipEventSynchronize(s->done);
    if (s->done) hipEventDestroy(s->done);
    if (s->ready) hipEventDestroy(s->ready);
    if (s->stream) hipStreamDestroy(s->stream);
    if (s->staging) hipFree(s->staging);
    if (s->host) hipHostFree(s->host);
    delete s;
}

int upload_impl(fullmag_fdm_backend *h, const void *m, uint64_t len, bool f64) {
    if (!h || !m) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(h);
    if (!c.ready) return FULLMAG_FDM_ERR_INVALID;
    if (len != 3 * c.n) {
        c.last_error = "magnetization length mismatch";
        return FULLMAG_FDM_ERR_INVALID;
    }
    // keep the field observables of the previous state until refresh/step
    if (!c.obs_frozen && !materialize_observables(c)) return FULLMAG_FDM_ERR_CUDA;
    c.obs_frozen = true;
    if (!ctx_upload_m(c, m, f64)) return FULLMAG_FDM_ERR_CUDA;
    c.hd_at_m = false;
    // an externally replaced state invalidates integrator history
    c.fsal_valid = false;
    c.pre_valid = false;
    c.abm_count = 0;
    c.abm_startup = 0;
    c.abm_last_dt = 0.0;
    return FULLMAG_FDM_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free preview: staging in one fixed device array per GPU.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocScratchBytes = FULLMAG_FDM_EXT_NOALLOC_SCRATCH_BYTES;
static_assert(kNoallocScratchBytes % sizeof(double) == 0, "whole doubles");

// Each GPU that loads the library holds one copy until the process ends.  In
// host code this name is only a handle for hipGetSymbolAddress: the device
// address of a GPU's copy is `noalloc_scratch(device)`.
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on `device`, resolved once per device with that
// device current (the caller's current device is restored).
void *noalloc_scratch(int device, std::string &err) {
    if (device < 0 || device >= kMaxScratchDevices) {
        err = "no-alloc preview: device index out of range";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = -1;
        if (!hip_ok(hipGetDevice(&prev), "hipGetDevice", err)) return nullptr;
        if (prev != device && !hip_ok(hipSetDevice(device), "hipSetDevice", err)) return nullptr;
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (prev != device) hipSetDevice(prev);
        if (!hip_ok(e, "hipGetSymbolAddress(scratch)", err)) return nullptr;
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

// copy_preview without any device allocation: the preview kernel writes into
// the GPU's `scratch` (every element of the output, so nothing is cleared),
// also for a full-resolution request.  H_ex / H_eff need the observable
// buffers that the first copy of either allocates.
template <typename T>
int copy_preview_noalloc(fullmag_fdm_backend *h, fullmag_fdm_observable obs, uint32_t pnx, uint32_t pny,
                         uint32_t pnz, uint32_t z0, uint32_t zs, T *out, uint64_t len) {
    if (!h || !out || pnx == 0 || pny == 0 || pnz == 0 || zs == 0) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(h);
    if (!c.ready) return FULLMAG_FDM_ERR_INVALID;
    const uint64_t count = (uint64_t)pnx * pny * pnz;
    if (len != 3 * count) {
        c.last_error = "preview out_len mismatch";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (len > kNoallocScratchBytes / sizeof(T)) {
        c.last_error = "no-alloc preview: 3 * preview cells * element size exceeds the " +
                       std::to_string(kNoallocScratchBytes) + "-byte scratch array";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (z0 >= c.nz) return FULLMAG_FDM_ERR_CUDA;  // as copy_preview
    if (obs == FULLMAG_FDM_OBSERVABLE_H_EXT) {
        preview_h_ext<T>(c, pnx, pny, pnz, z0, zs, out);
        return FULLMAG_FDM_OK;
    }
    if ((obs == FULLMAG_FDM_OBSERVABLE_H_EX || obs == FULLMAG_FDM_OBSERVABLE_H_EFF) &&
        (!c.h_ex_obs || !c.h_eff_obs)) {
        c.last_error = "no-alloc preview: the H_ex / H_eff observable buffers are not allocated yet (one "
                       "fullmag_fdm_backend_copy_field or copy_field_preview of H_ex or H_eff allocates them)";
        return FULLMAG_FDM_ERR_INVALID;
    }
    const void *buf = nullptr;
    size_t elem = 8;
    if (!field_source(c, obs, &buf, &elem)) {
        if (c.last_error.empty()) c.last_error = "unknown observable";
        return FULLMAG_FDM_ERR_CUDA;
    }
    if (!buf) {
        std::fill(out, out + len, T(0));
        return FULLMAG_FDM_OK;
    }
    void *stage = noalloc_scratch(c.device, c.last_error);
    if (!stage) return FULLMAG_FDM_ERR_CUDA;
    return preview_staged<T>(c, buf, elem, pnx, pny, pnz, z0, zs, stage, out, len);
}

// Layers of a stack are stepped, and owned, by their stack.
bool refuse_stack_layer(Context &c) {
    if (!c.stack) return false;
    c.last_error = "this backend is a layer of a multilayer stack: step the stack";
    return true;
}

void stack_free(Stack *s) {
    hipStream_t stream = s->conv.stream;
    if (stream) hipStreamSynchronize(stream);
    for (Context *c : s->layers) {
        if (c->stack) c->stream = nullptr;  // the stack's stream
        ctx_free_all(*c);
        delete c;
    }
    s->conv.stream = nullptr;
    ctx_free_all(s->conv);
    if (stream) hipStreamDestroy(stream);
    delete s;
}

Stack *stack_create_impl(const fullmag_fdm_ext_stack_desc &d, Stack *s) {
    if (d.layer_count == 0 || !d.layers || !d.layer_origins) {
        s->last_error = "a stack needs at least one layer, its plans and its origins";
        return s;
    }
    for (int k = 0; k < 3; ++k)
        if (d.common_cells[k] == 0 || !(d.convolution_cell_size[k] > 0.0)) {
            s->last_error = "common_cells and convolution_cell_size must be positive";
            return s;
        }
    const uint64_t ncv = (uint64_t)d.common_cells[0] * d.common_cells[1] * d.common_cells[2];
    if (ncv >= (uint64_t(1) << 32)) {
        s->last_error = "the common grid has too many cells";
        return s;
    }
    for (uint32_t l = 0; l < d.layer_count; ++l) {
        if (!d.layers[l]) {
            s->last_error = "layer " + std::to_string(l) + ": null plan";
            return s;
        }
        fullmag_fdm_plan_desc p = *d.layers[l];
        if (p.integrator != FULLMAG_FDM_INTEGRATOR_HEUN) {
            // the reference's multilayer planner and runner step Heun only
            s->last_error = "layer " + std::to_string(l) + ": multilayer stacks step with the Heun integrator";
            return s;
        }
        if (p.boundary_correction != FULLMAG_FDM_BOUNDARY_NONE && p.volume_fraction) {
            s->last_error = "layer " + std::to_string(l) + ": boundary tiers are not supported in a stack";
            return s;
        }
        p.enable_demag = 0;  // the stack owns the demag coupling
        p.demag_kernel_spectrum_len = 0;
        fullmag_fdm_ext_options opt{};
        opt.struct_size = sizeof(opt);
        opt.precision_override = d.precision_override;
        Context *c = create_impl(&p, &opt);
        if (!c) {
            s->last_error = "layer " + std::to_string(l) + ": allocation failed";
            return s;
        }
        if (!c->ready) {
            s->last_error = "layer " + std::to_string(l) + ": " + c->last_error;
            ctx_free_all(*c);
            delete c;
            return s;
        }
        s->layers.push_back(c);
        s->oz.push_back(d.layer_origins[3 * l + 2]);
        if (c->mode != s->layers[0]->mode) {
            s->last_error = "all layers of a stack must use one precision";
            return s;
        }
    }
    Context &cv = s->conv;
    cv.nx = d.common_cells[0];
    cv.ny = d.common_cells[1];
    cv.nz = d.common_cells[2];
    cv.n = ncv;
    cv.dx = d.convolution_cell_size[0];
    cv.dy = d.convolution_cell_size[1];
    cv.dz = d.convolution_cell_size[2];
    cv.mode = s->layers[0]->mode;
    cv.device = s->layers[0]->device;
    s->demag = d.enable_demag != 0;
    if (!hip_ok(hipStreamCreateWithFlags(&cv.stream, hipStreamNonBlocking), "hipStreamCreate", s->last_error))
        return s;
    for (Context *c : s->layers) {
        hipStreamSynchronize(c->stream);
        hipStreamDestroy(c->stream);
        c->stream = cv.stream;
        c->stack = s;
        if (s->demag) {
            c->demag = true;
            c->hd_at_m = false;
            if (!ctx_alloc(*c, &c->hd, 3 * c->n * c->real_bytes(), "H_demag")) {
                s->last_error = c->last_error;
                return s;
            }
        }
    }
    if (!stack_demag_setup(*s, s->last_error)) return s;
    if (!stack_refresh(*s)) return s;
    s->ready = true;
    return s;
}

}  // namespace

namespace fm {
Context *backend_context(fullmag_fdm_backend *h) { return as_ctx(h); }
int integrator_step_many(Context &c, double dt, uint64_t n, fullmag_fdm_step_stats *st);
}

extern "C" {

int fullmag_fdm_is_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    auto dev = env_device();
    if (dev && *dev >= count) return 0;
    return is_gfx950(dev ? *dev : 0) ? 1 : 0;
}

fullmag_fdm_backend *fullmag_fdm_backend_create(const fullmag_fdm_plan_desc *plan) {
    return as_handle(create_impl(plan, nullptr));
}

fullmag_fdm_backend *fullmag_fdm_ext_backend_create(const fullmag_fdm_plan_desc *plan,
                                                    const fullmag_fdm_ext_options *options) {
    if (options && options->struct_size != sizeof(fullmag_fdm_ext_options)) {
        Context *c = new (std::nothrow) Context();
        if (c) c->last_error = "fullmag_fdm_ext_options.struct_size mismatch";
        return as_handle(c);
    }
    return as_handle(create_impl(plan, options));
}

int fullmag_fdm_backend_step(fullmag_fdm_backend *handle, double dt_seconds, fullmag_fdm_step_stats *out_stats) {
    if (!handle || !out_stats) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.ready) {
        if (c.last_error.empty()) c.last_error = "backend is not initialised";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (!(dt_seconds > 0.0) || !std::isfinite(dt_seconds)) {
        c.last_error = "dt must be positive and finite";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (refuse_stack_layer(c)) return FULLMAG_FDM_ERR_INVALID;
    return integrator_step(c, dt_seconds, out_stats);
}

int fullmag_fdm_ext_step_many(fullmag_fdm_backend *handle, double dt_seconds, uint64_t n_steps,
                              fullmag_fdm_step_stats *out_stats) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.ready || !(dt_seconds > 0.0) || !std::isfinite(dt_seconds) || n_steps == 0) {
        c.last_error = "step_many needs an initialised backend, dt > 0 and n_steps > 0";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (refuse_stack_layer(c)) return FULLMAG_FDM_ERR_INVALID;
    return integrator_step_many(c, dt_seconds, n_steps, out_stats);
}

int fullmag_fdm_backend_copy_field_f64(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                       double *out_xyz, uint64_t out_len) {
    return copy_field<double>(handle, observable, out_xyz, out_len);
}

int fullmag_fdm_backend_copy_field_f32(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                       float *out_xyz, uint64_t out_len) {
    return copy_field<float>(handle, observable, out_xyz, out_len);
}

int fullmag_fdm_backend_copy_field_preview_f64(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                               uint32_t z_origin, uint32_t z_stride, double *out_xyz,
                                               uint64_t out_len) {
    return copy_preview<double>(handle, observable, preview_nx, preview_ny, preview_nz, z_origin, z_stride, out_xyz,
                                out_len);
}

int fullmag_fdm_backend_copy_field_preview_f32(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                               uint32_t z_origin, uint32_t z_stride, float *out_xyz,
                                               uint64_t out_len) {
    return copy_preview<float>(handle, observable, preview_nx, preview_ny, preview_nz, z_origin, z_stride, out_xyz,
                               out_len);
}

int fullmag_fdm_ext_copy_field_preview_noalloc(fullmag_fdm_backend *handle, fullmag_fdm_observable observable,
                                               uint32_t preview_nx, uint32_t preview_ny, uint32_t preview_nz,
                                               uint32_t z_origin, uint32_t z_stride, int out_f64, void *out_xyz,
                                               uint64_t out_len) {
    if (out_f64)
        return copy_preview_noalloc<double>(handle, observable, preview_nx, preview_ny, preview_nz, z_origin,
                                            z_stride, static_cast<double *>(out_xyz), out_len);
    return copy_preview_noalloc<float>(handle, observable, preview_nx, preview_ny, preview_nz, z_origin, z_stride,
                                       static_cast<float *>(out_xyz), out_len);
}

fullmag_fdm_field_snapshot *fullmag_fdm_backend_begin_field_snapshot(fullmag_fdm_backend *handle,
                                                                     fullmag_fdm_observable observable) {
    if (!handle) return nullptr;
    Context &c = *as_ctx(handle);
    if (!c.ready) return nullptr;
    Snapshot *s = new (std::nothrow) Snapshot();
    if (!s) {
        c.last_error = "failed to allocate async field snapshot";
        return nullptr;
    }
    s->cells = c.n;
    auto fail = [&](const std::string &msg) -> fullmag_fdm_field_snapshot * {
        c.last_error = msg;
        destroy_snapshot(s);
        return nullptr;
    };
    const void *buf = nullptr;
    size_t elem = c.real_bytes();
    const bool ext = observable == FULLMAG_FDM_OBSERVABLE_H_EXT;
    if (!ext) {
        if (!field_source(c, observable, &buf, &elem)) return fail("unsupported async snapshot observable");
    } else {
        elem = c.real_bytes();
    }
    s->scalar_bytes = (uint32_t)elem;
    s->bytes = 3 * c.n * elem;
    if (hipHostMalloc(&s->host, s->bytes, hipHostMallocDefault) != hipSuccess)
        return fail("hipHostMalloc(snapshot host buffer) failed");
    if (ext || !buf) {
        for (uint64_t i = 0; i < c.n; ++i) {
            const bool act = !c.has_mask || c.mask_host[i] != 0;
            for (int k = 0; k < 3; ++k) {
                const double v = (ext && c.has_ext && act) ? c.hext[k] : 0.0;
                if (elem == 8) static_cast<double *>(s->host)[k * c.n + i] = v;
                else static_cast<float *>(s->host)[k * c.n + i] = (float)v;
            }
        }
        return reinterpret_cast<fullmag_fdm_field_snapshot *>(s);
    }
    std::string err;
    if (!hip_ok(hipMalloc(&s->staging, s->bytes), "hipMalloc(snapshot staging)", err) ||
        !hip_ok(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate(snapshot)", err) ||
        !hip_ok(hipEventCreateWithFlags(&s->ready, hipEventDisableTiming), "hipEventCreate(snapshot)", err) ||
        !hip_ok(hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "hipEventCreate(snapshot)", err) ||
        !hip_ok(hipMemcpyAsync(s->staging, buf, s->bytes, hipMemcpyDeviceToDevice, c.stream), "snapshot staging",
                err) ||
        !hip_ok(hipEventRecord(s->ready, c.stream), "hipEventRecord(snapshot)", err) ||
        !hip_ok(hipStreamWaitEvent(s->stream, s->ready, 0), "hipStreamWaitEvent(snapshot)", err) ||
        !hip_ok(hipMemcpyAsync(s->host, s->staging, s->bytes, hipMemcpyDeviceToHost, s->stream), "snapshot copy",
                err) ||
        !hip_ok(hipEventRecord(s->done, s->stream), "hipEventRecord(snapshot)", err))
        return fail(err);
    s->needs_wait = true;
    return reinterpret_cast<fullmag_fdm_field_snapshot *>(s);
}

int fullmag_fdm_field_snapshot_wait(fullmag_fdm_field_snapshot *snapshot, const void **out_data,
                                    uint64_t *out_len_bytes, fullmag_fdm_snapshot_desc *out_desc) {
    if (!snapshot || !out_data || !out_len_bytes || !out_desc) return FULLMAG_FDM_ERR_INVALID;
    Snapshot *s = reinterpret_cast<Snapshot *>(snapshot);
    if (s->needs_wait) {
        if (hipEventSynchronize(s->done) != hipSuccess) return FULLMAG_FDM_ERR_CUDA;
        s->needs_wait = false;
    }
    *out_data = s->host;
    *out_len_bytes = s->bytes;
    out_desc->cell_count = s->cells;
    out_desc->component_count = 3;
    out_desc->scalar_bytes = s->scalar_bytes;
    out_desc->scalar_type = s->scalar_bytes == 4 ? FULLMAG_FDM_SNAPSHOT_SCALAR_F32 : FULLMAG_FDM_SNAPSHOT_SCALAR_F64;
    return FULLMAG_FDM_OK;
}

void fullmag_fdm_field_snapshot_destroy(fullmag_fdm_field_snapshot *snapshot) {
    destroy_snapshot(reinterpret_cast<Snapshot *>(snapshot));
}

int fullmag_fdm_backend_upload_magnetization_f64(fullmag_fdm_backend *handle, const double *m_xyz, uint64_t len) {
    return upload_impl(handle, m_xyz, len, true);
}

int fullmag_fdm_backend_upload_magnetization_f32(fullmag_fdm_backend *handle, const float *m_xyz, uint64_t len) {
    return upload_impl(handle, m_xyz, len, false);
}

int fullmag_fdm_backend_refresh_observables(fullmag_fdm_backend *handle) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.ready) return FULLMAG_FDM_ERR_INVALID;
    return refresh_observables(c) ? FULLMAG_FDM_OK : FULLMAG_FDM_ERR_CUDA;
}

int fullmag_fdm_backend_get_device_info(fullmag_fdm_backend *handle, fullmag_fdm_device_info *out_info) {
    if (!handle || !out_info) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.info_valid && !ctx_query_device_info(c)) return FULLMAG_FDM_ERR_CUDA;
    *out_info = c.info;
    return FULLMAG_FDM_OK;
}

const char *fullmag_fdm_backend_last_error(fullmag_fdm_backend *handle) {
    if (!handle) return "null handle";
    Context &c = *as_ctx(handle);
    return c.last_error.empty() ? nullptr : c.last_error.c_str();
}

void fullmag_fdm_backend_destroy(fullmag_fdm_backend *handle) {
    if (!handle) return;
    Context *c = as_ctx(handle);
    if (c->stack) return;  // destroyed with its stack
    ctx_free_all(*c);
    delete c;
}

int fullmag_fdm_ext_precision(fullmag_fdm_backend *handle) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    return as_ctx(handle)->mode;
}

int fullmag_fdm_ext_newell_spectra(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double dz,
                                   int thin_film_2d, double *out, uint64_t out_len) {
    if (!out || nx == 0 || ny == 0 || nz == 0) return FULLMAG_FDM_ERR_INVALID;
    if (thin_film_2d && nz != 1) return FULLMAG_FDM_ERR_INVALID;
    const uint64_t npad = thin_film_2d ? 4ull * nx * ny : 8ull * nx * ny * nz;
    if (out_len != 6 * 2 * npad) return FULLMAG_FDM_ERR_INVALID;
    std::string err;
    if (!newell_spectra_full(nx, ny, nz, dx, dy, dz, thin_film_2d != 0, out, err)) {
        std::fprintf(stderr, "[fullmag] newell spectra: %s\n", err.c_str());
        return FULLMAG_FDM_ERR_CUDA;
    }
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_ext_set_external_field(fullmag_fdm_backend *handle, const double h_am[3]) {
    if (!handle || !h_am) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    c.has_ext = true;
    for (int k = 0; k < 3; ++k) c.hext[k] = h_am[k];
    c.obs_valid = false;
    // stored slopes were evaluated with the old field
    c.fsal_valid = false;
    c.pre_valid = false;
    c.abm_count = 0;
    c.abm_startup = 0;
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_ext_heun_adjoint(fullmag_fdm_backend *handle, double dt_seconds, const double *m0_dev,
                                 const double *lam1_dev, double *lam0_dev, double *grad_ms_dev, double *grad_hext) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.ready || !(dt_seconds > 0.0) || !std::isfinite(dt_seconds) || !m0_dev || !lam1_dev || !lam0_dev) {
        c.last_error = "heun_adjoint needs an initialised backend, dt > 0 and device arrays m0, lam1, lam0";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (refuse_stack_layer(c)) return FULLMAG_FDM_ERR_INVALID;
    return heun_adjoint(c, dt_seconds, m0_dev, lam1_dev, lam0_dev, grad_ms_dev, grad_hext);
}

int fullmag_fdm_ext_copy_state_device(fullmag_fdm_backend *handle, double *dst_dev, uint64_t len) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.ready || !dst_dev || len != 3 * c.n || c.state_bytes() != 8) {
        c.last_error = "copy_state_device needs an FP64 state (double or mixed) and len = 3 * cells";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (!hip_ok(hipMemcpyAsync(dst_dev, c.m, len * sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                "copy_state_device", c.last_error) ||
        !hip_ok(hipStreamSynchronize(c.stream), "copy_state_device", c.last_error))
        return FULLMAG_FDM_ERR_CUDA;
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_ext_set_state_device(fullmag_fdm_backend *handle, const double *src_dev, uint64_t len) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (!c.ready || !src_dev || len != 3 * c.n || c.state_bytes() != 8) {
        c.last_error = "set_state_device needs an FP64 state (double or mixed) and len = 3 * cells";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (!hip_ok(hipMemcpyAsync(c.m, src_dev, len * sizeof(double), hipMemcpyDeviceToDevice, c.stream),
                "set_state_device", c.last_error) ||
        !hip_ok(hipStreamSynchronize(c.stream), "set_state_device", c.last_error))
        return FULLMAG_FDM_ERR_CUDA;
    c.hd_at_m = false;
    c.obs_valid = false;
    c.obs_frozen = false;
    c.fsal_valid = false;
    c.pre_valid = false;
    c.abm_count = 0;
    c.abm_startup = 0;
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_ext_set_timing(fullmag_fdm_backend *handle, int enable) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    c.timing = enable != 0;
    if (c.timing) c.timings.clear();
    return FULLMAG_FDM_OK;
}

int fullmag_fdm_ext_get_timing(fullmag_fdm_backend *handle, int index, const char **name, double *total_ms,
                               uint64_t *launches) {
    if (!handle) return FULLMAG_FDM_ERR_INVALID;
    Context &c = *as_ctx(handle);
    if (index < 0 || index >= (int)c.timings.size()) return FULLMAG_FDM_ERR_INVALID;
    if (name) *name = c.timings[index].name.c_str();
    if (total_ms) *total_ms = c.timings[index].total_ms;
    if (launches) *launches = c.timings[index].launches;
    return FULLMAG_FDM_OK;
}

void fullmag_fdm_ext_random_unit_vectors(uint64_t seed, uint64_t count, double *out_xyz) {
    if (!out_xyz) return;
    uint64_t s = seed;
    for (uint64_t i = 0; i < count; ++i) {
        double v[3];
        for (;;) {
            for (int k = 0; k < 3; ++k) {
                s ^= s << 13;
                s ^= s >> 7;
                s ^= s << 17;
                v[k] = ((double)s / (double)UINT64_MAX) * 2.0 - 1.0;
            }
            const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (r > 1e-10) {
                for (int k = 0; k < 3; ++k) out_xyz[3 * i + k] = v[k] / r;
                break;
            }
        }
    }
}

uint64_t fullmag_fdm_ext_device_bytes(fullmag_fdm_backend *handle) {
    return handle ? as_ctx(handle)->device_bytes : 0;
}

fullmag_fdm_stack *fullmag_fdm_ext_stack_create(const fullmag_fdm_ext_stack_desc *desc) {
    Stack *s = new (std::nothrow) Stack();
    if (!s) return nullptr;
    try {
        if (!desc || desc->struct_size != sizeof(fullmag_fdm_ext_stack_desc)) s->last_error = "fullmag_fdm_ext_stack_desc missing or of another size";
        else stack_create_impl(*desc, s);
    } catch (const std::bad_alloc &) {
        s->last_error = "host allocation failed";
    } catch (const std::exception &e) {
        s->last_error = e.what();
    }
    return reinterpret_cast<fullmag_fdm_stack *>(s);
}

const char *fullmag_fdm_ext_stack_last_error(fullmag_fdm_stack *stack) {
    if (!stack) return "null handle";
    Stack *s = reinterpret_cast<Stack *>(stack);
    return s->last_error.empty() ? nullptr : s->last_error.c_str();
}

int fullmag_fdm_ext_stack_step(fullmag_fdm_stack *stack, double dt_seconds, uint64_t n_steps,
                               fullmag_fdm_step_stats *out_stats) {
    if (!stack) return FULLMAG_FDM_ERR_INVALID;
    Stack &s = *reinterpret_cast<Stack *>(stack);
    if (!s.ready || !(dt_seconds > 0.0) || !std::isfinite(dt_seconds) || n_steps == 0) {
        s.last_error = "stack_step needs an initialised stack, dt > 0 and n_steps > 0";
        return FULLMAG_FDM_ERR_INVALID;
    }
    return stack_step(s, dt_seconds, n_steps, out_stats);
}

int fullmag_fdm_ext_stack_refresh(fullmag_fdm_stack *stack) {
    if (!stack) return FULLMAG_FDM_ERR_INVALID;
    Stack &s = *reinterpret_cast<Stack *>(stack);
    if (!s.ready) return FULLMAG_FDM_ERR_INVALID;
    return stack_refresh(s) ? FULLMAG_FDM_OK : FULLMAG_FDM_ERR_CUDA;
}

fullmag_fdm_backend *fullmag_fdm_ext_stack_layer(fullmag_fdm_stack *stack, uint32_t index) {
    if (!stack) return nullptr;
    Stack &s = *reinterpret_cast<Stack *>(stack);
    if (!s.ready || index >= s.layers.size()) return nullptr;
    return as_handle(s.layers[index]);
}

void fullmag_fdm_ext_stack_destroy(fullmag_fdm_stack *stack) {
    if (stack) stack_free(reinterpret_cast<Stack *>(stack));
}

}  // extern "C"
