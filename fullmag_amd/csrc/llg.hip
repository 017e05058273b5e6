// llg.hip -- fused field/torque/stage kernels and the time integrators.
//
// Replaces native/backends/fdm/src/{exchange,llg,llg_dp45,llg_rk4,llg_rk23,
// llg_abm3,reductions}_fp64.cu and their fp32 twins.  One "stage" kernel does
// per cell what the reference spreads over exchange, combine, rhs, predictor
// and reduction kernels:
//   H_eff = H_ex(src) + H_demag + H_ext          (stencil read from `src`)
//   k     = -gamma/(1+alpha^2) (p m x H + alpha m x (m x H))
//   out   = normalise(m0 + dt (sum_j a_j k_j + a_new k))
// optionally accumulating the step diagnostics (compensated energy sums and
// max norms) or the adaptive error norm.  Semantics follow the CPU reference
// crates/fullmag-engine/src/lib.rs (exchange :2247-2299, rhs :2584-2595,
// energies :2653-2703, Heun :1217-1249, RK4 :1254-1302, RK23 :1332-1407,
// DP45 :1412-1568, ABM3 :1573-1648, error norm :1653-1674).
#include "field.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <utility>

namespace fm {
namespace {

// out = normalise(m0 + dt sum_j a_j k_j), no field evaluation (DP45 FSAL
// first stage, ABM3 predictor).
template <typename State, typename Real, typename Out>
__global__ void __launch_bounds__(kBlock) combine_kernel(const State *m0, StageArgs<State, Real, State, Out> a) {
    const int64_t n = a.f.n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        State v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            State s = State(0);
            for (int j = 0; j < a.nprev; ++j) s += State(a.ap[j]) * State(a.kp[j][c * n + i]);
            v[c] = m0[c * n + i] + State(a.dt) * s;
        }
        const State nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const State inv = nrm > State(0) ? State(1) / nrm : State(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) a.out[c * n + i] = Out(v[c] * inv);
    }
}

// Field observables at m: H_ex and H_eff (Real); TIER: sub-cell boundary
// exchange.
template <typename State, typename Real, bool TIER>
__global__ void __launch_bounds__(kBlock) observe_kernel(FieldArgs f, const State *m, const Real *hd, Real *hex_out,
                                                         Real *heff_out) {
    const int64_t n = f.n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool act = !f.mask || f.mask[i];
        Real hex[3] = {Real(0), Real(0), Real(0)}, heff[3] = {Real(0), Real(0), Real(0)};
        if (act) {
            if (f.exchange) {
                State he[3];
                if (TIER) exchange_tier_at<State, State, Real>(f, m, i, he);
                else exchange_at<State, State>(f, m, i, he);
                for (int c = 0; c < 3; ++c) hex[c] = Real(he[c]);
            }
            const Real ext[3] = {f.has_ext ? Real(f.hx) : Real(0), f.has_ext ? Real(f.hy) : Real(0),
                                 f.has_ext ? Real(f.hz) : Real(0)};
            for (int c = 0; c < 3; ++c) heff[c] = (hex[c] + (hd ? hd[c * n + i] : Real(0))) + ext[c];
        }
        for (int c = 0; c < 3; ++c) {
            hex_out[c * n + i] = hex[c];
            heff_out[c * n + i] = heff[c];
        }
    }
}

// Box-average preview.  Bin (bx, by, bz) averages the cells
// [bx nx / pnx, (bx + 1) nx / pnx) x [by ny / pny, ...) (at least one cell
// per axis) over the z planes [z0 + bz zs, z0 + (bz + 1) zs) clamped to the
// grid, the bin edges and clamps of native/backends/fdm/src/context.cu:
// 222-286, summed in FP64.  One wavefront per bin: the lanes walk the bin's
// cells x-fastest with stride 64, so each row of the box is read by
// consecutive lanes, and a wave reduction combines the lane sums.
template <typename In, typename OutT>
__global__ void __launch_bounds__(256) preview_kernel(const In *fx, const In *fy, const In *fz, uint32_t nx,
                                                      uint32_t ny, uint32_t nz, uint32_t pnx, uint32_t pny,
                                                      uint32_t pnz, uint32_t z0, uint32_t zs, OutT *out) {
    const uint64_t bins = (uint64_t)pnx * pny * pnz;
    const uint64_t p = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // one bin per wavefront
    const int lane = (int)(threadIdx.x & 63);
    if (p >= bins) return;
    const uint32_t bx = (uint32_t)(p % pnx), by = (uint32_t)((p / pnx) % pny), bz = (uint32_t)(p / ((uint64_t)pnx * pny));
    auto edges = [](uint32_t b, uint32_t n, uint32_t pn, uint32_t &lo, uint32_t &hi) {
        lo = (uint32_t)((uint64_t)b * n / pn);
        hi = (uint32_t)((uint64_t)(b + 1) * n / pn);
        hi = hi > lo ? hi : lo + 1;
        hi = hi < n ? hi : n;
    };
    uint32_t xa, xb, ya, yb;
    edges(bx, nx, pnx, xa, xb);
    edges(by, ny, pny, ya, yb);
    const uint32_t za = z0 + bz * zs < nz ? z0 + bz * zs : nz - 1;
    uint32_t zb = z0 + (bz + 1) * zs;
    zb = zb > za ? zb : za + 1;
    zb = zb < nz ? zb : nz;
    const uint32_t wx = xb - xa, wy = yb - ya;
    const uint64_t cells = (uint64_t)wx * wy * (zb - za);
    double s[3] = {0.0, 0.0, 0.0};
    for (uint64_t k = (uint64_t)lane; k < cells; k += 64) {
        const uint64_t row = k / wx;
        const uint64_t i = ((uint64_t)(za + row / wy) * ny + (ya + row % wy)) * nx + xa + (k - row * wx);
        s[0] += double(fx[i]);
        s[1] += double(fy[i]);
        s[2] += double(fz[i]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += __shfl_down(s[c], off);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < 3; ++c) out[p * 3 + c] = OutT(s[c] / double(cells));
}

// H_demag of every layer of a stack at the layers' present m.
void stack_eval_at_m(Stack &s) {
    std::vector<const void *> src;
    for (Context *l : s.layers) src.push_back(l->m);
    stack_demag_eval(s, src.data(), true);
    for (Context *l : s.layers) {
        l->hd_at_m = true;
        l->obs_valid = false;
        l->obs_frozen = false;
    }
}

// Typed driver for one precision mode.
template <typename State, typename Real>
struct Driver {
    Context &c;
    explicit Driver(Context &cc) : c(cc) {}

    State *m() { return static_cast<State *>(c.m); }
    State *mst() { return static_cast<State *>(c.mst); }
    Real *hd() { return static_cast<Real *>(c.hd); }
    Real *mt(int i) { return static_cast<Real *>(c.mt[i]); }
    Real *k(int i) { return static_cast<Real *>(c.k[i]); }
    static constexpr bool kStateF64 = sizeof(State) == 8;
    static constexpr bool kRealF64 = sizeof(Real) == 8;

    unsigned grid_plain() const { return (unsigned)std::max<uint64_t>(1, cdiv(c.n, kBlock)); }
    unsigned grid_stage() const { return stage_blocks(c); }
    unsigned grid_reduce() const { return grid_stage(); }

    template <typename Src, typename Out>
    StageArgs<State, Real, Src, Out> args(const Src *src, const Real *hdp) {
        StageArgs<State, Real, Src, Out> a{};
        a.f = field_args(c);
        a.src = src;
        a.hd = c.demag ? hdp : nullptr;
        a.m0 = m();
        a.nprev = 0;
        a.anew = 0;
        a.dt = 0;
        a.kout = nullptr;
        a.out = nullptr;
        a.partials = c.d_partials;
        return a;
    }

    template <typename Src, typename Out>
    void stage(const Src *src, Real *kout, Out *out, std::initializer_list<std::pair<int, double>> prev, double anew,
               double dt) {
        StageArgs<State, Real, Src, Out> a = args<Src, Out>(src, hd());
        a.kout = kout;
        a.out = out;
        for (auto &p : prev) {
            a.kp[a.nprev] = k(p.first);
            a.ap[a.nprev] = p.second;
            ++a.nprev;
        }
        a.anew = anew;
        a.dt = dt;
        Timed t(c, "llg_stage");
        const dim3 g(grid_stage()), b(kBlock);
        // the uniform kernel (or its multi-instance twin while a batch records)
        auto uniform_m = [&](auto np, auto m0src, auto masked, auto tier) {
            constexpr int NP = decltype(np)::value;
            constexpr bool MS = decltype(m0src)::value, MK = decltype(masked)::value, TR = decltype(tier)::value;
            launch_rec(c,
                       reinterpret_cast<const void *>(
                           &stage_uniform_multi<State, Real, Src, Out, STAGE_PLAIN, NP, MS, true, MK, TR>),
                       g, b, 0, 1, a, [&] {
                           hipLaunchKernelGGL((stage_uniform_kernel<State, Real, Src, Out, STAGE_PLAIN, NP, MS, true, MK, TR>),
                                              g, b, 0, c.stream, a);
                       });
        };
        auto uniform = [&](auto np, auto m0src) {
            if (a.f.tier_ex) {
                if (a.f.mask) uniform_m(np, m0src, std::true_type{}, std::true_type{});
                else uniform_m(np, m0src, std::false_type{}, std::true_type{});
            } else {
                if (a.f.mask) uniform_m(np, m0src, std::true_type{}, std::false_type{});
                else uniform_m(np, m0src, std::false_type{}, std::false_type{});
            }
        };
        if (out && stage_uniform_ok(a.f, a.hd, a.nprev)) {
            constexpr bool src_is_state = sizeof(Src) == sizeof(State);
            const bool m0_is_src = src_is_state && static_cast<const void *>(src) == static_cast<const void *>(a.m0);
            if (a.nprev == 1) uniform(std::integral_constant<int, 1>{}, std::false_type{});
            else if (m0_is_src) uniform(std::integral_constant<int, 0>{}, std::integral_constant<bool, src_is_state>{});
            else uniform(std::integral_constant<int, 0>{}, std::false_type{});
            return;
        }
        launch_rec(c, nullptr, g, b, 0, 1, a, [&] {
            if (a.f.tier_ex)
                hipLaunchKernelGGL((stage_kernel<State, Real, Src, Out, STAGE_PLAIN, true>), g, b, 0, c.stream, a);
            else
                hipLaunchKernelGGL((stage_kernel<State, Real, Src, Out, STAGE_PLAIN>), g, b, 0, c.stream, a);
        });
    }

    template <typename Out>
    void combine(Out *out, std::initializer_list<std::pair<int, double>> prev, double dt) {
        StageArgs<State, Real, State, Out> a = args<State, Out>(m(), hd());
        a.out = out;
        for (auto &p : prev) {
            a.kp[a.nprev] = k(p.first);
            a.ap[a.nprev] = p.second;
            ++a.nprev;
        }
        a.dt = dt;
        Timed t(c, "llg_combine");
        hipLaunchKernelGGL((combine_kernel<State, Real, Out>), dim3(grid_plain()), dim3(kBlock), 0, c.stream, m(), a);
    }

    // Max over cells of |dt sum_j w_j k_j|, evaluating the last slope at src.
    double error_stage(const State *src, Real *kout, std::initializer_list<std::pair<int, double>> prev, double wnew,
                       double dt) {
        StageArgs<State, Real, State, Real> a = args<State, Real>(src, hd());
        a.kout = kout;
        for (auto &p : prev) {
            a.kp[a.nprev] = k(p.first);
            a.ap[a.nprev] = p.second;
            ++a.nprev;
        }
        a.anew = wnew;
        a.dt = dt;
        a.partials = c.h_partials;
        const unsigned g = grid_reduce();
        {
            Timed t(c, "llg_error");
            if (a.f.tier_ex)
                hipLaunchKernelGGL((stage_kernel<State, Real, State, Real, STAGE_ERROR, true>), dim3(g), dim3(kBlock),
                                   0, c.stream, a);
            else
                hipLaunchKernelGGL((stage_kernel<State, Real, State, Real, STAGE_ERROR>), dim3(g), dim3(kBlock), 0,
                                   c.stream, a);
        }
        hipStreamSynchronize(c.stream);
        timing_resolve(c);
        double e = 0.0;
        for (unsigned b = 0; b < g; ++b) e = std::fmax(e, c.h_partials[b]);
        return std::sqrt(e);
    }

    // A stack layer's field couples to the other layers: it is evaluated for
    // the whole stack at the layers' m (the stack steps its layers itself).
    void demag_at_state(const State *src) {
        if (c.stack) stack_eval_at_m(*c.stack);
        else if (c.demag) demag_eval(c, src, kStateF64, hd());
    }
    void demag_at_stage(const Real *src) {
        if (c.demag && !c.stack) demag_eval(c, src, kRealF64, hd());
    }
    void ensure_hd() {
        if (!c.hd_at_m) demag_at_state(m());
        c.hd_at_m = true;
    }

    // Diagnostics at m (hd must hold H_demag(m)); optionally stores dm/dt
    // and, with `out`, the predictor n(m + dt dm/dt) as stage() would.
    void stats(Real *kout, Real *out = nullptr, double dt = 0.0) {
        StageArgs<State, Real, State, Real> a = args<State, Real>(m(), hd());
        a.kout = kout;
        a.out = out;
        a.anew = 1.0;
        a.dt = dt;
        a.partials = c.h_partials;
        const unsigned g = grid_reduce();
        {
            Timed t(c, "llg_stats");
            if (stage_uniform_ok(a.f, a.hd, 0)) {
                auto launch = [&](auto has_out, auto masked, auto tier) {
                    constexpr bool HO = decltype(has_out)::value, MK = decltype(masked)::value,
                                   TR = decltype(tier)::value;
                    hipLaunchKernelGGL((stage_uniform_kernel<State, Real, State, Real, STAGE_STATS, 0, true, HO, MK, TR>),
                                       dim3(g), dim3(kBlock), 0, c.stream, a);
                };
                auto with_tier = [&](auto has_out, auto masked) {
                    if (a.f.tier_ex) launch(has_out, masked, std::true_type{});
                    else launch(has_out, masked, std::false_type{});
                };
                auto with_mask = [&](auto has_out) {
                    if (a.f.mask) with_tier(has_out, std::true_type{});
                    else with_tier(has_out, std::false_type{});
                };
                if (out) with_mask(std::true_type{});
                else with_mask(std::false_type{});
            } else if (a.f.tier_ex) {
                hipLaunchKernelGGL((stage_kernel<State, Real, State, Real, STAGE_STATS, true>), dim3(g), dim3(kBlock),
                                   0, c.stream, a);
            } else {
                hipLaunchKernelGGL((stage_kernel<State, Real, State, Real, STAGE_STATS>), dim3(g), dim3(kBlock), 0,
                                   c.stream, a);
            }
        }
        c.stats_blocks = g;
    }

    // The stats pass's workgroups store their partials (compensated energy
    // sums and maxima) in pinned host memory; once the stream has passed the
    // pass, the host merges them in workgroup order into h_result (the
    // merge of neu_merge), which spares the device a one-workgroup
    // reduction launch per step.
    void merge_stats() {
        double acc[kNStat];
        for (int q = 0; q < kNStat; ++q) acc[q] = 0.0;
        const double *p = c.h_partials;
        for (unsigned b = 0; b < c.stats_blocks; ++b, p += kNStat) {
            for (int e = 0; e < 3; ++e) {
                double &sum = acc[2 * e], &comp = acc[2 * e + 1];
                const double v = p[2 * e], t = sum + v;
                comp += std::fabs(sum) >= std::fabs(v) ? (sum - t) + v : (v - t) + sum;
                sum = t;
                comp += p[2 * e + 1];
            }
            for (int q = 6; q < kNStat; ++q) acc[q] = std::fmax(acc[q], p[q]);
        }
        for (int q = 0; q < kNStat; ++q) c.h_result[q] = acc[q];
    }

    // lib.rs:1217-1249.  With use_pre, a predictor left by the previous
    // step's stats pass (same m, same H, same dt) replaces the first stage.
    // The predictor's demag transforms may already be in the spectrum
    // (enqueued by the previous step after its stats; spec_pre): then only
    // the x C2R remains.
    void heun(double dt, bool use_pre = false) {
        ensure_hd();
        const bool reuse = use_pre && c.pre_valid && c.pre_dt == dt;
        if (!reuse) stage<State, Real>(m(), k(0), mt(0), {}, 1.0, dt);
        c.pre_valid = false;
        if (reuse && c.spec_pre) demag_eval(c, mt(0), kRealF64, hd(), true, kDemagFinish);
        else demag_at_stage(mt(0));
        c.spec_pre = false;
        stage<Real, State>(mt(0), nullptr, m(), {{0, 0.5}}, 0.5, dt);
        demag_at_state(m());
        c.hd_at_m = true;
    }

    // Stats at the new state.  For Heun the same pass stores the next step's
    // first slope and predictor for a step of the same dt.
    void heun_aware_stats(double dt) {
        if (c.integrator == FULLMAG_FDM_INTEGRATOR_HEUN) {
            stats(k(0), mt(0), dt);
            c.pre_valid = true;
            c.pre_dt = dt;
        } else {
            stats(nullptr);
        }
    }

    // lib.rs:1254-1302
    void rk4(double dt) {
        ensure_hd();
        stage<State, Real>(m(), k(0), mt(0), {}, 0.5, dt);
        demag_at_stage(mt(0));
        stage<Real, Real>(mt(0), k(1), mt(1), {}, 0.5, dt);
        demag_at_stage(mt(1));
        stage<Real, Real>(mt(1), k(2), mt(0), {}, 1.0, dt);
        demag_at_stage(mt(0));
        stage<Real, State>(mt(0), nullptr, m(), {{0, 1.0 / 6.0}, {1, 2.0 / 6.0}, {2, 2.0 / 6.0}}, 1.0 / 6.0, dt);
        demag_at_state(m());
        c.hd_at_m = true;
    }

    // lib.rs:1332-1407 (no FSAL); returns the suggested next dt.
    double rk23(double &dt) {
        dt = std::max(std::min(dt, c.dt_max), c.dt_min);
        for (;;) {
            ensure_hd();
            stage<State, Real>(m(), k(0), mt(0), {}, 0.5, dt);
            demag_at_stage(mt(0));
            stage<Real, Real>(mt(0), k(1), mt(1), {}, 0.75, dt);
            demag_at_stage(mt(1));
            stage<Real, State>(mt(1), k(2), mst(), {{0, 2.0 / 9.0}, {1, 1.0 / 3.0}}, 4.0 / 9.0, dt);
            demag_at_state(mst());
            const double err =
                error_stage(mst(), nullptr, {{0, -5.0 / 72.0}, {1, 1.0 / 12.0}, {2, 1.0 / 9.0}}, -1.0 / 8.0, dt);
            if (err <= c.max_err || dt <= c.dt_min) {
                std::swap(c.m, c.mst);
                c.hd_at_m = true;
                return std::min(std::max(c.headroom * dt * std::pow(c.max_err / std::max(err, 1e-30), 1.0 / 3.0),
                                         c.dt_min),
                                c.dt_max);
            }
            c.hd_at_m = false;
            dt = std::min(std::max(c.headroom * dt * std::pow(c.max_err / err, 1.0 / 3.0), c.dt_min), c.dt_max);
        }
    }

    // lib.rs:1412-1568 (Dormand-Prince 5(4), FSAL)
    double dp45(double &dt) {
        constexpr double A21 = 1.0 / 5.0;
        constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
        constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
        constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                         A54 = -212.0 / 729.0;
        constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0, A64 = 49.0 / 176.0,
                         A65 = -5103.0 / 18656.0;
        constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0, B5 = -2187.0 / 6784.0,
                         B6 = 11.0 / 84.0;
        constexpr double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0, E5 = -17253.0 / 339200.0,
                         E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;
        dt = std::max(std::min(dt, c.dt_max), c.dt_min);
        for (;;) {
            if (c.fsal_valid) {
                combine<Real>(mt(0), {{0, A21}}, dt);
            } else {
                ensure_hd();
                stage<State, Real>(m(), k(0), mt(0), {}, A21, dt);
            }
            c.fsal_valid = false;
            demag_at_stage(mt(0));
            stage<Real, Real>(mt(0), k(1), mt(1), {{0, A31}}, A32, dt);
            demag_at_stage(mt(1));
            stage<Real, Real>(mt(1), k(2), mt(0), {{0, A41}, {1, A42}}, A43, dt);
            demag_at_stage(mt(0));
            stage<Real, Real>(mt(0), k(3), mt(1), {{0, A51}, {1, A52}, {2, A53}}, A54, dt);
            demag_at_stage(mt(1));
            stage<Real, Real>(mt(1), k(4), mt(0), {{0, A61}, {1, A62}, {2, A63}, {3, A64}}, A65, dt);
            demag_at_stage(mt(0));
            stage<Real, State>(mt(0), k(5), mst(), {{0, B1}, {2, B3}, {3, B4}, {4, B5}}, B6, dt);
            demag_at_state(mst());
            const double err = error_stage(mst(), k(6), {{0, E1}, {2, E3}, {3, E4}, {4, E5}, {5, E6}}, E7, dt);
            if (err <= c.max_err || dt <= c.dt_min) {
                std::swap(c.m, c.mst);
                std::swap(c.k[0], c.k[6]);  // FSAL: k7 becomes the next k1
                c.fsal_valid = true;
                c.hd_at_m = true;
                return std::min(std::max(c.headroom * dt * std::pow(c.max_err / std::max(err, 1e-30), 0.2),
                                         c.dt_min),
                                c.dt_max);
            }
            c.hd_at_m = false;
            dt = std::min(std::max(c.headroom * dt * std::pow(c.max_err / err, 0.2), c.dt_min), c.dt_max);
        }
    }

    // lib.rs:1573-1648; the history push restarts it when dt changed by more
    // than 10% (lib.rs:663-674), after the step.  Also launches the stats
    // (the startup branch stores f(m_new) from them).
    void abm3(double dt) {
        const bool ready = c.abm_startup >= 3 && c.abm_count >= 3;
        const int fn = c.abm_slot[0], fn1 = c.abm_slot[1], fn2 = c.abm_slot[2];
        if (!ready) {
            heun(dt);
            stats(k(fn2));  // f(m_new) into the slot that is about to be recycled
        } else {
            combine<Real>(mt(0), {{fn, 23.0 / 12.0}, {fn1, -16.0 / 12.0}, {fn2, 5.0 / 12.0}}, dt);
            demag_at_stage(mt(0));
            stage<Real, State>(mt(0), k(fn2), m(), {{fn, 8.0 / 12.0}, {fn1, -1.0 / 12.0}}, 5.0 / 12.0, dt);
            demag_at_state(m());
            c.hd_at_m = true;
            stats(nullptr);
        }
        if (c.abm_last_dt > 0.0 && std::fabs(dt - c.abm_last_dt) / c.abm_last_dt > 0.1) {
            c.abm_count = 0;
            c.abm_startup = 0;
        }
        c.abm_slot[0] = fn2;
        c.abm_slot[1] = fn;
        c.abm_slot[2] = fn1;
        c.abm_count = std::min(c.abm_count + 1, 3);
        c.abm_startup = std::min(c.abm_startup + 1, 3);
        c.abm_last_dt = dt;
    }

    int step(double dt_in, fullmag_fdm_step_stats *st) {
        const auto t0 = std::chrono::steady_clock::now();
        double dt = dt_in, suggested = 0.0;
        bool stats_pending = false;
        switch (c.integrator) {
            case FULLMAG_FDM_INTEGRATOR_RK4: rk4(dt); break;
            case FULLMAG_FDM_INTEGRATOR_RK23: suggested = rk23(dt); break;
            case FULLMAG_FDM_INTEGRATOR_DP45: suggested = dp45(dt); break;
            case FULLMAG_FDM_INTEGRATOR_ABM3: abm3(dt); stats_pending = true; break;
            case FULLMAG_FDM_INTEGRATOR_HEUN:
            default: heun(dt, true); break;
        }
        if (!stats_pending) heun_aware_stats(dt);
        c.obs_valid = false;
        c.obs_frozen = false;
        // A Heun step leaves the next step's predictor (for the same dt).
        // Its demag transforms, which write only the spectrum, are enqueued
        // behind the stats and the host waits for the stats alone, so the
        // device keeps working while the caller takes the stats and issues
        // its next call.  Anything but that next step clears spec_pre (any
        // other demag evaluation) or pre_valid, and the work is discarded.
        const bool ahead = c.integrator == FULLMAG_FDM_INTEGRATOR_HEUN && c.pre_valid && c.demag && !c.stack &&
                           c.ev_stats && hipEventRecord(c.ev_stats, c.stream) == hipSuccess;
        if (ahead) {
            demag_eval(c, mt(0), kRealF64, nullptr, true, kDemagSpectrum);
            c.spec_pre = true;
        }
        if (!hip_ok(ahead ? hipEventSynchronize(c.ev_stats) : hipStreamSynchronize(c.stream), "integrator step",
                    c.last_error))
            return FULLMAG_FDM_ERR_CUDA;
        if (!hip_ok(hipGetLastError(), "integrator step launch", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
        timing_resolve(c);
        c.step_count++;
        c.time += dt;
        fill(st, dt, suggested);
        st->wall_time_ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                               std::chrono::steady_clock::now() - t0)
                               .count();
        return FULLMAG_FDM_OK;
    }

    // Runs n fixed-dt steps; diagnostics only after the last one.
    int step_many(double dt, uint64_t nsteps, fullmag_fdm_step_stats *st) {
        const auto t0 = std::chrono::steady_clock::now();
        double suggested = 0.0;
        for (uint64_t s = 0; s < nsteps; ++s) {
            double d = dt;
            switch (c.integrator) {
                case FULLMAG_FDM_INTEGRATOR_RK4: rk4(d); break;
                case FULLMAG_FDM_INTEGRATOR_RK23: suggested = rk23(d); break;
                case FULLMAG_FDM_INTEGRATOR_DP45: suggested = dp45(d); break;
                case FULLMAG_FDM_INTEGRATOR_ABM3: abm3(d); break;
                default: heun(d, s == 0); break;
            }
            c.step_count++;
            c.time += d;
            if (s + 1 == nsteps) {
                if (c.integrator != FULLMAG_FDM_INTEGRATOR_ABM3) heun_aware_stats(d);
                fill_pending_dt_ = d;
            }
        }
        c.obs_valid = false;
        c.obs_frozen = false;
        if (!hip_ok(hipStreamSynchronize(c.stream), "integrator steps", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
        if (!hip_ok(hipGetLastError(), "integrator steps launch", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
        timing_resolve(c);
        if (st) {
            fill(st, fill_pending_dt_, suggested);
            st->wall_time_ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                   std::chrono::steady_clock::now() - t0)
                                   .count();
        }
        return FULLMAG_FDM_OK;
    }
    double fill_pending_dt_ = 0.0;

    void fill(fullmag_fdm_step_stats *st, double dt, double suggested) {
        merge_stats();
        const double *r = c.h_result;
        const double v = c.dx * c.dy * c.dz;
        const double w = c.has_ms_cell ? 1.0 : c.ms;
        const double e_ex = c.exchange ? -0.5 * kMu0 * w * v * (r[0] + r[1]) : 0.0;
        const double e_d = c.demag ? -0.5 * kMu0 * w * v * (r[2] + r[3]) : 0.0;
        const double e_ext = c.has_ext ? -kMu0 * w * v * (r[4] + r[5]) : 0.0;
        st->step = c.step_count;
        st->time_seconds = c.time;
        st->dt_seconds = dt;
        st->exchange_energy_joules = e_ex;
        st->demag_energy_joules = e_d;
        st->external_energy_joules = e_ext;
        st->total_energy_joules = e_ex + e_d + e_ext;
        st->max_effective_field_amplitude = std::sqrt(r[6]);
        st->max_demag_field_amplitude = c.demag ? std::sqrt(r[7]) : 0.0;
        st->max_rhs_amplitude = std::sqrt(r[8]);
        st->suggested_next_dt = suggested;
    }

    bool refresh() {
        demag_at_state(m());
        c.hd_at_m = true;
        c.obs_valid = false;
        c.obs_frozen = false;
        return hip_ok(hipStreamSynchronize(c.stream), "refresh_observables", c.last_error);
    }

    bool observe() {
        if (!c.h_ex_obs && !ctx_alloc(c, &c.h_ex_obs, 3 * c.n * sizeof(Real), "H_ex observable")) return false;
        if (!c.h_eff_obs && !ctx_alloc(c, &c.h_eff_obs, 3 * c.n * sizeof(Real), "H_eff observable")) return false;
        FieldArgs f = field_args(c);
        const Real *hdp = c.demag ? hd() : nullptr;
        Real *hex = static_cast<Real *>(c.h_ex_obs), *heff = static_cast<Real *>(c.h_eff_obs);
        if (f.tier_ex)
            hipLaunchKernelGGL((observe_kernel<State, Real, true>), dim3(grid_plain()), dim3(kBlock), 0, c.stream, f,
                               m(), hdp, hex, heff);
        else
            hipLaunchKernelGGL((observe_kernel<State, Real, false>), dim3(grid_plain()), dim3(kBlock), 0, c.stream, f,
                               m(), hdp, hex, heff);
        c.obs_valid = true;
        return hip_ok(hipStreamSynchronize(c.stream), "observables", c.last_error);
    }
};

}  // namespace

bool alloc_integrator_buffers(Context &c) {
    const size_t sb = 3 * c.n * c.state_bytes(), rb = 3 * c.n * c.real_bytes();
    if (!ctx_alloc(c, &c.m, sb, "m")) return false;
    if (c.demag && !ctx_alloc(c, &c.hd, rb, "H_demag")) return false;
    int nk = 1, nmt = 1;
    bool need_mst = false;
    switch (c.integrator) {
        case FULLMAG_FDM_INTEGRATOR_RK4: nk = 3; nmt = 2; break;
        case FULLMAG_FDM_INTEGRATOR_RK23: nk = 3; nmt = 2; need_mst = true; break;
        case FULLMAG_FDM_INTEGRATOR_DP45: nk = 7; nmt = 2; need_mst = true; break;
        case FULLMAG_FDM_INTEGRATOR_ABM3: nk = 7; nmt = 1; break;
        default: break;
    }
    for (int i = 0; i < nk; ++i)
        if (!ctx_alloc(c, &c.k[i], rb, "stage slope")) return false;
    for (int i = 0; i < nmt; ++i)
        if (!ctx_alloc(c, &c.mt[i], rb, "stage point")) return false;
    if (need_mst && !ctx_alloc(c, &c.mst, sb, "candidate state")) return false;
    {
        // any z chunking (down to one plane per workgroup) fits
        int tl, tx, ty, zc, nch;
        stage_tiling(c, tl, tx, ty, zc, nch);
        c.max_partial_blocks = (int)std::max<uint64_t>((uint64_t)tx * ty * c.nz, cdiv(c.n, kBlock));
    }
    if (!ctx_alloc(c, (void **)&c.d_partials, (size_t)c.max_partial_blocks * kNStat * sizeof(double), "partials"))
        return false;
    if (!ctx_alloc(c, (void **)&c.d_result, 16 * sizeof(double), "result")) return false;
    // coherent: the stats and error passes store their workgroup partials here
    if (!hip_ok(hipHostMalloc((void **)&c.h_partials, (size_t)c.max_partial_blocks * kNStat * sizeof(double),
                              hipHostMallocCoherent),
                "hipHostMalloc(partials)", c.last_error))
        return false;
    if (!hip_ok(hipEventCreateWithFlags(&c.ev_stats, hipEventDisableTiming), "hipEventCreate", c.last_error))
        return false;
    if (!hip_ok(hipHostMalloc((void **)&c.h_result, 16 * sizeof(double), hipHostMallocDefault), "hipHostMalloc",
                c.last_error))
        return false;
    return true;
}

// -- graph-captured batches (fixed-dt Heun) ----------------------------------
bool integrator_prepare(Context &c) {
    if (c.integrator != FULLMAG_FDM_INTEGRATOR_HEUN) {
        c.last_error = "graph-captured batches support the Heun integrator only";
        return false;
    }
    c.timing = false;
    c.pre_valid = false;  // captured steps always run their first stage
    if (c.mode == MODE_DOUBLE) Driver<double, double>(c).ensure_hd();
    else if (c.mode == MODE_MIXED) Driver<double, float>(c).ensure_hd();
    else Driver<float, float>(c).ensure_hd();
    return hip_ok(hipStreamSynchronize(c.stream), "batch prepare", c.last_error);
}

void integrator_enqueue_heun(Context &c, double dt, int nsteps) {
    for (int s = 0; s < nsteps; ++s) {
        if (c.mode == MODE_DOUBLE) Driver<double, double>(c).heun(dt);
        else if (c.mode == MODE_MIXED) Driver<double, float>(c).heun(dt);
        else Driver<float, float>(c).heun(dt);
    }
}

int integrator_finish(Context &c, double dt, uint64_t nsteps, fullmag_fdm_step_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    if (c.mode == MODE_DOUBLE) Driver<double, double>(c).stats(nullptr);
    else if (c.mode == MODE_MIXED) Driver<double, float>(c).stats(nullptr);
    else Driver<float, float>(c).stats(nullptr);
    c.pre_valid = false;  // the replays moved m
    c.obs_valid = false;
    c.obs_frozen = false;
    if (!hip_ok(hipStreamSynchronize(c.stream), "batch steps", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
    if (!hip_ok(hipGetLastError(), "batch steps launch", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
    c.step_count += nsteps;
    c.time += dt * (double)nsteps;
    if (st) {
        if (c.mode == MODE_DOUBLE) Driver<double, double>(c).fill(st, dt, 0.0);
        else if (c.mode == MODE_MIXED) Driver<double, float>(c).fill(st, dt, 0.0);
        else Driver<float, float>(c).fill(st, dt, 0.0);
        st->wall_time_ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                               std::chrono::steady_clock::now() - t0)
                               .count();
    }
    return FULLMAG_FDM_OK;
}

// Device time of `reps` (demag + stats) evaluations at the current state with
// the context's present launch schedule (read-only for the state).
double integrator_probe_ms(Context &c, int reps) {
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    const bool timing = c.timing;
    c.timing = false;
    auto once = [&]() {
        if (c.mode == MODE_DOUBLE) {
            Driver<double, double> d(c);
            d.demag_at_state(d.m());
            d.stats(nullptr);
        } else if (c.mode == MODE_MIXED) {
            Driver<double, float> d(c);
            d.demag_at_state(d.m());
            d.stats(nullptr);
        } else {
            Driver<float, float> d(c);
            d.demag_at_state(d.m());
            d.stats(nullptr);
        }
    };
    once();  // warm
    hipEventRecord(a, c.stream);
    for (int r = 0; r < reps; ++r) once();
    hipEventRecord(b, c.stream);
    hipEventSynchronize(b);
    float ms = 0.0f;
    hipEventElapsedTime(&ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    c.timing = timing;
    c.hd_at_m = true;
    c.pre_valid = false;  // hd was recomputed under other schedules
    return (double)ms;
}

// Heun steps of a stack (crates/fullmag-runner/src/multilayer_reference.rs:
// 543-589): every stage evaluates all layers, then the coupled demag of all
// layers at the new points, then the next stage.  The stats pass at the new
// state also stores each layer's next first slope and predictor, as for one
// backend.
template <typename State, typename Real>
int stack_step_t(Stack &s, double dt, uint64_t nsteps, fullmag_fdm_step_stats *st) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t L = s.layers.size();
    std::vector<Driver<State, Real>> d;
    for (Context *c : s.layers) d.emplace_back(*c);
    auto eval = [&](bool at_m) {
        std::vector<const void *> src(L);
        for (size_t l = 0; l < L; ++l) src[l] = at_m ? (const void *)d[l].m() : (const void *)d[l].mt(0);
        stack_demag_eval(s, src.data(), at_m);
    };
    bool have_hd = true;
    for (Context *c : s.layers) have_hd = have_hd && c->hd_at_m;
    if (!have_hd) {
        // a layer's m changed since the last step: every layer's field, and
        // so every stored predictor, depends on it
        eval(true);
        for (Context *c : s.layers) c->pre_valid = false;
    }
    for (uint64_t k = 0; k < nsteps; ++k) {
        for (size_t l = 0; l < L; ++l) {
            Context &c = *s.layers[l];
            if (!(c.pre_valid && c.pre_dt == dt))
                d[l].template stage<State, Real>(d[l].m(), d[l].k(0), d[l].mt(0), {}, 1.0, dt);
            c.pre_valid = false;
        }
        eval(false);
        for (size_t l = 0; l < L; ++l) d[l].template stage<Real, State>(d[l].mt(0), nullptr, d[l].m(), {{0, 0.5}}, 0.5, dt);
        eval(true);
    }
    for (size_t l = 0; l < L; ++l) {
        Context &c = *s.layers[l];
        d[l].stats(d[l].k(0), d[l].mt(0), dt);
        c.pre_valid = true;
        c.pre_dt = dt;
    }
    Context &c0 = *s.layers[0];
    if (!hip_ok(hipStreamSynchronize(c0.stream), "stack steps", s.last_error)) return FULLMAG_FDM_ERR_CUDA;
    if (!hip_ok(hipGetLastError(), "stack steps launch", s.last_error)) return FULLMAG_FDM_ERR_CUDA;
    s.step_count += nsteps;
    s.time += dt * (double)nsteps;
    fullmag_fdm_step_stats sum{};
    for (size_t l = 0; l < L; ++l) {
        Context &c = *s.layers[l];
        timing_resolve(c);
        c.hd_at_m = true;
        c.obs_valid = false;
        c.obs_frozen = false;
        c.step_count += nsteps;
        c.time += dt * (double)nsteps;
        fullmag_fdm_step_stats one{};
        d[l].fill(&one, dt, 0.0);
        sum.exchange_energy_joules += one.exchange_energy_joules;
        sum.demag_energy_joules += one.demag_energy_joules;
        sum.external_energy_joules += one.external_energy_joules;
        sum.total_energy_joules += one.total_energy_joules;
        sum.max_effective_field_amplitude = std::max(sum.max_effective_field_amplitude, one.max_effective_field_amplitude);
        sum.max_demag_field_amplitude = std::max(sum.max_demag_field_amplitude, one.max_demag_field_amplitude);
        sum.max_rhs_amplitude = std::max(sum.max_rhs_amplitude, one.max_rhs_amplitude);
    }
    timing_resolve(s.conv);
    if (st) {
        *st = sum;
        st->step = s.step_count;
        st->time_seconds = s.time;
        st->dt_seconds = dt;
        st->suggested_next_dt = 0.0;
        st->wall_time_ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                               std::chrono::steady_clock::now() - t0)
                               .count();
    }
    return FULLMAG_FDM_OK;
}

int stack_step(Stack &s, double dt, uint64_t n_steps, fullmag_fdm_step_stats *st) {
    const int mode = s.layers[0]->mode;
    if (mode == MODE_DOUBLE) return stack_step_t<double, double>(s, dt, n_steps, st);
    if (mode == MODE_MIXED) return stack_step_t<double, float>(s, dt, n_steps, st);
    return stack_step_t<float, float>(s, dt, n_steps, st);
}

bool stack_refresh(Stack &s) {
    stack_eval_at_m(s);
    for (Context *c : s.layers) c->pre_valid = false;
    return hip_ok(hipStreamSynchronize(s.conv.stream), "stack refresh", s.last_error) &&
           hip_ok(hipGetLastError(), "stack refresh launch", s.last_error);
}

int integrator_step(Context &c, double dt, fullmag_fdm_step_stats *st) {
    if (c.mode == MODE_DOUBLE) return Driver<double, double>(c).step(dt, st);
    if (c.mode == MODE_MIXED) return Driver<double, float>(c).step(dt, st);
    return Driver<float, float>(c).step(dt, st);
}

int integrator_step_many(Context &c, double dt, uint64_t n, fullmag_fdm_step_stats *st) {
    if (c.mode == MODE_DOUBLE) return Driver<double, double>(c).step_many(dt, n, st);
    if (c.mode == MODE_MIXED) return Driver<double, float>(c).step_many(dt, n, st);
    return Driver<float, float>(c).step_many(dt, n, st);
}

bool refresh_observables(Context &c) {
    if (c.mode == MODE_DOUBLE) return Driver<double, double>(c).refresh();
    if (c.mode == MODE_MIXED) return Driver<double, float>(c).refresh();
    return Driver<float, float>(c).refresh();
}

bool materialize_observables(Context &c) {
    if (c.obs_valid || c.obs_frozen) return true;
    if (c.mode == MODE_DOUBLE) return Driver<double, double>(c).observe();
    if (c.mode == MODE_MIXED) return Driver<double, float>(c).observe();
    return Driver<float, float>(c).observe();
}

bool preview_field(Context &c, const void *field, bool field_f64, uint32_t pnx, uint32_t pny, uint32_t pnz,
                   uint32_t z0, uint32_t zs, void *dev_out, bool out_f64) {
    const uint64_t count = (uint64_t)pnx * pny * pnz;
    const dim3 grid((unsigned)cdiv(count * 64, 256)), block(256);  // a wavefront per bin
    const uint64_t n = c.n;
    if (field_f64) {
        const double *f = static_cast<const double *>(field);
        if (out_f64)
            hipLaunchKernelGGL((preview_kernel<double, double>), grid, block, 0, c.stream, f, f + n, f + 2 * n, c.nx,
                               c.ny, c.nz, pnx, pny, pnz, z0, zs, static_cast<double *>(dev_out));
        else
            hipLaunchKernelGGL((preview_kernel<double, float>), grid, block, 0, c.stream, f, f + n, f + 2 * n, c.nx,
                               c.ny, c.nz, pnx, pny, pnz, z0, zs, static_cast<float *>(dev_out));
    } else {
        const float *f = static_cast<const float *>(field);
        if (out_f64)
            hipLaunchKernelGGL((preview_kernel<float, double>), grid, block, 0, c.stream, f, f + n, f + 2 * n, c.nx,
                               c.ny, c.nz, pnx, pny, pnz, z0, zs, static_cast<double *>(dev_out));
        else
            hipLaunchKernelGGL((preview_kernel<float, float>), grid, block, 0, c.stream, f, f + n, f + 2 * n, c.nx,
                               c.ny, c.nz, pnx, pny, pnz, z0, zs, static_cast<float *>(dev_out));
    }
    return hip_ok(hipGetLastError(), "preview kernel", c.last_error);
}

}  // namespace fm
