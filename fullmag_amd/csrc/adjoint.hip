// adjoint.hip -- discrete adjoint of one Heun step (extension: the reference
// has no adjoint; this serves gradient-based inverse design through
// torch.autograd).
//
// Forward step (crates/fullmag-engine/src/lib.rs:1217-1249), per cell:
//   k1 = f(m0, H(m0)),  v = m0 + dt k1,  mt = n(v),
//   k2 = f(mt, H(mt)),  u = m0 + dt/2 (k1 + k2),  m1 = n(u),
// with f(m, H) = -gb (p m x H + alpha m x (m x H)), n(v) = v / |v| and
// H(m) = (2A / (mu0 Ms)) lap(m) - N * (Ms m) + H_ext on active cells.
// Reverse mode, given lam1 = dL/dm1:
//   (dn/dv)^T l   = (l - n (n.l)) / |v|
//   (df/dm)^T mu  = -gb (p H x mu + alpha ((m.H) mu + H (m.mu) - 2 m (H.mu)))
//   (df/dH)^T mu  = -gb (p mu x m + alpha (m (m.mu) - |m|^2 mu))  =: g
//   (dH/dm)^T g   = lap(pref g) + Ms (-N * g)   (lap and the Newell
//                   convolution are symmetric operators)
//   dL/dMs_i     += -g_i . H_ex,i / Ms_i + m_i . (-N * g)_i
//   dL/dH_ext    += sum over active cells of g.
// The forward recompute uses the integrator's own stage kernel, so k1 and
// mt are the ones the step produced; the two convolutions of g reuse the
// demag transforms with unit weighting.
#include "field.hpp"

#include <cmath>

namespace fm {
namespace {

template <typename Real>
struct AdjArgs {
    FieldArgs f;    // the backend's field parameters
    FieldArgs fq;   // the same stencil with unit prefactor (applied to q = pref g)
    double dt;
    const double *m0, *lam1;
    const Real *hd0, *hdt, *k1, *mt, *w;
    Real *g, *q;
    double *lu, *lmt, *lam0;
    double *gms;       // per-cell Ms gradient (accumulated) or null
    double *partials;  // [block][3] H_ext gradient sums of this pass
};

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// out += (df/dm)^T mu;  g = (df/dH)^T mu
__device__ __forceinline__ void torque_adjoint(const FieldArgs &f, const double m[3], const double h[3],
                                               const double mu[3], double out[3], double g[3]) {
    const double gb = f.gamma_bar, al = f.alpha, pe = f.precession ? 1.0 : 0.0;
    double hxm[3], mxm[3];
    cross3(h, mu, hxm);
    cross3(mu, m, mxm);
    const double mh = dot3(m, h), mm = dot3(m, mu), hm = dot3(h, mu), m2 = dot3(m, m);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        out[k] += -gb * (pe * hxm[k] + al * (mh * mu[k] + h[k] * mm - 2.0 * m[k] * hm));
        g[k] = -gb * (pe * mxm[k] + al * (m[k] * mm - m2 * mu[k]));
    }
}

// (l - y (y.l)) / |v|, y = v / |v|; zero for v = 0 (normalised zero stays zero)
__device__ __forceinline__ void norm_adjoint(const double v[3], const double l[3], double out[3]) {
    const double nv = sqrt(dot3(v, v));
    if (!(nv > 0.0)) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double inv = 1.0 / nv;
    const double y[3] = {v[0] * inv, v[1] * inv, v[2] * inv};
    const double yl = dot3(y, l);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = (l[k] - y[k] * yl) * inv;
}

__device__ __forceinline__ double ms_of(const FieldArgs &f, double ms_uniform, int64_t i) {
    return f.ms_cell ? f.ms_cell[i] : ms_uniform;
}

// Exchange prefactor of cell i: H_ex,i = pref_i (sum of the stencil).
__device__ __forceinline__ double pref_of(const FieldArgs &f, int64_t i) {
    if (f.region) return f.ms_cell ? f.two_over_mu0 / f.ms_cell[i] : f.pre_lut;
    return f.ms_cell ? f.two_over_mu0 * f.a_uniform / f.ms_cell[i] : f.pref;
}

// Block sums of g (3 components) into partials[blockIdx.x][3].
__device__ __forceinline__ void block_sum3(double s[3], double *partials) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] += __shfl_down(s[k], off);
    __shared__ double red[kBlock / 64][3];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int k = 0; k < 3; ++k) red[wave][k] = s[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) {
            double t = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w][k];
            partials[(int64_t)blockIdx.x * 3 + k] = t;
        }
    }
}

// Total field at cell i of state s (exchange in the arithmetic of S, as the
// stage kernel does) from the stored demag field.
template <typename S, typename Real>
__device__ __forceinline__ void field_at(const FieldArgs &f, const S *s, const Real *hd, int64_t i, double h[3],
                                         double hex[3]) {
    Real he[3] = {Real(0), Real(0), Real(0)};
    if (f.exchange) {
        S e[3];
        exchange_at<S, S>(f, s, i, e);
#pragma unroll
        for (int k = 0; k < 3; ++k) he[k] = Real(e[k]);
    }
    const Real ext[3] = {f.has_ext ? Real(f.hx) : Real(0), f.has_ext ? Real(f.hy) : Real(0),
                         f.has_ext ? Real(f.hz) : Real(0)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const Real d = hd ? hd[k * f.n + i] : Real(0);
        h[k] = double((he[k] + d) + ext[k]);
        hex[k] = double(he[k]);
    }
}

// Pass A: through m1 = n(u) and k2 = f(mt, H(mt)).
template <typename Real>
__global__ void __launch_bounds__(kBlock) adj_stage2_kernel(AdjArgs<Real> a, double ms_uniform) {
    const FieldArgs &f = a.f;
    const int64_t n = f.n, i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double gs[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const bool act = !f.mask || f.mask[i];
        double lu[3] = {0, 0, 0}, lmt[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        if (act) {
            double mt[3], ht[3], hext[3], k1[3], k2[3], m0[3], l1[3], u[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mt[k] = double(a.mt[k * n + i]);
                k1[k] = double(a.k1[k * n + i]);
                m0[k] = a.m0[k * n + i];
                l1[k] = a.lam1[k * n + i];
            }
            field_at<Real, Real>(f, a.mt, a.hdt, i, ht, hext);
            // k2 in the stage kernel's arithmetic (Real), u in FP64
            {
                Real mr[3] = {Real(mt[0]), Real(mt[1]), Real(mt[2])}, hr[3] = {Real(ht[0]), Real(ht[1]), Real(ht[2])},
                     kr[3];
                torque(f, mr, hr, kr);
#pragma unroll
                for (int k = 0; k < 3; ++k) k2[k] = double(kr[k]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) u[k] = m0[k] + a.dt * (0.5 * k2[k] + 0.5 * k1[k]);
            norm_adjoint(u, l1, lu);
            const double mu[3] = {0.5 * a.dt * lu[0], 0.5 * a.dt * lu[1], 0.5 * a.dt * lu[2]};
            torque_adjoint(f, mt, ht, mu, lmt, g);
            if (a.gms && f.exchange) a.gms[i] += -dot3(g, hext) / ms_of(f, ms_uniform, i);
            gs[0] = g[0];
            gs[1] = g[1];
            gs[2] = g[2];
        }
        const double pr = act ? pref_of(f, i) : 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.lu[k * n + i] = lu[k];
            a.lmt[k * n + i] = lmt[k];
            a.g[k * n + i] = Real(g[k]);
            a.q[k * n + i] = Real(pr * g[k]);
        }
    }
    block_sum3(gs, a.partials);
}

// Pass B: field adjoint at mt, through mt = n(v) and k1 = f(m0, H(m0)).
template <typename Real>
__global__ void __launch_bounds__(kBlock) adj_stage1_kernel(AdjArgs<Real> a, double ms_uniform) {
    const FieldArgs &f = a.f;
    const int64_t n = f.n, i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double gs[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const bool act = !f.mask || f.mask[i];
        double l0[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        if (act) {
            double lmt[3], lu[3], mt[3], m0[3], k1[3], v[3], h0[3], hex0[3];
            double lq[3] = {0, 0, 0};
            if (f.exchange) exchange_at<double, Real>(a.fq, a.q, i, lq);
            const double ms = ms_of(f, ms_uniform, i);
            double gms = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                mt[k] = double(a.mt[k * n + i]);
                m0[k] = a.m0[k * n + i];
                k1[k] = double(a.k1[k * n + i]);
                lu[k] = a.lu[k * n + i];
                const double w = a.w ? double(a.w[k * n + i]) : 0.0;
                lmt[k] = a.lmt[k * n + i] + lq[k] + ms * w;
                gms += mt[k] * w;
                v[k] = m0[k] + a.dt * k1[k];
            }
            if (a.gms && a.w) a.gms[i] += gms;
            double lv[3];
            norm_adjoint(v, lmt, lv);
            const double mu[3] = {0.5 * a.dt * lu[0] + a.dt * lv[0], 0.5 * a.dt * lu[1] + a.dt * lv[1],
                                  0.5 * a.dt * lu[2] + a.dt * lv[2]};
#pragma unroll
            for (int k = 0; k < 3; ++k) l0[k] = lu[k] + lv[k];
            field_at<double, Real>(f, a.m0, a.hd0, i, h0, hex0);
            torque_adjoint(f, m0, h0, mu, l0, g);
            if (a.gms && f.exchange) a.gms[i] += -dot3(g, hex0) / ms;
            gs[0] = g[0];
            gs[1] = g[1];
            gs[2] = g[2];
        }
        const double pr = act ? pref_of(f, i) : 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.lam0[k * n + i] = l0[k];
            a.g[k * n + i] = Real(g[k]);
            a.q[k * n + i] = Real(pr * g[k]);
        }
    }
    block_sum3(gs, a.partials);
}

// Pass C: field adjoint at m0.
template <typename Real>
__global__ void __launch_bounds__(kBlock) adj_field0_kernel(AdjArgs<Real> a, double ms_uniform) {
    const FieldArgs &f = a.f;
    const int64_t n = f.n, i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool act = !f.mask || f.mask[i];
    if (!act) {
#pragma unroll
        for (int k = 0; k < 3; ++k) a.lam0[k * n + i] = 0.0;
        return;
    }
    double lq[3] = {0, 0, 0};
    if (f.exchange) exchange_at<double, Real>(a.fq, a.q, i, lq);
    const double ms = ms_of(f, ms_uniform, i);
    double gms = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double w = a.w ? double(a.w[k * n + i]) : 0.0;
        a.lam0[k * n + i] += lq[k] + ms * w;
        gms += a.m0[k * n + i] * w;
    }
    if (a.gms && a.w) a.gms[i] += gms;
}

// Sums [nb][3] partials of `passes` consecutive passes in a fixed order.
__global__ void __launch_bounds__(kBlock) adj_sum3_kernel(const double *partials, int nb, int passes, double *out) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb * passes; b += kBlock)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] += partials[(int64_t)b * 3 + k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) s[k] += __shfl_down(s[k], off);
    __shared__ double red[kBlock / 64][3];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int k = 0; k < 3; ++k) red[wave][k] = s[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            double t = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) t += red[w][k];
            out[k] = t;
        }
}

template <typename Real>
int heun_adjoint_t(Context &c, double dt, const double *m0, const double *lam1, double *lam0, double *grad_ms,
                   double *grad_hext) {
    const uint64_t n = c.n;
    if (!c.adj_ready) {
        const size_t rb = 3 * n * sizeof(Real), db = 3 * n * sizeof(double);
        if (!ctx_alloc(c, &c.adj_hd0, rb, "adjoint H_demag(m0)") || !ctx_alloc(c, &c.adj_k1, rb, "adjoint k1") ||
            !ctx_alloc(c, &c.adj_mt, rb, "adjoint stage point") ||
            !ctx_alloc(c, &c.adj_hdt, rb, "adjoint H_demag(mt)") || !ctx_alloc(c, &c.adj_g, rb, "adjoint g") ||
            !ctx_alloc(c, &c.adj_q, rb, "adjoint q") || !ctx_alloc(c, &c.adj_w, rb, "adjoint w") ||
            !ctx_alloc(c, (void **)&c.adj_lu, db, "adjoint lam_u") ||
            !ctx_alloc(c, (void **)&c.adj_lmt, db, "adjoint lam_mt"))
            return FULLMAG_FDM_ERR_CUDA;
        c.adj_ready = true;
    }
    Real *hd0 = static_cast<Real *>(c.adj_hd0), *k1 = static_cast<Real *>(c.adj_k1);
    Real *mt = static_cast<Real *>(c.adj_mt), *hdt = static_cast<Real *>(c.adj_hdt);
    Real *g = static_cast<Real *>(c.adj_g), *q = static_cast<Real *>(c.adj_q), *w = static_cast<Real *>(c.adj_w);
    constexpr bool kRealF64 = sizeof(Real) == 8;

    // forward recompute with the integrator's stage kernel
    if (c.demag) demag_eval(c, m0, true, hd0);
    StageArgs<double, Real, double, Real> s{};
    s.f = field_args(c);
    s.src = m0;
    s.hd = c.demag ? hd0 : nullptr;
    s.m0 = m0;
    s.nprev = 0;
    s.anew = 1.0;
    s.dt = dt;
    s.kout = k1;
    s.out = mt;
    hipLaunchKernelGGL((stage_kernel<double, Real, double, Real, STAGE_PLAIN>), dim3(stage_blocks(c)),
                       dim3(kBlock), 0, c.stream, s);
    if (c.demag) demag_eval(c, mt, kRealF64, hdt);

    AdjArgs<Real> a{};
    a.f = field_args(c);
    a.fq = a.f;
    a.fq.pref = 1.0;
    a.fq.pre_lut = 1.0;
    a.fq.ms_cell = nullptr;
    a.dt = dt;
    a.m0 = m0;
    a.lam1 = lam1;
    a.hd0 = c.demag ? hd0 : nullptr;
    a.hdt = c.demag ? hdt : nullptr;
    a.k1 = k1;
    a.mt = mt;
    a.w = nullptr;
    a.g = g;
    a.q = q;
    a.lu = c.adj_lu;
    a.lmt = c.adj_lmt;
    a.lam0 = lam0;
    a.gms = grad_ms;
    const unsigned nb = (unsigned)std::max<uint64_t>(1, cdiv(n, kBlock));
    a.partials = c.d_partials;
    hipLaunchKernelGGL((adj_stage2_kernel<Real>), dim3(nb), dim3(kBlock), 0, c.stream, a, c.ms);
    if (c.demag) {
        demag_eval(c, g, kRealF64, w, false);
        a.w = w;
    }
    a.partials = c.d_partials + 3 * (size_t)nb;
    hipLaunchKernelGGL((adj_stage1_kernel<Real>), dim3(nb), dim3(kBlock), 0, c.stream, a, c.ms);
    if (c.demag) demag_eval(c, g, kRealF64, w, false);
    hipLaunchKernelGGL((adj_field0_kernel<Real>), dim3(nb), dim3(kBlock), 0, c.stream, a, c.ms);
    hipLaunchKernelGGL(adj_sum3_kernel, dim3(1), dim3(kBlock), 0, c.stream, c.d_partials, (int)nb, 2, c.d_result);
    if (!hip_ok(hipMemcpyAsync(c.h_result, c.d_result, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream),
                "adjoint result copy", c.last_error))
        return FULLMAG_FDM_ERR_CUDA;
    if (!hip_ok(hipStreamSynchronize(c.stream), "heun adjoint", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
    if (!hip_ok(hipGetLastError(), "heun adjoint launch", c.last_error)) return FULLMAG_FDM_ERR_CUDA;
    if (grad_hext && c.has_ext)
        for (int k = 0; k < 3; ++k) grad_hext[k] += c.h_result[k];
    return FULLMAG_FDM_OK;
}

}  // namespace

int heun_adjoint(Context &c, double dt, const double *m0, const double *lam1, double *lam0, double *grad_ms,
                 double *grad_hext) {
    if (c.mode == MODE_SINGLE) {
        c.last_error = "the Heun adjoint needs an FP64 state (precision double or mixed)";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (c.tier) {
        c.last_error = "the Heun adjoint does not cover the sub-cell boundary tiers";
        return FULLMAG_FDM_ERR_INVALID;
    }
    if (c.mode == MODE_DOUBLE) return heun_adjoint_t<double>(c, dt, m0, lam1, lam0, grad_ms, grad_hext);
    return heun_adjoint_t<float>(c, dt, m0, lam1, lam0, grad_ms, grad_hext);
}

}  // namespace fm
