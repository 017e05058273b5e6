// field.hpp -- per-cell field and torque building blocks shared by the
// integrator (llg.hip) and the adjoint (adjoint.hip): exchange stencils,
// torque, compensated sums, workgroup reductions, XCD-aware workgroup
// numbering, the fused stage kernel and its column tiling, kernel timing.
// Semantics follow crates/fullmag-engine/src/lib.rs (exchange :2247-2299,
// rhs :2584-2595).
#pragma once

#include "context.hpp"

#include <algorithm>

namespace fm {
namespace {

constexpr int kBlock = 256;
constexpr int kNStat = 9;  // s_ex, c_ex, s_d, c_d, s_ext, c_ext, max|Heff|^2, max|Hd|^2, max|k|^2
// Read in place of an absent per-cell Ms, so that the read needs no branch.
__device__ const double kUnitStandIn = 1.0;

struct FieldArgs {
    int nx, ny, nz;
    int64_t n;
    FastDiv div_nx, div_nxy;  // cell index -> (x, y, z); n < 2^32
    // column tiling of the stage kernels: a workgroup owns a TX x TY tile of
    // (x, y) columns (TX TY = kBlock) and marches z through zc planes
    int tx_log2, tiles_x, tiles_y, zc, nchunks;
    int xcd;  // renumber workgroups per XCD
    double idx2, idy2, idz2;  // 1 / d^2
    double pref;              // 2 A / (mu0 Ms) for uniform A and Ms
    double a_uniform;         // A when ms_cell is set and no table
    double two_over_mu0;
    double pre_lut;           // 2 / (mu0 Ms) for the per-pair table
    const uint8_t *mask;
    const uint32_t *region;  // non-null: per-pair A from `lut`
    const double *lut;
    const double *ms_cell;
    int exchange, has_ext, precession;
    double hx, hy, hz;
    double gamma_bar, alpha;
    // sub-cell boundary exchange: tier_ex 1 (T0 face links) or 2 (T1
    // intersection distances) replaces the stencil above
    int tier_ex;
    const double *phi;     // volume fraction
    const float *phi32;    // the same in FP32, in its place under mixed precision
    const double *tier_w;  // 6 n: xp, xm, yp, ym, zp, zm face links or distances
    const float *tier_w32; // the same in FP32, in their place under mixed precision
    double phi_floor, delta_min;
    double tier_pref;      // 2 A / (mu0 Ms) with the plan's scalar A
    double a_tier;         // the plan's scalar A (per-cell Ms)
    double cell[3];        // dx, dy, dz
};

// Exchange field at cell i in the arithmetic type C of the source, with
// clamped (Neumann) neighbours and inactive neighbours replaced by the centre
// (lib.rs:2270-2283).
template <typename C, typename S>
__device__ __forceinline__ void exchange_at(const FieldArgs &f, const S *__restrict__ m, int64_t i, C h[3]) {
    const int64_t nxy = (int64_t)f.nx * f.ny;
    const uint32_t ui = (uint32_t)i;
    const uint32_t z = f.div_nxy.div(ui);
    const uint32_t r = ui - z * (uint32_t)nxy;
    const uint32_t y = f.div_nx.div(r);
    const uint32_t x = r - y * (uint32_t)f.nx;
    int64_t xm = x > 0 ? i - 1 : i, xp = (int)x < f.nx - 1 ? i + 1 : i;
    int64_t ym = y > 0 ? i - f.nx : i, yp = (int)y < f.ny - 1 ? i + f.nx : i;
    int64_t zm = z > 0 ? i - nxy : i, zp = (int)z < f.nz - 1 ? i + nxy : i;
    if (f.mask) {
        if (!f.mask[xm]) xm = i;
        if (!f.mask[xp]) xp = i;
        if (!f.mask[ym]) ym = i;
        if (!f.mask[yp]) yp = i;
        if (!f.mask[zm]) zm = i;
        if (!f.mask[zp]) zp = i;
    }
    const C ix = C(f.idx2), iy = C(f.idy2), iz = C(f.idz2);
    if (!f.region) {
        const C pref = f.ms_cell ? C(f.two_over_mu0 * f.a_uniform / f.ms_cell[i]) : C(f.pref);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const S *mc = m + c * f.n;
            const C cc = C(mc[i]);
            const C lap = (C(mc[xp]) - C(2) * cc + C(mc[xm])) * ix + (C(mc[yp]) - C(2) * cc + C(mc[ym])) * iy +
                          (C(mc[zp]) - C(2) * cc + C(mc[zm])) * iz;
            h[c] = pref * lap;
        }
    } else {
        // per-pair coupling (native/backends/fdm/src/exchange_fp64.cu:80-114)
        const uint32_t ri = f.region[i];
        const double *row = f.lut + (size_t)ri * FULLMAG_FDM_MAX_EXCHANGE_REGIONS;
        const C axm = C(row[f.region[xm]]), axp = C(row[f.region[xp]]);
        const C aym = C(row[f.region[ym]]), ayp = C(row[f.region[yp]]);
        const C azm = C(row[f.region[zm]]), azp = C(row[f.region[zp]]);
        const double msi = f.ms_cell ? f.ms_cell[i] : 0.0;
        const C pre = f.ms_cell ? C(f.two_over_mu0 / msi) : C(f.pre_lut);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const S *mc = m + c * f.n;
            const C cc = C(mc[i]);
            const C s = (axp * (C(mc[xp]) - cc) + axm * (C(mc[xm]) - cc)) * ix +
                        (ayp * (C(mc[yp]) - cc) + aym * (C(mc[ym]) - cc)) * iy +
                        (azp * (C(mc[zp]) - cc) + azm * (C(mc[zm]) - cc)) * iz;
            h[c] = pre * s;
        }
    }
}

// exchange_at for a z-marching thread: the z neighbours' values arrive in
// registers (vzm, vzp as stored at i -+ nx ny, ignored where the neighbour is
// clamped or inactive); x and y neighbours are loaded.  Same operations in
// the same order as exchange_at.
template <typename C, typename S>
__device__ __forceinline__ void exchange_col(const FieldArgs &f, const S *__restrict__ m, int64_t i, int x, int y,
                                             int z, const S cc_s[3], const S vzm[3], const S vzp[3], C h[3]) {
    const int64_t nxy = (int64_t)f.nx * f.ny;
    int64_t xm = x > 0 ? i - 1 : i, xp = x < f.nx - 1 ? i + 1 : i;
    int64_t ym = y > 0 ? i - f.nx : i, yp = y < f.ny - 1 ? i + f.nx : i;
    bool zm_in = z > 0, zp_in = z < f.nz - 1;
    if (f.mask) {
        if (!f.mask[xm]) xm = i;
        if (!f.mask[xp]) xp = i;
        if (!f.mask[ym]) ym = i;
        if (!f.mask[yp]) yp = i;
        if (zm_in && !f.mask[i - nxy]) zm_in = false;
        if (zp_in && !f.mask[i + nxy]) zp_in = false;
    }
    const C ix = C(f.idx2), iy = C(f.idy2), iz = C(f.idz2);
    if (!f.region) {
        const C pref = f.ms_cell ? C(f.two_over_mu0 * f.a_uniform / f.ms_cell[i]) : C(f.pref);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const S *mc = m + c * f.n;
            const C cc = C(cc_s[c]);
            const C zmv = zm_in ? C(vzm[c]) : cc, zpv = zp_in ? C(vzp[c]) : cc;
            const C lap = (C(mc[xp]) - C(2) * cc + C(mc[xm])) * ix + (C(mc[yp]) - C(2) * cc + C(mc[ym])) * iy +
                          (zpv - C(2) * cc + zmv) * iz;
            h[c] = pref * lap;
        }
    } else {
        const int64_t zm = zm_in ? i - nxy : i, zp = zp_in ? i + nxy : i;
        const uint32_t ri = f.region[i];
        const double *row = f.lut + (size_t)ri * FULLMAG_FDM_MAX_EXCHANGE_REGIONS;
        const C axm = C(row[f.region[xm]]), axp = C(row[f.region[xp]]);
        const C aym = C(row[f.region[ym]]), ayp = C(row[f.region[yp]]);
        const C azm = C(row[f.region[zm]]), azp = C(row[f.region[zp]]);
        const double msi = f.ms_cell ? f.ms_cell[i] : 0.0;
        const C pre = f.ms_cell ? C(f.two_over_mu0 / msi) : C(f.pre_lut);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const S *mc = m + c * f.n;
            const C cc = C(cc_s[c]);
            const C zmv = zm_in ? C(vzm[c]) : cc, zpv = zp_in ? C(vzp[c]) : cc;
            const C s = (axp * (C(mc[xp]) - cc) + axm * (C(mc[xm]) - cc)) * ix +
                        (ayp * (C(mc[yp]) - cc) + aym * (C(mc[ym]) - cc)) * iy +
                        (azp * (zpv - cc) + azm * (zmv - cc)) * iz;
            h[c] = pre * s;
        }
    }
}

template <typename C, typename S>
__device__ __forceinline__ void exchange_tier_vals(const FieldArgs &f, const bool in[6], const S v[6][3],
                                                   const double wt[6], double phi0, double pre, const S cc[3],
                                                   C h[3]);

// Face link or distance of side s (-x, +x, -y, +y, -z, +z) at cell i: side s
// reads array s ^ 1 of xp, xm, yp, ym, zp, zm.  W = float (mixed precision,
// whose fields are FP32) reads the FP32 copy, W = double the FP64 arrays.
template <typename W>
__device__ __forceinline__ double tier_weight(const FieldArgs &f, int s, int64_t i) {
    if constexpr (sizeof(W) == 4) return double(f.tier_w32[(int64_t)(s ^ 1) * f.n + i]);
    else return f.tier_w[(int64_t)(s ^ 1) * f.n + i];
}
// Volume fraction at cell i, read like tier_weight.
template <typename W>
__device__ __forceinline__ double tier_phi(const FieldArgs &f, int64_t i) {
    if constexpr (sizeof(W) == 4) return double(f.phi32[i]);
    else return f.phi[i];
}

// Sub-cell boundary exchange at cell (x, y, z) = i, in the arithmetic type C
// of the source.  T0 (tier_ex 1, native/backends/fdm/src/exchange_t0_fp64.cu:
// 19-133): H = 2A / (mu0 Ms) / max(phi, phi_floor) sum f (m_j - m) / d^2 over
// the faces whose link f > 0; T1 (tier_ex 2, exchange_t1_fp64.cu:29-175):
// H = 2A / (mu0 Ms) sum 2 (m_j - m) / (d (d + delta')) over the neighbours with
// nonzero m, delta = max(delta_side, delta_min) and delta' the opposite side's.
// Neighbours beyond the grid add nothing (z only when nz > 1); cells with
// phi <= 0 or m = 0 get zero.  The active mask and the region table do not
// enter.  cc, vzm, vzp: m at i and at i -+ nx ny (read only where that
// neighbour is inside the grid).
template <typename C, typename S, typename W = double>
__device__ __forceinline__ void exchange_tier(const FieldArgs &f, const S *__restrict__ m, int64_t i, int x, int y,
                                              int z, const S cc[3], const S vzm[3], const S vzp[3], C h[3]) {
    const int64_t n = f.n;
    // sides in the order -x, +x, -y, +y, -z, +z; side s reads weight s ^ 1
    // of the xp, xm, yp, ym, zp, zm arrays.  Every load is issued
    // unconditionally (clamped to the cell itself beyond the grid) and a
    // side that does not count enters with coefficient zero, which adds
    // exactly nothing.
    const bool in[6] = {x > 0, x < f.nx - 1, y > 0, y < f.ny - 1, f.nz > 1 && z > 0, f.nz > 1 && z < f.nz - 1};
    const int64_t off[4] = {-1, 1, -(int64_t)f.nx, (int64_t)f.nx};
    S v[6][3];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int64_t j = in[s] ? i + off[s] : i;
#pragma unroll
        for (int q = 0; q < 3; ++q) v[s][q] = m[q * n + j];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        v[4][q] = vzm[q];
        v[5][q] = vzp[q];
    }
    double wt[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) wt[s] = tier_weight<W>(f, s, i);
    const double phi0 = tier_phi<W>(f, i);
    const double pre = f.ms_cell ? 2.0 * f.a_tier / (kMu0 * f.ms_cell[i]) : f.tier_pref;
    exchange_tier_vals<C, S>(f, in, v, wt, phi0, pre, cc, h);
}

// exchange_tier's arithmetic on loaded values: v[s] the neighbour m of side
// s (-x, +x, -y, +y, -z, +z; any value where in[s] is false), wt[s] its face
// link or distance, phi0 the cell's volume fraction, pre its 2 A / (mu0 Ms).
template <typename C, typename S>
__device__ __forceinline__ void exchange_tier_vals(const FieldArgs &f, const bool in[6], const S v[6][3],
                                                   const double wt[6], double phi0, double pre, const S cc[3],
                                                   C h[3]) {
    const C c0[3] = {C(cc[0]), C(cc[1]), C(cc[2])};
    C b[3] = {C(0), C(0), C(0)};
    double scale;
    if (f.tier_ex == 1) {
        const double id2[3] = {f.idx2, f.idy2, f.idz2};
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const C k = (wt[s] > 0.0 && in[s]) ? C(wt[s] * id2[s >> 1]) : C(0);
#pragma unroll
            for (int q = 0; q < 3; ++q) b[q] += k * (C(v[s][q]) - c0[q]);
        }
        scale = pre * (1.0 / (phi0 > f.phi_floor ? phi0 : f.phi_floor));
    } else {
        double dl[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) dl[s] = fmax(wt[s], f.delta_min);
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const bool nz = !(v[s][0] == S(0) && v[s][1] == S(0) && v[s][2] == S(0));
            const double d = f.cell[s >> 1];
            const C k = (in[s] && dl[s] > 0.0 && nz) ? C(2.0 / (d * (d + dl[s ^ 1]))) : C(0);
#pragma unroll
            for (int q = 0; q < 3; ++q) b[q] += k * (C(v[s][q]) - c0[q]);
        }
        scale = pre;
    }
    // empty cells (phi <= 0) and zero m get no field
    const bool live = phi0 > 0.0 && !(cc[0] == S(0) && cc[1] == S(0) && cc[2] == S(0));
#pragma unroll
    for (int q = 0; q < 3; ++q) h[q] = live ? C(scale) * b[q] : C(0);
}

// exchange_tier at cell i with every neighbour loaded from memory.
template <typename C, typename S, typename W = double>
__device__ __forceinline__ void exchange_tier_at(const FieldArgs &f, const S *__restrict__ m, int64_t i, C h[3]) {
    const int64_t nxy = (int64_t)f.nx * f.ny;
    const uint32_t ui = (uint32_t)i;
    const uint32_t z = f.div_nxy.div(ui);
    const uint32_t r = ui - z * (uint32_t)nxy;
    const uint32_t y = f.div_nx.div(r);
    const uint32_t x = r - y * (uint32_t)f.nx;
    const int64_t zm = z > 0 ? i - nxy : i, zp = (int)z < f.nz - 1 ? i + nxy : i;
    S cc[3], vzm[3], vzp[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        cc[q] = m[q * f.n + i];
        vzm[q] = m[q * f.n + zm];
        vzp[q] = m[q * f.n + zp];
    }
    exchange_tier<C, S, W>(f, m, i, (int)x, (int)y, (int)z, cc, vzm, vzp, h);
}

// Workgroup b runs on XCD b mod 8 (round-robin dispatch over the 8 XCDs of
// an MI355X).  Renumbered so that each XCD receives one contiguous range of
// logical workgroups: neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g) {
    const uint32_t x = b & 7u, q = g >> 3, r = g & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

template <typename R>
__device__ __forceinline__ void torque(const FieldArgs &f, const R m[3], const R h[3], R k[3]) {
    const R p0 = m[1] * h[2] - m[2] * h[1];
    const R p1 = m[2] * h[0] - m[0] * h[2];
    const R p2 = m[0] * h[1] - m[1] * h[0];
    const R d0 = m[1] * p2 - m[2] * p1;
    const R d1 = m[2] * p0 - m[0] * p2;
    const R d2 = m[0] * p1 - m[1] * p0;
    const R a = R(f.alpha), g = R(-f.gamma_bar);
    const R pe = f.precession ? R(1) : R(0);
    k[0] = (pe * p0 + d0 * a) * g;
    k[1] = (pe * p1 + d1 * a) * g;
    k[2] = (pe * p2 + d2 * a) * g;
}

__device__ __forceinline__ void neu_add(double &s, double &c, double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
}

enum { STAGE_PLAIN = 0, STAGE_STATS = 1, STAGE_ERROR = 2 };

// Combines two compensated partial sums (s, c) += (s2, c2).
__device__ __forceinline__ void neu_merge(double &s, double &c, double s2, double c2) {
    neu_add(s, c, s2);
    c += c2;
}

// Workgroup reduction of the per-lane accumulators: wave shuffles, then the
// wavefront partials through LDS; lane 0 writes the result to `out`.
template <int MODE>
__device__ __forceinline__ void block_reduce(double *acc, double *out) {
    constexpr int NQ = MODE == STAGE_STATS ? kNStat : 1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        if (MODE == STAGE_STATS) {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const double s2 = __shfl_down(acc[2 * e], off), c2 = __shfl_down(acc[2 * e + 1], off);
                neu_merge(acc[2 * e], acc[2 * e + 1], s2, c2);
            }
#pragma unroll
            for (int q = 6; q < 9; ++q) acc[q] = fmax(acc[q], __shfl_down(acc[q], off));
        } else {
            acc[0] = fmax(acc[0], __shfl_down(acc[0], off));
        }
    }
    __shared__ double red[kBlock / 64][NQ];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int q = 0; q < NQ; ++q) red[wave][q] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
            if (MODE == STAGE_STATS) {
                for (int e = 0; e < 3; ++e) neu_merge(red[0][2 * e], red[0][2 * e + 1], red[w][2 * e], red[w][2 * e + 1]);
                for (int q = 6; q < 9; ++q) red[0][q] = fmax(red[0][q], red[w][q]);
            } else {
                red[0][0] = fmax(red[0][0], red[w][0]);
            }
        }
        for (int q = 0; q < NQ; ++q) out[q] = red[0][q];
    }
}

template <typename State, typename Real, typename Src, typename Out>
struct StageArgs {
    FieldArgs f;
    const Src *src;
    const Real *hd;    // demag field at src, or null
    const State *m0;   // base point of the update (MODE_PLAIN with out)
    const Real *kp[6];
    double ap[6];
    int nprev;
    double anew, dt;
    Real *kout;
    Out *out;
    double *partials;
};

// Per-cell body of the stage kernel; cc, vzm, vzp: src at i and at its z
// neighbours (see exchange_col).  TIER: the exchange is the sub-cell boundary
// stencil of f.tier_ex (a separate instantiation, so that the plain kernels
// keep their register budget).
template <typename State, typename Real, typename Src, typename Out, int MODE, bool TIER = false>
__device__ __forceinline__ void stage_cell(const StageArgs<State, Real, Src, Out> &a, int64_t i, int x, int y, int z,
                                           const Src cc[3], const Src vzm[3], const Src vzp[3], double *acc) {
    const FieldArgs &f = a.f;
    const int64_t n = f.n;
    const bool act = !f.mask || f.mask[i];
    Real m[3], hex[3] = {Real(0), Real(0), Real(0)}, hdv[3] = {Real(0), Real(0), Real(0)}, h[3], k[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = Real(cc[c]);
    if (act) {
        if (f.exchange) {
            Src he[3];
            if (TIER) {
                exchange_tier<Src, Src, Real>(f, a.src, i, x, y, z, cc, vzm, vzp, he);
            } else {
                exchange_col<Src, Src>(f, a.src, i, x, y, z, cc, vzm, vzp, he);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) hex[c] = Real(he[c]);
        }
        if (a.hd) {
#pragma unroll
            for (int c = 0; c < 3; ++c) hdv[c] = a.hd[c * n + i];
        }
        const Real he0 = f.has_ext ? Real(f.hx) : Real(0), he1 = f.has_ext ? Real(f.hy) : Real(0),
                   he2 = f.has_ext ? Real(f.hz) : Real(0);
        h[0] = (hex[0] + hdv[0]) + he0;
        h[1] = (hex[1] + hdv[1]) + he1;
        h[2] = (hex[2] + hdv[2]) + he2;
        torque(f, m, h, k);
    } else {
        h[0] = h[1] = h[2] = Real(0);
        k[0] = k[1] = k[2] = Real(0);
    }
    if (a.kout) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.kout[c * n + i] = k[c];
    }
    if (MODE != STAGE_ERROR && a.out) {
        State v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            State s = State(a.anew) * State(k[c]);
            for (int j = 0; j < a.nprev; ++j) s += State(a.ap[j]) * State(a.kp[j][c * n + i]);
            v[c] = a.m0[c * n + i] + State(a.dt) * s;
        }
        const State nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const State inv = nrm > State(0) ? State(1) / nrm : State(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) a.out[c * n + i] = Out(v[c] * inv);
    }
    if (MODE == STAGE_STATS && act) {
        // m . H per term in FP64 with the state's own m
        double md[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) md[c] = double(cc[c]);
        const double w = f.ms_cell ? f.ms_cell[i] : 1.0;
        // with a boundary tier the exchange energy stays that of the standard
        // masked stencil (native/backends/fdm/src/reductions_fp64.cu:123-200
        // sums the pair differences without face links or distances)
        // a lane's column (at most one z chunk) is summed plainly; lanes and
        // workgroups merge with compensation (block_reduce, then the host)
        if (TIER && f.exchange && !f.region) {
            // the pair form of the standard masked stencil's energy
            // (native/backends/fdm/src/reductions_fp64.cu:150-190): over the
            // cell's +x, +y, +z pairs with an active neighbour,
            // sum_i w m_i . H_i = -(2A / mu0)(w / Ms) sum |m_j - m_i|^2 / d^2
            const int64_t nxy = (int64_t)f.nx * f.ny;
            const bool up[3] = {x < f.nx - 1, y < f.ny - 1, z < f.nz - 1};
            const int64_t nb[3] = {i + 1, i + f.nx, i + nxy};
            const double id2[3] = {f.idx2, f.idy2, f.idz2};
            double e = 0.0;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int64_t j = up[t] ? nb[t] : i;
                const bool use = up[t] && (!f.mask || f.mask[j]);
                double d2 = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double dv = double(t == 2 ? vzp[c] : a.src[c * n + j]) - md[c];
                    d2 += dv * dv;
                }
                e += use ? d2 * id2[t] : 0.0;
            }
            acc[0] -= (f.ms_cell ? f.two_over_mu0 * f.a_uniform : f.pref) * e;
        } else {
            double he[3] = {double(hex[0]), double(hex[1]), double(hex[2])};
            if (TIER && f.exchange) {
                Src hp[3];
                exchange_col<Src, Src>(f, a.src, i, x, y, z, cc, vzm, vzp, hp);
#pragma unroll
                for (int c = 0; c < 3; ++c) he[c] = double(Real(hp[c]));
            }
            acc[0] += w * (md[0] * he[0] + md[1] * he[1] + md[2] * he[2]);
        }
        acc[2] += w * (md[0] * double(hdv[0]) + md[1] * double(hdv[1]) + md[2] * double(hdv[2]));
        if (f.has_ext) acc[4] += w * (md[0] * f.hx + md[1] * f.hy + md[2] * f.hz);
        const double h2 = double(h[0]) * h[0] + double(h[1]) * h[1] + double(h[2]) * h[2];
        const double d2 = double(hdv[0]) * hdv[0] + double(hdv[1]) * hdv[1] + double(hdv[2]) * hdv[2];
        const double k2 = double(k[0]) * k[0] + double(k[1]) * k[1] + double(k[2]) * k[2];
        acc[6] = fmax(acc[6], h2);
        acc[7] = fmax(acc[7], d2);
        acc[8] = fmax(acc[8], k2);
    }
    if (MODE == STAGE_ERROR) {
        // lib.rs:1653-1674: |dt sum_j E_j k_j| in FP64
        double e[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s = a.anew * double(k[c]);
            for (int j = 0; j < a.nprev; ++j) s += a.ap[j] * double(a.kp[j][c * n + i]);
            e[c] = s * a.dt;
        }
        acc[0] = fmax(acc[0], e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    }
}

// One workgroup per (tile of TX x TY columns, chunk of zc planes); each lane
// walks its column upward keeping src at z-1, z, z+1 in registers, so every
// plane of src is read from memory once per chunk (plus one halo plane on
// each side) and the x/y neighbours come from the same or an adjacent tile
// on the same XCD.
template <typename State, typename Real, typename Src, typename Out, int MODE, bool TIER = false>
__global__ void __launch_bounds__(kBlock) stage_kernel(StageArgs<State, Real, Src, Out> a) {
    const FieldArgs &f = a.f;
    const int64_t n = f.n;
    double acc[kNStat];
#pragma unroll
    for (int q = 0; q < kNStat; ++q) acc[q] = 0.0;
    const uint32_t tile = f.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int chunk = (int)(tile % (uint32_t)f.nchunks);
    const uint32_t t2 = tile / (uint32_t)f.nchunks;
    const int txi = (int)(t2 % (uint32_t)f.tiles_x), tyi = (int)(t2 / (uint32_t)f.tiles_x);
    const int TX = 1 << f.tx_log2;
    const int x = txi * TX + (int)(threadIdx.x & (TX - 1));
    const int y = tyi * (kBlock >> f.tx_log2) + (int)(threadIdx.x >> f.tx_log2);
    const int z0 = chunk * f.zc, z1 = min(f.nz, z0 + f.zc);
    if (x < f.nx && y < f.ny && z0 < z1) {
        const int64_t nxy = (int64_t)f.nx * f.ny;
        int64_t i = ((int64_t)z0 * f.ny + y) * f.nx + x;
        Src prev[3], cur[3], nxt[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) cur[c] = a.src[c * n + i];
#pragma unroll
        for (int c = 0; c < 3; ++c) prev[c] = z0 > 0 ? a.src[c * n + i - nxy] : cur[c];
        for (int z = z0; z < z1; ++z, i += nxy) {
#pragma unroll
            for (int c = 0; c < 3; ++c) nxt[c] = z + 1 < f.nz ? a.src[c * n + i + nxy] : cur[c];
            stage_cell<State, Real, Src, Out, MODE, TIER>(a, i, x, y, z, cur, prev, nxt, acc);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                prev[c] = cur[c];
                cur[c] = nxt[c];
            }
        }
    }
    if (MODE == STAGE_PLAIN) return;
    block_reduce<MODE>(acc, a.partials + (int64_t)blockIdx.x * (MODE == STAGE_STATS ? kNStat : 1));
}

// Uniform-material form of stage_kernel: no region table, no per-cell Ms,
// no boundary tier, demag and exchange on, at most one earlier slope (the
// Heun stages and the stats pass; RK4 / RK23 / DP45 / ABM3 stages with more
// slopes use stage_kernel).  The per-cell arithmetic is stage_cell's, in the
// same order, so results are bit-identical; what changes is that every load
// of a plane (the z neighbour ahead, the four x/y neighbours, H_d, the base
// point, the earlier slope, and with MASK their mask bytes) is issued in one
// branch-free batch before any arithmetic, where stage_kernel's
// optional-input branches split a plane's loads into several serial waits.
// M0_IS_SRC: the base point is src itself (predictor stages and the stats
// pass), so it is not loaded again.  MASK: an active mask is present; an
// inactive neighbour enters as the centre value (exchange_col) and an
// inactive cell gets zero field and slope and adds nothing to the sums.
// TIER: the exchange is the sub-cell boundary stencil of f.tier_ex
// (exchange_tier on the plane's loaded neighbours, face weights, volume
// fraction and per-cell Ms, which it allows; the mask then gates only the
// cell itself), and the stats pass sums the standard masked stencil's
// exchange energy in stage_cell's pair form from the same loaded values.
template <typename State, typename Real, typename Src, typename Out, int MODE, int NPREV, bool M0_IS_SRC,
          bool HAS_OUT, bool MASK = false, bool TIER = false>
__device__ __forceinline__ void stage_uniform_body(const StageArgs<State, Real, Src, Out> &a) {
    static_assert(MODE != STAGE_ERROR && NPREV <= 1, "plain and stats stages with at most one earlier slope");
    static_assert(!M0_IS_SRC || sizeof(Src) == sizeof(State), "a base point read from src is a state");
    const FieldArgs &f = a.f;
    const int64_t n = f.n;
    double acc[kNStat];
#pragma unroll
    for (int q = 0; q < kNStat; ++q) acc[q] = 0.0;
    const uint32_t tile = f.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int chunk = (int)(tile % (uint32_t)f.nchunks);
    const uint32_t t2 = tile / (uint32_t)f.nchunks;
    const int txi = (int)(t2 % (uint32_t)f.tiles_x), tyi = (int)(t2 / (uint32_t)f.tiles_x);
    const int TX = 1 << f.tx_log2;
    const int x = txi * TX + (int)(threadIdx.x & (TX - 1));
    const int y = tyi * (kBlock >> f.tx_log2) + (int)(threadIdx.x >> f.tx_log2);
    const int z0 = chunk * f.zc, z1 = min(f.nz, z0 + f.zc);
    if (x < f.nx && y < f.ny && z0 < z1) {
        const int64_t nxy = (int64_t)f.nx * f.ny;
        // clamped (Neumann) neighbours as offsets: 0 at the grid edge
        const int64_t oxm = x > 0 ? -1 : 0, oxp = x < f.nx - 1 ? 1 : 0;
        const int64_t oym = y > 0 ? -(int64_t)f.nx : 0, oyp = y < f.ny - 1 ? (int64_t)f.nx : 0;
        const Src *__restrict__ src = a.src;
        const Real *__restrict__ hd = a.hd;
        const State *__restrict__ m0p = a.m0;
        const Real *__restrict__ kp = a.kp[0];
        int64_t i = ((int64_t)z0 * f.ny + y) * f.nx + x;
        Src prev[3], cur[3];
        const int64_t ib = z0 > 0 ? i - nxy : i;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cur[c] = src[c * n + i];
            prev[c] = src[c * n + ib];
        }
        // mask bytes marched along z like src (unused without MASK)
        const uint8_t *__restrict__ mk = f.mask;
        uint8_t kcur = 1, kprev = 1;
        if constexpr (MASK) {
            kcur = mk[i];
            kprev = mk[ib];
        }
        const Src ix = Src(f.idx2), iy = Src(f.idy2), iz = Src(f.idz2), pref = Src(f.pref);
        const Real he[3] = {f.has_ext ? Real(f.hx) : Real(0), f.has_ext ? Real(f.hy) : Real(0),
                            f.has_ext ? Real(f.hz) : Real(0)};
        for (int z = z0; z < z1; ++z, i += nxy) {
            const int64_t ia = z + 1 < f.nz ? i + nxy : i;
            Src nxt[3], xm[3], xp[3], ym[3], yp[3];
            Real hdv[3], kpv[3];
            State m0v[3];
            uint8_t knxt = 1, kxm = 1, kxp = 1, kym = 1, kyp = 1;
            if constexpr (MASK) {
                knxt = mk[ia];
                kxm = mk[i + oxm];
                kxp = mk[i + oxp];
                kym = mk[i + oym];
                kyp = mk[i + oyp];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const Src *sc = src + c * n;
                nxt[c] = sc[ia];
                xm[c] = sc[i + oxm];
                xp[c] = sc[i + oxp];
                ym[c] = sc[i + oym];
                yp[c] = sc[i + oyp];
                hdv[c] = hd[c * n + i];
                if (HAS_OUT && !M0_IS_SRC) m0v[c] = m0p[c * n + i];
                if (HAS_OUT && NPREV == 1) kpv[c] = kp[c * n + i];
            }
            double wt[6], phi0 = 0.0, msc = 1.0;
            if constexpr (TIER) {
#pragma unroll
                for (int s = 0; s < 6; ++s) wt[s] = tier_weight<Real>(f, s, i);
                phi0 = tier_phi<Real>(f, i);
                msc = *(f.ms_cell ? f.ms_cell + i : &kUnitStandIn);
            }
            const bool act = !MASK || kcur != 0;
            Src het[3];
            if constexpr (TIER) {
                const bool in[6] = {x > 0, x < f.nx - 1, y > 0, y < f.ny - 1, f.nz > 1 && z > 0,
                                    f.nz > 1 && z < f.nz - 1};
                Src v[6][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    v[0][c] = xm[c];
                    v[1][c] = xp[c];
                    v[2][c] = ym[c];
                    v[3][c] = yp[c];
                    v[4][c] = prev[c];
                    v[5][c] = nxt[c];
                }
                const double pre = f.ms_cell ? 2.0 * f.a_tier / (kMu0 * msc) : f.tier_pref;
                exchange_tier_vals<Src, Src>(f, in, v, wt, phi0, pre, cur, het);
            }
            Real m[3], hex[3], h[3], k[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                m[c] = Real(cur[c]);
                if constexpr (TIER) {
                    hex[c] = Real(het[c]);
                } else {
                    // exchange_col's arithmetic (z neighbours clamped at the
                    // grid edge, inactive neighbours replaced by the centre)
                    const Src cc = cur[c];
                    const Src zmv = (z > 0 && kprev) ? prev[c] : cc, zpv = (z < f.nz - 1 && knxt) ? nxt[c] : cc;
                    const Src xmv = kxm ? xm[c] : cc, xpv = kxp ? xp[c] : cc;
                    const Src ymv = kym ? ym[c] : cc, ypv = kyp ? yp[c] : cc;
                    const Src lap = (xpv - Src(2) * cc + xmv) * ix + (ypv - Src(2) * cc + ymv) * iy +
                                    (zpv - Src(2) * cc + zmv) * iz;
                    hex[c] = Real(pref * lap);
                }
                h[c] = (hex[c] + hdv[c]) + he[c];
            }
            torque(f, m, h, k);
            if (MASK && !act) {
#pragma unroll
                for (int c = 0; c < 3; ++c) hex[c] = hdv[c] = h[c] = k[c] = Real(0);
            }
            if (a.kout) {
#pragma unroll
                for (int c = 0; c < 3; ++c) a.kout[c * n + i] = k[c];
            }
            if constexpr (HAS_OUT) {
                State v[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    State s = State(a.anew) * State(k[c]);
                    if constexpr (NPREV == 1) s += State(a.ap[0]) * State(kpv[c]);
                    const State base = M0_IS_SRC ? State(cur[c]) : m0v[c];
                    v[c] = base + State(a.dt) * s;
                }
                const State nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                const State inv = nrm > State(0) ? State(1) / nrm : State(0);
#pragma unroll
                for (int c = 0; c < 3; ++c) a.out[c * n + i] = Out(v[c] * inv);
            }
            if (MODE == STAGE_STATS && act) {
                double md[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) md[c] = double(cur[c]);
                const double w = (TIER && f.ms_cell) ? msc : 1.0;
                if constexpr (TIER) {
                    // stage_cell's pair form over the cell's +x, +y, +z pairs
                    const bool up[3] = {x < f.nx - 1, y < f.ny - 1, z < f.nz - 1};
                    const uint8_t kn[3] = {kxp, kyp, knxt};
                    const double id2[3] = {f.idx2, f.idy2, f.idz2};
                    double e = 0.0;
#pragma unroll
                    for (int t = 0; t < 3; ++t) {
                        const bool use = up[t] && kn[t] != 0;
                        double d2 = 0.0;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const double dv = double(t == 0 ? xp[c] : t == 1 ? yp[c] : nxt[c]) - md[c];
                            d2 += dv * dv;
                        }
                        e += use ? d2 * id2[t] : 0.0;
                    }
                    acc[0] -= (f.ms_cell ? f.two_over_mu0 * f.a_uniform : f.pref) * e;
                } else {
                    acc[0] += w * (md[0] * double(hex[0]) + md[1] * double(hex[1]) + md[2] * double(hex[2]));
                }
                acc[2] += w * (md[0] * double(hdv[0]) + md[1] * double(hdv[1]) + md[2] * double(hdv[2]));
                if (f.has_ext) acc[4] += w * (md[0] * f.hx + md[1] * f.hy + md[2] * f.hz);
                const double h2 = double(h[0]) * h[0] + double(h[1]) * h[1] + double(h[2]) * h[2];
                const double d2 = double(hdv[0]) * hdv[0] + double(hdv[1]) * hdv[1] + double(hdv[2]) * hdv[2];
                const double k2 = double(k[0]) * k[0] + double(k[1]) * k[1] + double(k[2]) * k[2];
                acc[6] = fmax(acc[6], h2);
                acc[7] = fmax(acc[7], d2);
                acc[8] = fmax(acc[8], k2);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                prev[c] = cur[c];
                cur[c] = nxt[c];
            }
            kprev = kcur;
            kcur = knxt;
        }
    }
    if constexpr (MODE == STAGE_PLAIN) return;
    block_reduce<MODE>(acc, a.partials + (int64_t)blockIdx.x * kNStat);
}

template <typename State, typename Real, typename Src, typename Out, int MODE, int NPREV, bool M0_IS_SRC,
          bool HAS_OUT, bool MASK = false, bool TIER = false>
__global__ void __launch_bounds__(kBlock) stage_uniform_kernel(StageArgs<State, Real, Src, Out> a) {
    stage_uniform_body<State, Real, Src, Out, MODE, NPREV, M0_IS_SRC, HAS_OUT, MASK, TIER>(a);
}
// Multi-instance form: instance blockIdx.z / zdiv of the table.
template <typename State, typename Real, typename Src, typename Out, int MODE, int NPREV, bool M0_IS_SRC,
          bool HAS_OUT, bool MASK = false, bool TIER = false>
__global__ void __launch_bounds__(kBlock) stage_uniform_multi(const StageArgs<State, Real, Src, Out> *tab, int zdiv) {
    const StageArgs<State, Real, Src, Out> a = tab[blockIdx.z / zdiv];
    stage_uniform_body<State, Real, Src, Out, MODE, NPREV, M0_IS_SRC, HAS_OUT, MASK, TIER>(a);
}

// Whether a stage with these arguments may run as stage_uniform_kernel
// (with MASK = f.mask != nullptr and TIER = f.tier_ex != 0; per-cell Ms only
// with a tier).
inline bool stage_uniform_ok(const FieldArgs &f, const void *hd, int nprev) {
    return !f.region && (f.tier_ex || !f.ms_cell) && hd && f.exchange && nprev <= 1;
}

// Column tiling of the stage kernels: tile width TX = 64 (or the smallest
// power of two >= nx), TY = kBlock / TX rows; z is split into chunks only
// when the tiles alone give fewer than about 1024 workgroups (longer chunks
// re-read fewer halo planes: measured faster at 512x512x8).
void stage_tiling(const Context &c, int &tx_log2, int &tiles_x, int &tiles_y, int &zc, int &nchunks) {
    tx_log2 = 0;
    while ((1u << tx_log2) < c.nx && tx_log2 < 6) ++tx_log2;
    const int tx = 1 << tx_log2, ty = kBlock / tx;
    tiles_x = (int)cdiv(c.nx, tx);
    tiles_y = (int)cdiv(c.ny, ty);
    const uint64_t tiles = (uint64_t)tiles_x * tiles_y;
    const int want = (int)std::min<uint64_t>(c.nz, std::max<uint64_t>(1, cdiv(1024, tiles)));
    zc = c.dp.tune.zc > 0 ? std::min<int>(c.dp.tune.zc, (int)c.nz) : (int)cdiv(c.nz, want);
    nchunks = (int)cdiv(c.nz, zc);
}

unsigned stage_blocks(const Context &c) {
    int tl, tx, ty, zc, nch;
    stage_tiling(c, tl, tx, ty, zc, nch);
    return (unsigned)((uint64_t)tx * ty * nch);
}

FieldArgs field_args(const Context &c) {
    FieldArgs f{};
    f.nx = (int)c.nx;
    f.ny = (int)c.ny;
    f.nz = (int)c.nz;
    f.n = (int64_t)c.n;
    f.div_nx = FastDiv(c.nx);
    f.div_nxy = FastDiv(c.nx * c.ny);
    stage_tiling(c, f.tx_log2, f.tiles_x, f.tiles_y, f.zc, f.nchunks);
    f.xcd = c.dp.tune.xcd;
    f.idx2 = 1.0 / (c.dx * c.dx);
    f.idy2 = 1.0 / (c.dy * c.dy);
    f.idz2 = 1.0 / (c.dz * c.dz);
    f.a_uniform = c.use_lut ? c.a : c.a_exchange;
    f.pref = 2.0 * f.a_uniform / (kMu0 * c.ms);
    f.two_over_mu0 = 2.0 / kMu0;
    f.pre_lut = 2.0 / (kMu0 * c.ms);
    f.mask = c.d_mask;
    f.region = c.use_lut ? c.d_region : nullptr;
    f.lut = c.use_lut ? c.d_lut : nullptr;
    f.ms_cell = c.d_ms_cell;
    f.exchange = c.exchange ? 1 : 0;
    f.has_ext = c.has_ext ? 1 : 0;
    f.precession = c.precession ? 1 : 0;
    f.hx = c.hext[0];
    f.hy = c.hext[1];
    f.hz = c.hext[2];
    f.gamma_bar = c.gamma / (1.0 + c.alpha * c.alpha);
    f.alpha = c.alpha;
    f.tier_ex = c.tier_ex;
    f.phi = c.d_phi;
    f.phi32 = c.d_phi32;
    f.tier_w = c.d_tier_w;
    f.tier_w32 = c.d_tier_w32;
    f.phi_floor = c.phi_floor;
    f.delta_min = c.delta_min;
    f.tier_pref = 2.0 * c.a / (kMu0 * c.ms);
    f.a_tier = c.a;
    f.cell[0] = c.dx;
    f.cell[1] = c.dy;
    f.cell[2] = c.dz;
    return f;
}

struct Timed {
    Context &c;
    int entry = -1;
    hipEvent_t ev{};
    Timed(Context &cc, const char *name) : c(cc) {
        if (c.timing) {
            entry = timing_entry(c, name);
            timing_begin(c, entry, &ev);
        }
    }
    ~Timed() {
        if (entry >= 0) timing_end(c, entry, ev);
    }
};

}  // namespace
}  // namespace fm
