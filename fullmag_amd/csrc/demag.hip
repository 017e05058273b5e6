// demag.hip -- zero-padded Newell-tensor demag convolution on gfx950.
//
// Replaces the cuFFT pipeline of native/backends/fdm/src/demag_fp64.cu:264-381
// (pack, 3 forward Z2Z transforms over the fully padded grid, 6-component
// complex tensor multiply, 3 inverse transforms, unpack).  The mathematics is
// the same as crates/fullmag-engine/src/lib.rs:2306-2400:
//     H = IFFT( -K . FFT(Ms m) ) / Npad   on a (2nx, 2ny, 2nz) grid
// but the transform is organised around what is actually non-zero:
//   x  rows   real-to-complex: two real rows share one complex transform and
//             only nx of the 2nx inputs are loaded; nx+1 outputs are kept.
//   y  cols   forward transform reads only ny of the 2ny rows.
//   z  cols   forward transform of nz inputs, 3x3 tensor multiply, inverse
//             transform, only nz outputs written (3-D); for thin films the
//             tensor multiply sits between the forward and inverse y
//             transforms of one kernel and no z transform exists.
//   y  cols   inverse transform writes only ny rows (3-D).
//   x  rows   complex-to-real: two rows from one complex transform, only nx
//             outputs, scaled by 1/Npad and masked.
// The Newell spectra are real (the padded kernel is even or odd along every
// axis), so the tensor is stored as six real values on the non-negative
// octant (kx <= nx, ky <= ny, kz <= nz) and mirrored with parity signs.
#include "context.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace fm {
namespace {

constexpr int kPad = 1;         // LDS line padding (complex slots)
constexpr int kMaxBlock = 512;  // lanes per workgroup for FFT kernels

// ---------------------------------------------------------------------------
// Tensor access
// ---------------------------------------------------------------------------
template <typename Real>
struct KernelView {
    const Real *k = nullptr;  // octant tensor, nullptr for the projection fallback
    int kcomp = 6, kz_n = 1;
    int ny = 0, nz = 0, py = 0, pz = 1, hx = 0;
    double kxs = 0, kys = 0, kzs = 0;  // 2 pi / (p d) per axis (projection)
};

// t = (xx, yy, zz, xy, xz, yz) at frequency (kx, ky, kz); kx <= nx.
template <typename Real>
__device__ inline void tensor_at(const KernelView<Real> &K, int kx, int ky, int kz, Real t[6]) {
    if (K.k) {
        const bool ylo = ky <= K.ny;
        const int oy = ylo ? ky : K.py - ky;
        const Real sy = ylo ? Real(1) : Real(-1);
        int oz = 0;
        Real sz = Real(1);
        if (K.kz_n > 1) {
            const bool zlo = kz <= K.nz;
            oz = zlo ? kz : K.pz - kz;
            sz = zlo ? Real(1) : Real(-1);
        }
        const int64_t plane = (int64_t)(K.ny + 1) * K.hx;
        const int64_t off = (int64_t)oz * plane + (int64_t)oy * K.hx + kx;
        const int64_t cs = (int64_t)K.kz_n * plane;
        t[0] = K.k[off];
        t[1] = K.k[cs + off];
        t[2] = K.k[2 * cs + off];
        t[3] = sy * K.k[3 * cs + off];
        if (K.kcomp == 6) {
            t[4] = sz * K.k[4 * cs + off];
            t[5] = sy * sz * K.k[5 * cs + off];
        } else {
            t[4] = Real(0);
            t[5] = Real(0);
        }
    } else {
        // native/backends/fdm/src/demag_fp64.cu:102-148: H(k) = -k (k.M) / k^2
        const int fy = ky <= K.py / 2 ? ky : ky - K.py;
        const int fz = kz <= K.pz / 2 ? kz : kz - K.pz;
        const double qx = kx * K.kxs, qy = fy * K.kys, qz = fz * K.kzs;
        const double q2 = qx * qx + qy * qy + qz * qz;
        if (q2 == 0.0) {
            for (int i = 0; i < 6; ++i) t[i] = Real(0);
            return;
        }
        const double r = 1.0 / q2;
        t[0] = Real(qx * qx * r); t[1] = Real(qy * qy * r); t[2] = Real(qz * qz * r);
        t[3] = Real(qx * qy * r); t[4] = Real(qx * qz * r); t[5] = Real(qy * qz * r);
    }
}

// ---------------------------------------------------------------------------
// Row (x) passes
// ---------------------------------------------------------------------------
template <typename Real>
struct XArgs {
    const cplx<Real> *tw;
    LinePlan lp;
    int64_t lines;  // real lines
    int lpw;        // complex lines (= pairs of real lines) per workgroup
    int n_in;       // loaded inputs per real line (forward)
    int n_out;      // stored outputs per real line (inverse)
    int hx;         // half-spectrum length kept = px/2 + 1
    Real scale;     // inverse: 1/Npad
};

// Rows of the half spectrum: line l -> (component, plane, row).
template <typename Real>
struct SpecRows {
    cplx<Real> *base;
    int64_t lines_per_comp, comp_stride;
    int lines_per_plane, rows, hx;
    __device__ cplx<Real> *row(int64_t l) const {
        const int64_t c = l / lines_per_comp, r = l - c * lines_per_comp;
        const int64_t z = r / lines_per_plane, y = r - z * lines_per_plane;
        return base + c * comp_stride + (z * rows + y) * hx;
    }
};

// Ms m on line l = (c, z, y); zero on inactive cells.
template <typename Real, typename S>
struct MagSource {
    const S *m;
    const uint8_t *mask;
    const double *ms_cell;
    double ms;
    int nx;
    int64_t per_comp, n;
    __device__ Real operator()(int64_t l, int x) const {
        const int64_t c = l / per_comp;
        const int64_t i = (l - c * per_comp) * nx + x;
        if (mask && !mask[i]) return Real(0);
        const double w = ms_cell ? ms_cell[i] : ms;
        return Real(double(m[c * n + i]) * w);
    }
};

// Padded real-space Newell kernel on line l = (c, z, y) of the (px, py, pz)
// grid, mirrored from the octant with the placement signs of
// crates/fullmag-fdm-demag/src/newell.rs:397-426 (index n of each padded
// axis is never written there and stays zero).
struct NewellSource {
    const double *oct;  // [6][nz][ny][nx]
    int nx, ny, nz, px, py, pz;
    __device__ double operator()(int64_t l, int x) const {
        const int64_t per_comp = (int64_t)py * pz;
        const int c = (int)(l / per_comp);
        const int64_t r = l - (int64_t)c * per_comp;
        const int zz = (int)(r / py), yy = (int)(r - (int64_t)zz * py);
        if (x == nx || yy == ny || (pz > 1 && zz == nz)) return 0.0;
        const int xi = x < nx ? x : px - x;
        const int yi = yy < ny ? yy : py - yy;
        const int zi = (pz == 1 || zz < nz) ? zz : pz - zz;
        const double sx = x < nx ? 1.0 : -1.0, sy = yy < ny ? 1.0 : -1.0;
        const double sz = (pz == 1 || zz < nz) ? 1.0 : -1.0;
        double s = 1.0;
        if (c == 3) s = sx * sy;
        else if (c == 4) s = sx * sz;
        else if (c == 5) s = sy * sz;
        return s * oct[(((int64_t)c * nz + zi) * ny + yi) * nx + xi];
    }
};

template <typename Real>
struct FieldSink {
    Real *out;
    const uint8_t *mask;
    int nx;
    int64_t per_comp, n;
    __device__ void operator()(int64_t l, int x, Real v) const {
        const int64_t c = l / per_comp;
        const int64_t i = (l - c * per_comp) * nx + x;
        out[c * n + i] = (mask && !mask[i]) ? Real(0) : v;
    }
};

// Forward real-to-complex along x.  Real lines 2p and 2p+1 are packed as
// z = a + i b; with Z = FFT(z):  A_k = (Z_k + conj Z_{N-k}) / 2,
// B_k = (Z_k - conj Z_{N-k}) / (2i), for k = 0..N/2.
template <typename Real, class Src>
__global__ void __launch_bounds__(kMaxBlock) xfwd_kernel(XArgs<Real> a, Src src, SpecRows<Real> dst) {
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int n = a.lp.n, ls = n + kPad, T = a.lp.threads;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int64_t pairs = (a.lines + 1) / 2;
    const int64_t p0 = (int64_t)blockIdx.x * a.lpw;
    for (int idx = tid; idx < a.lpw * n; idx += nthr) {
        const int l = idx / n, x = idx - l * n;
        const int64_t p = p0 + l;
        cplx<Real> v = {Real(0), Real(0)};
        if (x < a.n_in && p < pairs) {
            v.x = src(2 * p, x);
            if (2 * p + 1 < a.lines) v.y = src(2 * p + 1, x);
        }
        lds[l * ls + x] = v;
    }
    __syncthreads();
    line_fft(lds + (tid / T) * ls, a.lp, tid % T, a.tw, false);
    for (int idx = tid; idx < a.lpw * a.hx; idx += nthr) {
        const int l = idx / a.hx, k = idx - l * a.hx;
        const int64_t p = p0 + l;
        if (p >= pairs) continue;
        const cplx<Real> zk = lds[l * ls + k];
        const cplx<Real> zn = lds[l * ls + (k == 0 ? 0 : n - k)];
        const Real h = Real(0.5);
        dst.row(2 * p)[k] = {h * (zk.x + zn.x), h * (zk.y - zn.y)};
        if (2 * p + 1 < a.lines) dst.row(2 * p + 1)[k] = {h * (zk.y + zn.y), h * (zn.x - zk.x)};
    }
}

// Inverse complex-to-real along x: Z_k = A_k + i B_k with the Hermitian
// continuation A_{N-k} = conj A_k; z = IFFT(Z); a = Re z, b = Im z.
template <typename Real, class Sink>
__global__ void __launch_bounds__(kMaxBlock) xinv_kernel(XArgs<Real> a, SpecRows<Real> srcr, Sink sink) {
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int n = a.lp.n, ls = n + kPad, T = a.lp.threads;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int64_t pairs = (a.lines + 1) / 2;
    const int64_t p0 = (int64_t)blockIdx.x * a.lpw;
    for (int idx = tid; idx < a.lpw * n; idx += nthr) {
        const int l = idx / n, k = idx - l * n;
        const int64_t p = p0 + l;
        cplx<Real> z = {Real(0), Real(0)};
        if (p < pairs) {
            const bool lo = k < a.hx;
            const int kk = lo ? k : n - k;
            cplx<Real> A = srcr.row(2 * p)[kk];
            cplx<Real> B = {Real(0), Real(0)};
            if (2 * p + 1 < a.lines) B = srcr.row(2 * p + 1)[kk];
            if (!lo) {
                A.y = -A.y;
                B.y = -B.y;
            }
            z = {A.x - B.y, A.y + B.x};
        }
        lds[l * ls + k] = z;
    }
    __syncthreads();
    line_fft(lds + (tid / T) * ls, a.lp, tid % T, a.tw, true);
    for (int idx = tid; idx < a.lpw * a.n_out; idx += nthr) {
        const int l = idx / a.n_out, x = idx - l * a.n_out;
        const int64_t p = p0 + l;
        if (p >= pairs) continue;
        const cplx<Real> z = lds[l * ls + x];
        sink(2 * p, x, z.x * a.scale);
        if (2 * p + 1 < a.lines) sink(2 * p + 1, x, z.y * a.scale);
    }
}

// ---------------------------------------------------------------------------
// Column (y or z) passes, in place on the half spectrum
// ---------------------------------------------------------------------------
template <typename Real>
struct ColArgs {
    cplx<Real> *data;
    const cplx<Real> *tw;
    LinePlan lp;
    int n_in, n_out;
    int64_t estride;  // element stride along the transformed axis
    int ncols;        // columns along the contiguous axis (= half-spectrum width)
    int tile;         // columns per workgroup
    int q_div;        // blockIdx.y = q -> offset (q / q_div) s_hi + (q % q_div) s_lo
    int64_t s_hi, s_lo;
    int inverse;
    int64_t comp_stride;  // CONV: distance between components
    KernelView<Real> K;
};

// CONV = 0: plain transform; 2: forward y, tensor (ky = e, kz = 0), inverse y;
// 3: forward z, tensor (ky = blockIdx.y, kz = e), inverse z.
template <typename Real, int CONV>
__global__ void __launch_bounds__(kMaxBlock) col_kernel(ColArgs<Real> a) {
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    constexpr int NC = CONV ? 3 : 1;
    const int n = a.lp.n, ls = n + kPad, T = a.lp.threads;
    const int nslots = NC * a.tile;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int kx0 = blockIdx.x * a.tile;
    const int q = blockIdx.y;
    int64_t qoff;
    if (CONV) qoff = (int64_t)q * a.s_lo;
    else qoff = (int64_t)(q / a.q_div) * a.s_hi + (int64_t)(q % a.q_div) * a.s_lo;

    for (int idx = tid; idx < nslots * n; idx += nthr) {
        const int s = idx % nslots, e = idx / nslots;
        const int c = s / a.tile, j = s - c * a.tile;
        const int kx = kx0 + j;
        cplx<Real> v = {Real(0), Real(0)};
        if (e < a.n_in && kx < a.ncols) v = a.data[c * a.comp_stride + qoff + (int64_t)e * a.estride + kx];
        lds[s * ls + e] = v;
    }
    __syncthreads();
    line_fft(lds + (tid / T) * ls, a.lp, tid % T, a.tw, CONV ? false : (a.inverse != 0));
    if (CONV) {
        for (int idx = tid; idx < a.tile * n; idx += nthr) {
            const int j = idx % a.tile, e = idx / a.tile;
            const int kx = kx0 + j;
            if (kx >= a.ncols) continue;
            const int ky = CONV == 2 ? e : q;
            const int kz = CONV == 2 ? 0 : e;
            Real t[6];
            tensor_at(a.K, kx, ky, kz, t);
            cplx<Real> &mx = lds[j * ls + e];
            cplx<Real> &my = lds[(a.tile + j) * ls + e];
            cplx<Real> &mz = lds[(2 * a.tile + j) * ls + e];
            const cplx<Real> X = mx, Y = my, Z = mz;
            mx = {-(t[0] * X.x + t[3] * Y.x + t[4] * Z.x), -(t[0] * X.y + t[3] * Y.y + t[4] * Z.y)};
            my = {-(t[3] * X.x + t[1] * Y.x + t[5] * Z.x), -(t[3] * X.y + t[1] * Y.y + t[5] * Z.y)};
            mz = {-(t[4] * X.x + t[5] * Y.x + t[2] * Z.x), -(t[4] * X.y + t[5] * Y.y + t[2] * Z.y)};
        }
        __syncthreads();
        line_fft(lds + (tid / T) * ls, a.lp, tid % T, a.tw, true);
    }
    for (int idx = tid; idx < nslots * a.n_out; idx += nthr) {
        const int s = idx % nslots, e = idx / nslots;
        const int c = s / a.tile, j = s - c * a.tile;
        const int kx = kx0 + j;
        if (kx < a.ncols) a.data[c * a.comp_stride + qoff + (int64_t)e * a.estride + kx] = lds[s * ls + e];
    }
}

// Octant extraction after the setup transform: kern[c][oz][oy][kx] =
// Re spec[c][oz][oy][kx] for oy <= ny, oz <= nz (or oz = 0 in 2-D).
template <typename Real>
__global__ void extract_octant_kernel(const cplx<double> *spec, Real *kern, int kcomp, int kz_n, int ny, int hx,
                                      int py, int64_t spec_comp, int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t plane = (int64_t)(ny + 1) * hx;
    const int64_t per_comp = (int64_t)kz_n * plane;
    const int c = (int)(i / per_comp);
    const int64_t r = i - (int64_t)c * per_comp;
    const int oz = (int)(r / plane);
    const int64_t r2 = r - (int64_t)oz * plane;
    const int oy = (int)(r2 / hx), kx = (int)(r2 - (int64_t)oy * hx);
    (void)kcomp;  // 2-D keeps only the first four components
    kern[i] = Real(spec[c * spec_comp + ((int64_t)oz * py + oy) * hx + kx].x);
}

// ---------------------------------------------------------------------------
// Launch geometry
// ---------------------------------------------------------------------------
size_t lds_budget() {
    static size_t b = 0;
    if (b == 0) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t p;
        size_t lim = 65536;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.sharedMemPerBlock > 0)
            lim = p.sharedMemPerBlock;
        b = std::min<size_t>(lim, 98304);
    }
    return b;
}

int pick_tile(const LinePlan &lp, int ncomp, size_t elem, int ncols) {
    int tile = 64;
    const size_t budget = lds_budget();
    while (tile > 1 && ((size_t)ncomp * tile * (lp.n + kPad) * elem > budget ||
                        ncomp * tile * lp.threads > kMaxBlock))
        tile >>= 1;
    while (tile > 1 && tile / 2 >= ncols) tile >>= 1;
    return tile;
}

int pick_lpw(const LinePlan &lp, size_t elem) {
    int lpw = std::max(1, 256 / std::max(1, lp.threads));
    const size_t budget = lds_budget();
    while (lpw > 1 && ((size_t)lpw * (lp.n + kPad) * elem > budget || lpw * lp.threads > kMaxBlock)) --lpw;
    return lpw;
}

template <typename Real>
bool geometry_ok(const DemagPlan &dp, std::string &err) {
    const size_t e = sizeof(cplx<Real>);
    auto fits = [&](const LinePlan &lp, int ncomp, int tile) {
        return (size_t)ncomp * tile * (lp.n + kPad) * e <= lds_budget() && ncomp * tile * lp.threads <= kMaxBlock;
    };
    if (!fits(dp.lx, 1, dp.x_lines_per_wg)) { err = "x line FFT does not fit in LDS"; return false; }
    if (dp.three_d) {
        if (!fits(dp.ly, 1, dp.y_tile)) { err = "y line FFT does not fit in LDS"; return false; }
        if (!fits(dp.lz, 3, dp.z_tile)) { err = "z line FFT does not fit in LDS"; return false; }
    } else if (!fits(dp.ly, 3, dp.yconv_tile)) {
        err = "thin-film y convolution does not fit in LDS";
        return false;
    }
    return true;
}

template <typename Real>
XArgs<Real> xargs(const DemagPlan &dp, int64_t lines, int lpw, int n_in, int n_out, Real scale) {
    XArgs<Real> a;
    a.tw = static_cast<const cplx<Real> *>(dp.tw_x);
    a.lp = dp.lx;
    a.lines = lines;
    a.lpw = lpw;
    a.n_in = n_in;
    a.n_out = n_out;
    a.hx = dp.hx;
    a.scale = scale;
    return a;
}

template <typename Real>
KernelView<Real> kview(const Context &c) {
    const DemagPlan &dp = c.dp;
    KernelView<Real> K;
    K.k = dp.projection ? nullptr : static_cast<const Real *>(dp.kern);
    K.kcomp = dp.kcomp;
    K.kz_n = dp.kz_n;
    K.ny = dp.ny;
    K.nz = dp.nz;
    K.py = dp.py;
    K.pz = dp.pz;
    K.hx = dp.hx;
    K.kxs = 2.0 * kPi / (dp.px * c.dx);
    K.kys = 2.0 * kPi / (dp.py * c.dy);
    K.kzs = 2.0 * kPi / (dp.pz * c.dz);
    return K;
}

template <typename Real>
void launch_col(Context &c, const char *name, int mode, const ColArgs<Real> &a, dim3 grid) {
    const int ncomp = mode ? 3 : 1;
    const int threads = ncomp * a.tile * a.lp.threads;
    const size_t lds = (size_t)ncomp * a.tile * (a.lp.n + kPad) * sizeof(cplx<Real>);
    hipEvent_t ev{};
    if (c.timing) timing_begin(c, timing_entry(c, name), &ev);
    if (mode == 0) hipLaunchKernelGGL((col_kernel<Real, 0>), grid, dim3(threads), lds, c.stream, a);
    else if (mode == 2) hipLaunchKernelGGL((col_kernel<Real, 2>), grid, dim3(threads), lds, c.stream, a);
    else hipLaunchKernelGGL((col_kernel<Real, 3>), grid, dim3(threads), lds, c.stream, a);
    if (c.timing) timing_end(c, timing_entry(c, name), ev);
}

template <typename Real>
SpecRows<Real> spec_rows(cplx<Real> *base, int64_t lines_per_comp, int64_t comp_stride, int lines_per_plane, int rows,
                         int hx) {
    SpecRows<Real> r;
    r.base = base;
    r.lines_per_comp = lines_per_comp;
    r.comp_stride = comp_stride;
    r.lines_per_plane = lines_per_plane;
    r.rows = rows;
    r.hx = hx;
    return r;
}

template <typename Real, typename S>
void demag_eval_t(Context &c, const S *src, Real *out) {
    DemagPlan &dp = c.dp;
    cplx<Real> *spec = static_cast<cplx<Real> *>(dp.spec);
    const int64_t per_comp_lines = (int64_t)dp.ny * dp.nz;
    const int64_t lines = 3 * per_comp_lines;
    const int64_t comp_stride = (int64_t)dp.nz * dp.rows * dp.hx;
    SpecRows<Real> rows = spec_rows<Real>(spec, per_comp_lines, comp_stride, dp.ny, dp.rows, dp.hx);

    // x forward
    {
        MagSource<Real, S> ms{src, c.d_mask, c.d_ms_cell, c.ms, dp.nx, per_comp_lines, (int64_t)c.n};
        XArgs<Real> a = xargs<Real>(dp, lines, dp.x_lines_per_wg, dp.nx, 0, Real(1));
        const int64_t pairs = (lines + 1) / 2;
        dim3 grid((unsigned)cdiv(pairs, a.lpw));
        const size_t lds = (size_t)a.lpw * (dp.lx.n + kPad) * sizeof(cplx<Real>);
        hipEvent_t ev{};
        if (c.timing) timing_begin(c, timing_entry(c, "demag_x_r2c"), &ev);
        hipLaunchKernelGGL((xfwd_kernel<Real, MagSource<Real, S>>), grid, dim3(a.lpw * dp.lx.threads), lds, c.stream,
                           a, ms, rows);
        if (c.timing) timing_end(c, timing_entry(c, "demag_x_r2c"), ev);
    }
    ColArgs<Real> ca{};
    ca.data = spec;
    ca.ncols = dp.hx;
    ca.K = kview<Real>(c);
    if (dp.three_d) {
        // y forward: planes q = (component, z)
        ca.tw = static_cast<const cplx<Real> *>(dp.tw_y);
        ca.lp = dp.ly;
        ca.n_in = dp.ny;
        ca.n_out = dp.py;
        ca.estride = dp.hx;
        ca.tile = dp.y_tile;
        ca.q_div = dp.nz;
        ca.s_hi = comp_stride;
        ca.s_lo = (int64_t)dp.rows * dp.hx;
        ca.inverse = 0;
        ca.comp_stride = 0;
        launch_col(c, "demag_y_fwd", 0, ca, dim3((unsigned)cdiv(dp.hx, ca.tile), 3 * dp.nz));
        // z forward, tensor, z inverse: lines (ky, kx), three components
        ColArgs<Real> cz = ca;
        cz.tw = static_cast<const cplx<Real> *>(dp.tw_z);
        cz.lp = dp.lz;
        cz.n_in = dp.nz;
        cz.n_out = dp.nz;
        cz.estride = (int64_t)dp.rows * dp.hx;
        cz.tile = dp.z_tile;
        cz.q_div = 1;
        cz.s_hi = 0;
        cz.s_lo = dp.hx;
        cz.comp_stride = comp_stride;
        launch_col(c, "demag_z_conv", 3, cz, dim3((unsigned)cdiv(dp.hx, cz.tile), dp.py));
        // y inverse
        ca.n_in = dp.py;
        ca.n_out = dp.ny;
        ca.inverse = 1;
        launch_col(c, "demag_y_inv", 0, ca, dim3((unsigned)cdiv(dp.hx, ca.tile), 3 * dp.nz));
    } else {
        ca.tw = static_cast<const cplx<Real> *>(dp.tw_y);
        ca.lp = dp.ly;
        ca.n_in = dp.ny;
        ca.n_out = dp.ny;
        ca.estride = dp.hx;
        ca.tile = dp.yconv_tile;
        ca.q_div = 1;
        ca.s_hi = 0;
        ca.s_lo = 0;
        ca.inverse = 0;
        ca.comp_stride = comp_stride;
        launch_col(c, "demag_y_conv", 2, ca, dim3((unsigned)cdiv(dp.hx, ca.tile), 1));
    }
    // x inverse
    {
        FieldSink<Real> sink{out, c.d_mask, dp.nx, per_comp_lines, (int64_t)c.n};
        XArgs<Real> a = xargs<Real>(dp, lines, dp.x_lines_per_wg, 0, dp.nx, Real(dp.inv_npad));
        const int64_t pairs = (lines + 1) / 2;
        dim3 grid((unsigned)cdiv(pairs, a.lpw));
        const size_t lds = (size_t)a.lpw * (dp.lx.n + kPad) * sizeof(cplx<Real>);
        hipEvent_t ev{};
        if (c.timing) timing_begin(c, timing_entry(c, "demag_x_c2r"), &ev);
        hipLaunchKernelGGL((xinv_kernel<Real, FieldSink<Real>>), grid, dim3(a.lpw * dp.lx.threads), lds, c.stream,
                           a, rows, sink);
        if (c.timing) timing_end(c, timing_entry(c, "demag_x_c2r"), ev);
    }
}

template <typename Real>
bool upload_twiddles(Context &c, int n, void **dst) {
    std::vector<cplx<double>> w = twiddle_table(n);
    const size_t bytes = w.size() * sizeof(cplx<Real>);
    if (!ctx_alloc(c, dst, bytes, "twiddle table")) return false;
    if (sizeof(Real) == sizeof(double)) {
        return hip_ok(hipMemcpy(*dst, w.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(twiddles)", c.last_error);
    }
    std::vector<cplx<float>> wf(w.size());
    for (size_t i = 0; i < w.size(); ++i) wf[i] = {(float)w[i].x, (float)w[i].y};
    return hip_ok(hipMemcpy(*dst, wf.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(twiddles)", c.last_error);
}

// Shared plan construction for the demag convolution and the Newell setup.
bool plan_lines(DemagPlan &dp, std::string &err) {
    // column passes prefer radices >= 8 (fewer lanes per line, wider tiles)
    if (!plan_line(dp.px, dp.lx, err, false)) return false;
    if (!plan_line(dp.py, dp.ly, err, true)) return false;
    if (dp.three_d && !plan_line(dp.pz, dp.lz, err, true)) return false;
    return true;
}

// Runs the padded forward transform of the Newell kernel (FP64 throughout)
// into a temporary half spectrum [6][pz][py][hx].
bool newell_forward(Context &c, const DemagPlan &dp, const std::vector<double> &oct, cplx<double> **spec_out,
                    std::string &err) {
    void *d_oct = nullptr, *d_spec = nullptr, *twx = nullptr, *twy = nullptr, *twz = nullptr;
    const int pz = dp.three_d ? dp.pz : 1;
    const int64_t spec_comp = (int64_t)pz * dp.py * dp.hx;
    const size_t spec_bytes = 6 * (size_t)spec_comp * sizeof(cplx<double>);
    bool ok = hip_ok(hipMalloc(&d_oct, oct.size() * sizeof(double)), "hipMalloc(newell octant)", err) &&
              hip_ok(hipMalloc(&d_spec, spec_bytes), "hipMalloc(newell spectrum)", err) &&
              hip_ok(hipMemcpy(d_oct, oct.data(), oct.size() * sizeof(double), hipMemcpyHostToDevice),
                     "hipMemcpy(newell octant)", err);
    auto up = [&](int n, void **dst) {
        std::vector<cplx<double>> w = twiddle_table(n);
        return hip_ok(hipMalloc(dst, w.size() * sizeof(cplx<double>)), "hipMalloc(twiddles)", err) &&
               hip_ok(hipMemcpy(*dst, w.data(), w.size() * sizeof(cplx<double>), hipMemcpyHostToDevice),
                      "hipMemcpy(twiddles)", err);
    };
    ok = ok && up(dp.px, &twx) && up(dp.py, &twy) && (!dp.three_d || up(dp.pz, &twz));
    if (ok) {
        cplx<double> *spec = static_cast<cplx<double> *>(d_spec);
        const int kz_src = dp.three_d ? dp.nz : 1;
        NewellSource src{static_cast<const double *>(d_oct), dp.nx, dp.ny, kz_src, dp.px, dp.py, pz};
        const int64_t lines = 6 * (int64_t)dp.py * pz;
        SpecRows<double> rows = spec_rows<double>(spec, (int64_t)dp.py * pz, spec_comp, dp.py, dp.py, dp.hx);
        XArgs<double> a{};
        a.tw = static_cast<const cplx<double> *>(twx);
        a.lp = dp.lx;
        a.lines = lines;
        a.lpw = pick_lpw(dp.lx, sizeof(cplx<double>));
        a.n_in = dp.px;
        a.n_out = 0;
        a.hx = dp.hx;
        a.scale = 1.0;
        const size_t lds = (size_t)a.lpw * (dp.lx.n + kPad) * sizeof(cplx<double>);
        hipLaunchKernelGGL((xfwd_kernel<double, NewellSource>), dim3((unsigned)cdiv((lines + 1) / 2, a.lpw)),
                           dim3(a.lpw * dp.lx.threads), lds, c.stream, a, src, rows);
        ColArgs<double> ca{};
        ca.data = spec;
        ca.ncols = dp.hx;
        ca.tw = static_cast<const cplx<double> *>(twy);
        ca.lp = dp.ly;
        ca.n_in = dp.py;
        ca.n_out = dp.py;
        ca.estride = dp.hx;
        ca.tile = pick_tile(dp.ly, 1, sizeof(cplx<double>), dp.hx);
        ca.q_div = pz;
        ca.s_hi = spec_comp;
        ca.s_lo = (int64_t)dp.py * dp.hx;
        ca.inverse = 0;
        if ((size_t)ca.tile * (dp.ly.n + kPad) * sizeof(cplx<double>) > lds_budget()) {
            err = "Newell setup: y line does not fit in LDS";
            ok = false;
        }
        if (ok) {
            hipLaunchKernelGGL((col_kernel<double, 0>), dim3((unsigned)cdiv(dp.hx, ca.tile), 6 * pz),
                               dim3(ca.tile * dp.ly.threads), (size_t)ca.tile * (dp.ly.n + kPad) * sizeof(cplx<double>),
                               c.stream, ca);
        }
        if (ok && dp.three_d) {
            ColArgs<double> cz = ca;
            cz.tw = static_cast<const cplx<double> *>(twz);
            cz.lp = dp.lz;
            cz.n_in = dp.pz;
            cz.n_out = dp.pz;
            cz.estride = (int64_t)dp.py * dp.hx;
            cz.tile = pick_tile(dp.lz, 1, sizeof(cplx<double>), dp.hx);
            cz.q_div = dp.py;
            cz.s_hi = spec_comp;
            cz.s_lo = dp.hx;
            hipLaunchKernelGGL((col_kernel<double, 0>), dim3((unsigned)cdiv(dp.hx, cz.tile), 6 * dp.py),
                               dim3(cz.tile * dp.lz.threads),
                               (size_t)cz.tile * (dp.lz.n + kPad) * sizeof(cplx<double>), c.stream, cz);
        }
        ok = ok && hip_ok(hipStreamSynchronize(c.stream), "Newell setup transform", err);
        ok = ok && hip_ok(hipGetLastError(), "Newell setup launch", err);
    }
    if (d_oct) hipFree(d_oct);
    if (twx) hipFree(twx);
    if (twy) hipFree(twy);
    if (twz) hipFree(twz);
    if (!ok) {
        if (d_spec) hipFree(d_spec);
        return false;
    }
    *spec_out = static_cast<cplx<double> *>(d_spec);
    return true;
}

bool host_octant_from_spectra(const DemagPlan &dp, const fullmag_fdm_plan_desc &plan, std::vector<double> &kern,
                              std::string &err) {
    // Full spectra are [pz][py][px] interleaved complex; keep Re on the octant.
    const double *src[6] = {plan.demag_kernel_xx_spectrum, plan.demag_kernel_yy_spectrum,
                            plan.demag_kernel_zz_spectrum, plan.demag_kernel_xy_spectrum,
                            plan.demag_kernel_xz_spectrum, plan.demag_kernel_yz_spectrum};
    const int pz = dp.three_d ? dp.pz : 1;
    const int64_t plane = (int64_t)(dp.ny + 1) * dp.hx;
    kern.assign((size_t)dp.kcomp * dp.kz_n * plane, 0.0);
    // parity signs of the six components along (x, y, z)
    const int odd[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    double kmax = 0.0, bad = 0.0;
    for (int c = 0; c < 6; ++c) {
        const double *s = src[c];
        for (int z = 0; z < pz; ++z)
            for (int y = 0; y < dp.py; ++y)
                for (int x = 0; x < dp.px; ++x) {
                    const int64_t i = ((int64_t)z * dp.py + y) * dp.px + x;
                    const double re = s[2 * i], im = s[2 * i + 1];
                    kmax = std::max(kmax, std::fabs(re));
                    bad = std::max(bad, std::fabs(im));
                    // parity against the mirrored octant entry
                    const int ox = x <= dp.nx ? x : dp.px - x;
                    const int oy = y <= dp.ny ? y : dp.py - y;
                    const int oz = (pz == 1 || z <= dp.nz) ? z : pz - z;
                    double sg = 1.0;
                    if (odd[c][0] && x > dp.nx) sg = -sg;
                    if (odd[c][1] && y > dp.ny) sg = -sg;
                    if (odd[c][2] && pz > 1 && z > dp.nz) sg = -sg;
                    const int64_t j = ((int64_t)oz * dp.py + oy) * dp.px + ox;
                    bad = std::max(bad, std::fabs(re - sg * s[2 * j]));
                    if (x <= dp.nx && y <= dp.ny && (pz == 1 || z <= dp.nz) && c < dp.kcomp)
                        kern[((size_t)c * dp.kz_n + z) * plane + (int64_t)y * dp.hx + x] = re;
                }
    }
    if (bad > 1e-8 * std::max(kmax, 1e-300)) {
        err = "demag kernel spectra are not real with Newell parity symmetry (deviation " + std::to_string(bad) +
              " vs max " + std::to_string(kmax) + ")";
        return false;
    }
    if (dp.kcomp == 4) {
        // thin film: xz and yz vanish on the film plane; check they are noise
        double m2 = 0.0;
        for (int c = 4; c < 6; ++c)
            for (int64_t i = 0; i < (int64_t)dp.px * dp.py; ++i) m2 = std::max(m2, std::fabs(src[c][2 * i]));
        if (m2 > 1e-8 * std::max(kmax, 1e-300)) {
            err = "thin-film demag spectra carry non-zero xz/yz components";
            return false;
        }
    }
    return true;
}

template <typename Real>
bool upload_kern(Context &c, const std::vector<double> &kern) {
    DemagPlan &dp = c.dp;
    dp.kern_elems = kern.size();
    const size_t bytes = kern.size() * sizeof(Real);
    if (!ctx_alloc(c, &dp.kern, bytes, "demag tensor")) return false;
    if (sizeof(Real) == sizeof(double))
        return hip_ok(hipMemcpy(dp.kern, kern.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(tensor)",
                      c.last_error);
    std::vector<float> f(kern.size());
    for (size_t i = 0; i < kern.size(); ++i) f[i] = (float)kern[i];
    return hip_ok(hipMemcpy(dp.kern, f.data(), bytes, hipMemcpyHostToDevice), "hipMemcpy(tensor)", c.last_error);
}

template <typename Real>
bool setup_t(Context &c, const fullmag_fdm_plan_desc &plan, bool newell_on_device) {
    DemagPlan &dp = c.dp;
    const size_t ce = sizeof(cplx<Real>);
    dp.x_lines_per_wg = pick_lpw(dp.lx, ce);
    if (dp.three_d) {
        dp.y_tile = pick_tile(dp.ly, 1, ce, dp.hx);
        dp.z_tile = pick_tile(dp.lz, 3, ce, dp.hx);
    } else {
        dp.yconv_tile = pick_tile(dp.ly, 3, ce, dp.hx);
    }
    if (!geometry_ok<Real>(dp, c.last_error)) return false;
    if (!upload_twiddles<Real>(c, dp.px, &dp.tw_x) || !upload_twiddles<Real>(c, dp.py, &dp.tw_y)) return false;
    if (dp.three_d && !upload_twiddles<Real>(c, dp.pz, &dp.tw_z)) return false;
    dp.spec_elems = 3ull * dp.nz * dp.rows * dp.hx;
    if (!ctx_alloc(c, &dp.spec, dp.spec_elems * ce, "demag spectrum")) return false;
    if (dp.projection) return true;
    if (plan.demag_kernel_spectrum_len != 0) {
        std::vector<double> kern;
        if (!host_octant_from_spectra(dp, plan, kern, c.last_error)) return false;
        return upload_kern<Real>(c, kern);
    }
    (void)newell_on_device;
    // In-library Newell build: octant on the host, padded transform on the device.
    std::vector<double> oct;
    newell_octant(dp.nx, dp.ny, dp.three_d ? dp.nz : 1, c.dx, c.dy, c.dz, oct);
    cplx<double> *spec = nullptr;
    if (!newell_forward(c, dp, oct, &spec, c.last_error)) return false;
    dp.kern_elems = (uint64_t)dp.kcomp * dp.kz_n * (dp.ny + 1) * dp.hx;
    bool ok = ctx_alloc(c, &dp.kern, dp.kern_elems * sizeof(Real), "demag tensor");
    if (ok) {
        const int pz = dp.three_d ? dp.pz : 1;
        const int64_t spec_comp = (int64_t)pz * dp.py * dp.hx;
        hipLaunchKernelGGL((extract_octant_kernel<Real>), dim3((unsigned)cdiv(dp.kern_elems, 256)), dim3(256), 0,
                           c.stream, spec, static_cast<Real *>(dp.kern), dp.kcomp, dp.kz_n, dp.ny, dp.hx, dp.py,
                           spec_comp, (int64_t)dp.kern_elems);
        ok = hip_ok(hipStreamSynchronize(c.stream), "extract octant", c.last_error);
    }
    hipFree(spec);
    return ok;
}

}  // namespace

bool demag_setup(Context &c, const fullmag_fdm_plan_desc &plan, bool newell_on_device) {
    DemagPlan &dp = c.dp;
    dp = DemagPlan{};
    if (!c.demag) return true;
    dp.enabled = true;
    dp.nx = c.nx;
    dp.ny = c.ny;
    dp.nz = c.nz;
    const uint64_t len2d = 2ull * (2ull * c.nx) * (2ull * c.ny);
    const bool have_spectra = plan.demag_kernel_spectrum_len != 0;
    dp.projection = !have_spectra && !newell_on_device;
    dp.three_d = !(c.nz == 1 && ((have_spectra && plan.demag_kernel_spectrum_len == len2d) ||
                                 (!have_spectra && newell_on_device)));
    dp.px = 2 * c.nx;
    dp.py = 2 * c.ny;
    dp.pz = dp.three_d ? 2 * c.nz : 1;
    dp.hx = c.nx + 1;
    dp.rows = dp.three_d ? dp.py : dp.ny;
    dp.kcomp = dp.three_d ? 6 : 4;
    dp.kz_n = dp.three_d ? dp.nz + 1 : 1;
    dp.inv_npad = 1.0 / ((double)dp.px * dp.py * dp.pz);
    if (!plan_lines(dp, c.last_error)) return false;
    if (c.mode == MODE_DOUBLE) return setup_t<double>(c, plan, newell_on_device);
    return setup_t<float>(c, plan, newell_on_device);
}

void demag_eval(Context &c, const void *src, bool src_f64, void *out) {
    if (c.mode == MODE_DOUBLE) {
        demag_eval_t<double, double>(c, static_cast<const double *>(src), static_cast<double *>(out));
    } else if (src_f64) {
        demag_eval_t<float, double>(c, static_cast<const double *>(src), static_cast<float *>(out));
    } else {
        demag_eval_t<float, float>(c, static_cast<const float *>(src), static_cast<float *>(out));
    }
}

bool newell_spectra_full(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double dz, bool thin_film_2d,
                         double *out, std::string &err) {
    Context c;
    c.nx = nx; c.ny = ny; c.nz = nz; c.dx = dx; c.dy = dy; c.dz = dz;
    c.mode = MODE_DOUBLE;
    if (!hip_ok(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate", err)) return false;
    DemagPlan dp;
    dp.nx = nx; dp.ny = ny; dp.nz = nz;
    dp.three_d = !thin_film_2d;
    dp.px = 2 * nx; dp.py = 2 * ny; dp.pz = dp.three_d ? 2 * nz : 1;
    dp.hx = nx + 1;
    bool ok = plan_lines(dp, err);
    std::vector<double> oct;
    cplx<double> *spec = nullptr;
    if (ok) {
        newell_octant(nx, ny, dp.three_d ? nz : 1, dx, dy, dz, oct);
        ok = newell_forward(c, dp, oct, &spec, err);
    }
    if (ok) {
        const int pz = dp.pz;
        const int64_t spec_comp = (int64_t)pz * dp.py * dp.hx;
        std::vector<cplx<double>> h((size_t)6 * spec_comp);
        ok = hip_ok(hipMemcpy(h.data(), spec, h.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost),
                    "hipMemcpy(newell spectrum)", err);
        hipFree(spec);
        if (ok) {
            const int64_t npad = (int64_t)dp.px * dp.py * pz;
            for (int comp = 0; comp < 6; ++comp)
                for (int z = 0; z < pz; ++z)
                    for (int y = 0; y < dp.py; ++y)
                        for (int x = 0; x < dp.px; ++x) {
                            cplx<double> v;
                            if (x < dp.hx) {
                                v = h[comp * spec_comp + ((int64_t)z * dp.py + y) * dp.hx + x];
                            } else {
                                // Hermitian continuation of a real kernel
                                const int zz = (pz - z) % pz, yy = (dp.py - y) % dp.py, xx = dp.px - x;
                                v = h[comp * spec_comp + ((int64_t)zz * dp.py + yy) * dp.hx + xx];
                                v.y = -v.y;
                            }
                            const int64_t i = ((int64_t)z * dp.py + y) * dp.px + x;
                            out[comp * 2 * npad + 2 * i] = v.x;
                            out[comp * 2 * npad + 2 * i + 1] = v.y;
                        }
        }
    }
    hipStreamDestroy(c.stream);
    return ok;
}

}  // namespace fm
