// demag.hip -- zero-padded Newell-tensor demag convolution on gfx950.
//
// Replaces the cuFFT pipeline of native/backends/fdm/src/demag_fp64.cu:264-381
// (pack, 3 forward Z2Z transforms over the fully padded grid, 6-component
// complex tensor multiply, 3 inverse transforms, unpack).  The mathematics is
// that of crates/fullmag-engine/src/lib.rs:2306-2400,
//     H = IFFT( -K . FFT(Ms m) ) / Npad   on a (2nx, 2ny, 2nz) grid,
// organised around what is non-zero and around coalesced HBM access:
//
//   x fwd  R2C of 16 real rows per workgroup (two rows per complex transform,
//          only nx of 2nx inputs loaded); the nx+1 outputs of each row are
//          written transposed, so ky runs are contiguous in the spectrum.
//   y fwd  contiguous lines, ny of 2ny inputs loaded (3-D only).
//   z conv per (kx, ky): 3 components x (nz inputs -> 2nz), 3x3 tensor,
//          inverse, nz outputs; in registers for 2nz <= 16 (3-D only).
//   y inv  contiguous lines, 2ny inputs, ny outputs written (3-D only).
//   x inv  C2R of 16 rows, reading transposed runs, nx outputs per row,
//          scaled by 1/Npad and masked.
// Thin films (nz = 1, 2-D spectra): x fwd, y conv (forward y, tensor at
// kz = 0, inverse y in one kernel), x inv.
//
// Spectrum layout: spec[c][z][kx][ky] with z < nz (3-D) or a single plane
// (2-D), kx <= nx, ky < pitch (2ny in 3-D, ny in 2-D).
// The Newell spectra are real (the padded kernel is even or odd along every
// axis); the tensor is stored on the non-negative octant as
// blocks per (kx, ky) column (kz <= nz, ky <= ny, kern_at) and mirrored with
// parity signs.
#include "context.hpp"

#include <atomic>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace fm {
namespace {

constexpr int kPad = 1;         // LDS line padding (complex slots)

// Stand-ins read in place of an absent optional per-cell array, so that the
// read itself needs no branch: an active-mask byte of 1 and a value of 1.
__device__ const uint8_t kNoMask = 1;
__device__ const double kNoArrayF64 = 1.0;
__device__ const float kNoArrayF32 = 1.0f;
template <typename T>
__device__ __forceinline__ const T *no_array() {
    if constexpr (sizeof(T) == 4) return &kNoArrayF32;
    else return &kNoArrayF64;
}
constexpr int kMaxBlock = 512;  // lanes per workgroup for FFT kernels
// largest padded z length done in registers (a lane pair per column holds
// 3 x pz complex values)
template <typename Real>
constexpr int z_reg_max() {
    return 16;
}

// Workgroup b runs on XCD b mod 8 (round-robin dispatch over the 8 XCDs of
// an MI355X).  Renumbered so that each XCD receives one contiguous range of
// logical workgroups.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g) {
    const uint32_t x = b & 7u, q = g >> 3, r = g & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// ---------------------------------------------------------------------------
// Tensor access
// ---------------------------------------------------------------------------
template <typename Real>
struct KernelView {
    const Real *k = nullptr;  // octant tensor, nullptr for the projection fallback
    int kcomp = 6, kz_n = 1;
    int kb = 0, kodd = 0;     // block per (kx, oy) and offset of its odd-oz half
    int ny = 0, nz = 0, py = 0, pz = 1, hx = 0;
    double kxs = 0, kys = 0, kzs = 0;  // 2 pi / (p d) per axis (projection)
};

// Octant tensor layout: one block of `kblock` reals per (kx, oy) column,
// holding the entries of even oz (0, 2, ...) and, from `kodd`, those of odd
// oz, each entry `kcomp` consecutive components, each half padded to 16
// bytes.  A z-transform lane that owns kz of one parity reads its whole half
// with 16-byte loads (kz and pz - kz share an entry and a parity).
__host__ __device__ __forceinline__ int64_t kern_at(int kblock, int kodd, int kcomp, int64_t ly, int c, int oz,
                                                    int kx, int oy) {
    return ((int64_t)kx * ly + oy) * kblock + ((oz & 1) ? kodd : 0) + (int64_t)(oz >> 1) * kcomp + c;
}

// t = (xx, yy, zz, xy, xz, yz) at frequency (kx, ky, kz); kx <= nx.
template <typename Real>
__device__ __forceinline__ void tensor_at(const KernelView<Real> &K, int kx, int ky, int kz, Real t[6]) {
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
        const Real *e = K.k + kern_at(K.kb, K.kodd, K.kcomp, K.ny + 1, 0, oz, kx, oy);
        t[0] = e[0];
        t[1] = e[1];
        t[2] = e[2];
        t[3] = sy * e[3];
        if (K.kcomp == 6) {
            t[4] = sz * e[4];
            t[5] = sy * sz * e[5];
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

template <typename Real>
__device__ __forceinline__ void apply_tensor(const Real t[6], cplx<Real> &mx, cplx<Real> &my, cplx<Real> &mz) {
    const cplx<Real> X = mx, Y = my, Z = mz;
    mx = {-(t[0] * X.x + t[3] * Y.x + t[4] * Z.x), -(t[0] * X.y + t[3] * Y.y + t[4] * Z.y)};
    my = {-(t[3] * X.x + t[1] * Y.x + t[5] * Z.x), -(t[3] * X.y + t[1] * Y.y + t[5] * Z.y)};
    mz = {-(t[4] * X.x + t[5] * Y.x + t[2] * Z.x), -(t[4] * X.y + t[5] * Y.y + t[2] * Z.y)};
}

// ---------------------------------------------------------------------------
// Spectrum addressing: spec[c][z][kx][ky], `planes` z-planes, ky pitch.
// ---------------------------------------------------------------------------
struct SpecGeom {
    int hx;       // half-spectrum width nx + 1
    int planes;   // z planes stored
    int pitch;    // ky line length stored
    __device__ __host__ int64_t at(int c, int z, int kx, int ky) const {
        return (((int64_t)c * planes + z) * hx + kx) * pitch + ky;
    }
};

// ---------------------------------------------------------------------------
// x passes
// ---------------------------------------------------------------------------
template <typename Real>
struct XArgs {
    const cplx<Real> *tw;
    LinePlan lp;      // length px
    int xr, xr_log2;  // real rows per workgroup (power of two) and its log2
    int ls;           // LDS line stride (complex slots)
    int rows;         // valid real rows per (c, z) plane
    int n_in;         // forward: inputs loaded per row
    int n_out;        // inverse: outputs stored per row
    Real scale;       // inverse: 1/Npad
    SpecGeom g;
};

// Ms m at (c, z, y, x) of the state; zero on inactive cells.  GATHER: the
// x pass reads the mask, volume fraction and per-cell Ms beside m with
// unconditional loads (a separate instantiation, so that the kernel for
// uniform material keeps its register budget).
template <typename Real, typename S, bool GATHER = false>
struct MagSource {
    const S *m;
    const uint8_t *mask;
    const double *ms_cell;
    double ms;
    int nx, ny, nz;
    int64_t n;
    // boundary tiers: every cell packs phi Ms m, mask or not (the volume
    // fraction in the field's precision)
    const Real *phi;
    __device__ Real operator()(int c, int z, int y, int x) const {
        const int64_t i = ((int64_t)z * ny + y) * nx + x;
        if (phi) return Real(double(m[(int64_t)c * n + i]) * (double(phi[i]) * (ms_cell ? ms_cell[i] : ms)));
        if (mask && !mask[i]) return Real(0);
        const double w = ms_cell ? ms_cell[i] : ms;
        return Real(double(m[(int64_t)c * n + i]) * w);
    }
    // No mask and no per-cell Ms: the value is one unconditional load, so a
    // lane's loads can all be issued before any of them is used.
    __device__ bool uniform() const { return !mask && !ms_cell && !phi; }
    __device__ Real uniform_at(int c, int z, int y, int x) const {
        return Real(double(m[(int64_t)c * n + ((int64_t)z * ny + y) * nx + x]) * ms);
    }
    // operator()'s value with every load unconditional (an absent array is
    // read from a stand-in and its value replaced after the load), so a
    // lane's loads all issue before the first is used.
    static constexpr bool kGather = GATHER;
    __device__ Real gather_at(int c, int z, int y, int x) const {
        const int64_t i = ((int64_t)z * ny + y) * nx + x;
        const double mv = double(m[(int64_t)c * n + i]);
        const double pv = double(*(phi ? phi + i : no_array<Real>()));
        const double cv = *(ms_cell ? ms_cell + i : &kNoArrayF64);
        const uint8_t kv = *(mask ? mask + i : &kNoMask);
        const double w = ms_cell ? cv : ms;
        const bool live = phi || kv != 0;
        return live ? Real(mv * (phi ? pv * w : w)) : Real(0);
    }
};

// Padded real-space Newell kernel at (c, z, y, x) of the (px, py, pz) grid,
// mirrored from the octant with the placement signs of
// crates/fullmag-fdm-demag/src/newell.rs:397-426 (index n of each padded
// axis is never written there and stays zero).
struct NewellSource {
    const double *oct;  // [6][nz][ny][nx]
    int nx, ny, nz, px, py, pz;
    int z_even;  // the cross-layer kernel: mirrored z slots keep every sign (newell.rs:630-647)
    __device__ bool uniform() const { return false; }
    __device__ double uniform_at(int, int, int, int) const { return 0.0; }
    [[maybe_unused]] static constexpr bool kGather = false;  // read by the fixed-length x pass only
    __device__ double gather_at(int, int, int, int) const { return 0.0; }
    __device__ double operator()(int c, int zz, int yy, int x) const {
        if (x == nx || yy == ny || (pz > 1 && zz == nz)) return 0.0;
        const int xi = x < nx ? x : px - x;
        const int yi = yy < ny ? yy : py - yy;
        const int zi = (pz == 1 || zz < nz) ? zz : pz - zz;
        const double sx = x < nx ? 1.0 : -1.0, sy = yy < ny ? 1.0 : -1.0;
        const double sz = (pz == 1 || zz < nz || z_even) ? 1.0 : -1.0;
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
    int nx, ny;
    int64_t n;
    __device__ void operator()(int c, int z, int y, int x, Real v) const {
        const int64_t i = ((int64_t)z * ny + y) * nx + x;
        out[(int64_t)c * n + i] = (mask && !mask[i]) ? Real(0) : v;
    }
    // With a mask, operator() in two halves, so that a lane reads all its
    // mask bytes before its first store (a byte read may alias the stores,
    // so the compiler keeps each one behind the store before it).
    __device__ bool masked() const { return mask != nullptr; }
    __device__ uint8_t keep(int z, int y, int x) const { return mask[((int64_t)z * ny + y) * nx + x]; }
    __device__ void put(int c, int z, int y, int x, Real v, uint8_t keep) const {
        out[(int64_t)c * n + ((int64_t)z * ny + y) * nx + x] = keep ? v : Real(0);
    }
};

// Forward real-to-complex along x for xr rows (y0 .. y0+xr-1) of plane
// (c = blockIdx.z, z = blockIdx.y).  Rows 2j, 2j+1 are packed as a + i b;
// with Z = FFT(a + i b):  A_k = (Z_k + conj Z_{N-k}) / 2,
// B_k = (Z_k - conj Z_{N-k}) / (2i), k = 0..N/2, stored at spec[c][z][k][y].
template <typename Real, class Src, int N>
__device__ __forceinline__ void xfwd_body(const XArgs<Real> &a, const Src &src, cplx<Real> *spec, int c) {
    using L = LineOps<N>;
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int LINES = a.xr >> 1;
    const int n = L::n(a.lp), ls = a.ls, T = L::threads(a.lp);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int y0 = blockIdx.x * a.xr, z = blockIdx.y;
    // the T lanes of packed line j (= tid / T, blockDim.x = LINES * T) load it
    const int jl = tid / T, t = tid - jl * T;
    cplx<Real> *line = lds + jl * ls;
    const int lx = L::line_xor(jl);
    {
        const int ya = y0 + 2 * jl;
        if constexpr (N > 0) {
            // fixed length: all 16 loads of a lane in flight before the LDS writes
            cplx<Real> v[16];
            if (!Src::kGather && src.uniform() && 2 * a.n_in <= N) {
                // branch-free: clamped addresses, every load issued before
                // any is used; elements t + e N/16 with e >= 8 lie in the
                // zero padding (n_in <= N/2) and are not loaded
                const bool oka = ya < a.rows, okb = ya + 1 < a.rows;
                const int yA = oka ? ya : 0, yB = okb ? ya + 1 : 0;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int x = t + e * (N / 16);
                    const bool okx = x < a.n_in;
                    const int xs = okx ? x : 0;
                    const Real va = src.uniform_at(c, z, yA, xs), vb = src.uniform_at(c, z, yB, xs);
                    v[e] = {(okx && oka) ? va : Real(0), (okx && okb) ? vb : Real(0)};
                }
#pragma unroll
                for (int e = 8; e < 16; ++e) v[e] = {Real(0), Real(0)};
            } else if (Src::kGather && 2 * a.n_in <= N) {
                // the same with the mask, volume fraction and per-cell Ms
                // read beside each element, all unconditionally
                const bool oka = ya < a.rows, okb = ya + 1 < a.rows;
                const int yA = oka ? ya : 0, yB = okb ? ya + 1 : 0;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int x = t + e * (N / 16);
                    const bool okx = x < a.n_in;
                    const int xs = okx ? x : 0;
                    const Real va = src.gather_at(c, z, yA, xs), vb = src.gather_at(c, z, yB, xs);
                    v[e] = {(okx && oka) ? va : Real(0), (okx && okb) ? vb : Real(0)};
                }
#pragma unroll
                for (int e = 8; e < 16; ++e) v[e] = {Real(0), Real(0)};
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int x = t + e * (N / 16);
                    v[e] = {Real(0), Real(0)};
                    if (x < a.n_in) {
                        if (ya < a.rows) v[e].x = src(c, z, ya, x);
                        if (ya + 1 < a.rows) v[e].y = src(c, z, ya + 1, x);
                    }
                }
            }
            if constexpr (p2_first_from_regs<N>()) {
                // the loaded elements are the first radix-16 butterfly's inputs
                p2_first_pass_regs<Real, N>(line, t, v, a.tw, false, lx);
                p2_passes<Real, N, 1, 16>(line, t, a.tw, false, lx);
            } else {
#pragma unroll
                for (int e = 0; e < 16; ++e) line[L::slot(a.lp, t + e * (N / 16), lx)] = v[e];
                __syncthreads();
                L::template fft<Real>(line, a.lp, t, a.tw, false, lx);
            }
        } else {
            for (int x = t; x < n; x += T) {
                cplx<Real> v = {Real(0), Real(0)};
                if (x < a.n_in) {
                    if (ya < a.rows) v.x = src(c, z, ya, x);
                    if (ya + 1 < a.rows) v.y = src(c, z, ya + 1, x);
                }
                line[L::slot(a.lp, x, lx)] = v;
            }
            __syncthreads();
            L::template fft<Real>(line, a.lp, t, a.tw, false, lx);
        }
    }
    __syncthreads();  // the unpack reads every line
    const int hx = a.g.hx;
    // lane (j, k) splits line j into rows 2j (A) and 2j + 1 (B) at kx = k;
    // the two are adjacent in memory (one 16-byte store in FP32 when the
    // row pitch is even)
    if constexpr (sizeof(Real) == 8) {
        // FP64: one 16-byte complex per lane, consecutive lanes on consecutive rows
        for (int idx = tid; idx < hx * a.xr; idx += nthr) {
            const int yl = idx & (a.xr - 1), k = idx >> a.xr_log2;
            const int y = y0 + yl;
            if (y >= a.rows) continue;
            const int j = yl >> 1, jx = L::line_xor(j);
            const cplx<Real> zk = lds[j * ls + L::slot(a.lp, k, jx)];
            const cplx<Real> zn = lds[j * ls + L::slot(a.lp, k == 0 ? 0 : n - k, jx)];
            const Real h = Real(0.5);
            cplx<Real> v;
            if ((yl & 1) == 0) v = {h * (zk.x + zn.x), h * (zk.y - zn.y)};
            else v = {h * (zk.y + zn.y), h * (zn.x - zk.x)};
            spec[a.g.at(c, z, k, y)] = v;
        }
        return;
    }
    const bool pair16 = (a.g.pitch & 1) == 0;
    for (int idx = tid; idx < hx * LINES; idx += nthr) {
        const int j = idx & (LINES - 1), k = idx >> (a.xr_log2 - 1);
        const int y = y0 + 2 * j;
        if (y >= a.rows) continue;
        const int jx = L::line_xor(j);
        const cplx<Real> zk = lds[j * ls + L::slot(a.lp, k, jx)];
        const cplx<Real> zn = lds[j * ls + L::slot(a.lp, k == 0 ? 0 : n - k, jx)];
        const Real h = Real(0.5);
        const cplx<Real> A = {h * (zk.x + zn.x), h * (zk.y - zn.y)};
        const cplx<Real> B = {h * (zk.y + zn.y), h * (zn.x - zk.x)};
        cplx<Real> *dst = spec + a.g.at(c, z, k, y);
        if (y + 1 < a.rows) {
            if (pair16) {
                *reinterpret_cast<float4 *>(dst) = float4{float(A.x), float(A.y), float(B.x), float(B.y)};
            } else {
                dst[0] = A;
                dst[1] = B;
            }
        } else {
            dst[0] = A;
        }
    }
}

// Component c = blockIdx.z.
template <typename Real, class Src, int N = 0>
__global__ void __launch_bounds__(kMaxBlock) xfwd_kernel(XArgs<Real> a, Src src, cplx<Real> *spec) {
    xfwd_body<Real, Src, N>(a, src, spec, (int)blockIdx.z);
}

// Multi-instance form: instance blockIdx.z / zdiv of the table, component
// blockIdx.z % zdiv.  Each multi-instance kernel copies its instance's
// arguments into a local before the body runs, so the body sees invariant
// values exactly as the plain kernel does (a reference into the table would be
// re-read after the body's stores, which changes how the compiler contracts
// the arithmetic, and with it the last bit of FP32 results).
template <typename Real, class Src>
struct XFwdPack {
    XArgs<Real> a;
    Src src;
    cplx<Real> *spec;
};
template <typename Real, class Src, int N = 0>
__global__ void __launch_bounds__(kMaxBlock) xfwd_multi(const XFwdPack<Real, Src> *tab, int zdiv) {
    const XFwdPack<Real, Src> p = tab[blockIdx.z / zdiv];
    xfwd_body<Real, Src, N>(p.a, p.src, p.spec, (int)(blockIdx.z % zdiv));
}

// Inverse complex-to-real along x: Z_k = A_k + i B_k for k <= N/2 and the
// Hermitian continuation A_{N-k} = conj A_k; z = IFFT(Z); rows a = Re z,
// b = Im z.  Reads spec[c][z][k][y0..y0+xr-1].
template <typename Real, class Sink, int N>
__device__ __forceinline__ void xinv_body(const XArgs<Real> &a, const cplx<Real> *spec, const Sink &sink, int c) {
    using L = LineOps<N>;
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int LINES = a.xr >> 1;
    const int n = L::n(a.lp), ls = a.ls, T = L::threads(a.lp);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int y0 = blockIdx.x * a.xr, z = blockIdx.y;
    const int hx = a.g.hx;
    // consecutive lanes read consecutive rows of one kx column
    auto put = [&](int idx, cplx<Real> A, cplx<Real> B) {
        const int j = idx & (LINES - 1), k = idx >> (a.xr_log2 - 1);
        const int jx = L::line_xor(j);
        lds[j * ls + L::slot(a.lp, k, jx)] = {A.x - B.y, A.y + B.x};
        if (k > 0 && k < n - k) lds[j * ls + L::slot(a.lp, n - k, jx)] = {A.x + B.y, B.x - A.y};
    };
    if constexpr (N > 0) {
        // fixed length: every load of the lane (hx / T <= U row pairs) in
        // flight before the LDS writes; a row pair is one 16-byte load in
        // FP32 when the row pitch is even
        constexpr int U = (N / 2 + 1 + N / 16 - 1) / (N / 16);
        const bool pair16 = sizeof(Real) == 4 && (a.g.pitch & 1) == 0;
        cplx<Real> A[U], B[U];
        // branch-free: clamped addresses, selects after the loads
        if (pair16) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = tid + u * nthr;
                const int j = idx & (LINES - 1), k = idx >> (a.xr_log2 - 1);
                const int ya = y0 + 2 * j;
                const bool oka = k < hx && ya < a.rows, okb = oka && ya + 1 < a.rows;
                // an even row pitch and an even ya keep the pair 16-byte aligned
                const float4 q = *reinterpret_cast<const float4 *>(spec + (oka ? a.g.at(c, z, k, ya) : 0));
                A[u] = oka ? cplx<Real>{Real(q.x), Real(q.y)} : cplx<Real>{Real(0), Real(0)};
                B[u] = okb ? cplx<Real>{Real(q.z), Real(q.w)} : cplx<Real>{Real(0), Real(0)};
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = tid + u * nthr;
                const int j = idx & (LINES - 1), k = idx >> (a.xr_log2 - 1);
                const int ya = y0 + 2 * j;
                const bool oka = k < hx && ya < a.rows, okb = oka && ya + 1 < a.rows;
                const cplx<Real> qa = spec[oka ? a.g.at(c, z, k, ya) : 0];
                const cplx<Real> qb = spec[okb ? a.g.at(c, z, k, ya + 1) : 0];
                A[u] = oka ? qa : cplx<Real>{Real(0), Real(0)};
                B[u] = okb ? qb : cplx<Real>{Real(0), Real(0)};
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = tid + u * nthr;
            if (idx < hx * LINES) put(idx, A[u], B[u]);
        }
    } else {
        // batches of four loads per lane in flight before their LDS writes
        constexpr int U = 4;
        for (int base = tid; base < hx * LINES; base += U * nthr) {
            cplx<Real> A[U], B[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = base + u * nthr;
                const int j = idx & (LINES - 1), k = idx >> (a.xr_log2 - 1);
                const int ya = y0 + 2 * j;
                A[u] = B[u] = {Real(0), Real(0)};
                if (idx < hx * LINES) {
                    if (ya < a.rows) A[u] = spec[a.g.at(c, z, k, ya)];
                    if (ya + 1 < a.rows) B[u] = spec[a.g.at(c, z, k, ya + 1)];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = base + u * nthr;
                if (idx >= hx * LINES) break;
                put(idx, A[u], B[u]);
            }
        }
    }
    __syncthreads();
    const int jl = tid / T, t = tid - jl * T;
    const cplx<Real> *line = lds + jl * ls;
    const int lx = L::line_xor(jl);
    L::template fft<Real>(lds + jl * ls, a.lp, t, a.tw, true, lx);
    // the lanes of line jl store its rows 2 jl (real part) and 2 jl + 1
    const int ya = y0 + 2 * jl;
    if constexpr (N > 0 && std::is_same<Sink, FieldSink<float>>::value) {
        if (2 * a.n_out <= N && sink.masked() && (a.n_out & 3) == 0 && (sink.n & 3) == 0) {
            // four consecutive outputs of a row per lane: their mask bytes in
            // one 4-byte load, every mask word read before the first store,
            // and the values in one 16-byte store (rows start on 16-byte
            // boundaries: nx = n_out and n are multiples of four)
            constexpr int TT = N / 16, Q = (N / 2 + 4 * TT - 1) / (4 * TT);
            const bool oka = ya < a.rows, okb = ya + 1 < a.rows;
            const int yA = oka ? ya : 0, yB = okb ? ya + 1 : 0;
            uint32_t ka[Q], kb[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int x = 4 * (t + q * TT);
                const int xs = x < a.n_out ? x : 0;
                ka[q] = *reinterpret_cast<const uint32_t *>(sink.mask + ((int64_t)z * sink.ny + yA) * sink.nx + xs);
                kb[q] = *reinterpret_cast<const uint32_t *>(sink.mask + ((int64_t)z * sink.ny + yB) * sink.nx + xs);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int x = 4 * (t + q * TT);
                if (x < a.n_out) {
                    cplx<Real> v[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = line[L::slot(a.lp, x + e, lx)];
                    auto row = [&](uint32_t k, int part) {
                        float o[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            o[e] = ((k >> (8 * e)) & 0xffu) ? (part ? v[e].y : v[e].x) * a.scale : 0.0f;
                        return float4{o[0], o[1], o[2], o[3]};
                    };
                    const int64_t i = ((int64_t)z * sink.ny + ya) * sink.nx + x;
                    if (oka) *reinterpret_cast<float4 *>(sink.out + (int64_t)c * sink.n + i) = row(ka[q], 0);
                    if (okb)
                        *reinterpret_cast<float4 *>(sink.out + (int64_t)c * sink.n + i + sink.nx) = row(kb[q], 1);
                }
            }
            return;
        }
    }
    if constexpr (N > 0) {
        if (2 * a.n_out <= N && sink.masked()) {
            // at most 8 outputs per row and lane: every mask byte read
            // (clamped addresses) before the first store
            const bool oka = ya < a.rows, okb = ya + 1 < a.rows;
            const int yA = oka ? ya : 0, yB = okb ? ya + 1 : 0;
            uint8_t ka[8], kb[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int x = t + e * T;
                const int xs = x < a.n_out ? x : 0;
                ka[e] = sink.keep(z, yA, xs);
                kb[e] = sink.keep(z, yB, xs);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int x = t + e * T;
                if (x < a.n_out) {
                    const cplx<Real> v = line[L::slot(a.lp, x, lx)];
                    if (oka) sink.put(c, z, ya, x, v.x * a.scale, ka[e]);
                    if (okb) sink.put(c, z, ya + 1, x, v.y * a.scale, kb[e]);
                }
            }
            return;
        }
    }
    if constexpr (std::is_same<Sink, FieldSink<float>>::value) {
        if (!sink.masked() && (a.n_out & 3) == 0 && (sink.n & 3) == 0) {
            // four consecutive outputs of a row per lane and one 16-byte
            // store (rows start on 16-byte boundaries: nx = n_out and n are
            // multiples of four)
            for (int x = 4 * t; x < a.n_out; x += 4 * T) {
                cplx<Real> v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = line[L::slot(a.lp, x + e, lx)];
                const int64_t i = ((int64_t)z * sink.ny + ya) * sink.nx + x;
                if (ya < a.rows)
                    *reinterpret_cast<float4 *>(sink.out + (int64_t)c * sink.n + i) =
                        float4{v[0].x * a.scale, v[1].x * a.scale, v[2].x * a.scale, v[3].x * a.scale};
                if (ya + 1 < a.rows)
                    *reinterpret_cast<float4 *>(sink.out + (int64_t)c * sink.n + i + sink.nx) =
                        float4{v[0].y * a.scale, v[1].y * a.scale, v[2].y * a.scale, v[3].y * a.scale};
            }
            return;
        }
    }
    for (int x = t; x < a.n_out; x += T) {
        const cplx<Real> v = line[L::slot(a.lp, x, lx)];
        if (ya < a.rows) sink(c, z, ya, x, v.x * a.scale);
        if (ya + 1 < a.rows) sink(c, z, ya + 1, x, v.y * a.scale);
    }
}

template <typename Real, class Sink, int N = 0>
__global__ void __launch_bounds__(kMaxBlock) xinv_kernel(XArgs<Real> a, const cplx<Real> *spec, Sink sink) {
    xinv_body<Real, Sink, N>(a, spec, sink, (int)blockIdx.z);
}

template <typename Real, class Sink>
struct XInvPack {
    XArgs<Real> a;
    const cplx<Real> *spec;
    Sink sink;
};
template <typename Real, class Sink, int N = 0>
__global__ void __launch_bounds__(kMaxBlock) xinv_multi(const XInvPack<Real, Sink> *tab, int zdiv) {
    const XInvPack<Real, Sink> p = tab[blockIdx.z / zdiv];
    xinv_body<Real, Sink, N>(p.a, p.spec, p.sink, (int)(blockIdx.z % zdiv));
}

// ---------------------------------------------------------------------------
// Contiguous line passes (y; and the Newell setup's z when long)
// ---------------------------------------------------------------------------
template <typename Real>
struct LineArgs {
    cplx<Real> *data;
    const cplx<Real> *tw;
    LinePlan lp;
    int64_t lines;      // number of lines
    int64_t line_pitch; // distance between consecutive lines
    int lpw;            // lines per workgroup
    int n_in, n_out;
    int inverse;
    int iters;          // line_pipe_kernel: line groups per workgroup
};

template <typename Real, int N = 0>
__global__ void __launch_bounds__(kMaxBlock) line_kernel(LineArgs<Real> a) {
    using L = LineOps<N>;
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int n = L::n(a.lp), ls = n + kPad, T = L::threads(a.lp);
    // line j = tid / T of the workgroup belongs to its T lanes (blockDim.x = lpw T)
    const int j = threadIdx.x / T, t = threadIdx.x - j * T;
    const int64_t l = (int64_t)blockIdx.x * a.lpw + j;
    cplx<Real> *line = lds + j * ls;
    cplx<Real> *g = a.data + l * a.line_pitch;
    const bool live = l < a.lines;
    for (int e = t; e < n; e += T) {
        cplx<Real> v = {Real(0), Real(0)};
        if (e < a.n_in && live) v = g[e];
        line[L::slot(a.lp, e)] = v;
    }
    __syncthreads();
    L::template fft<Real>(line, a.lp, t, a.tw, a.inverse != 0);
    if (live)
        for (int e = t; e < a.n_out; e += T) g[e] = line[L::slot(a.lp, e)];
}

// Contiguous line pass for a fixed power-of-two length N: a workgroup holds
// G = max(1, 64 / T) lines at a time (T = N / 16 lanes per line) and walks
// `iters` groups of lines.  With PIPE the global loads of the next group are
// issued into registers before the current group's LDS transform, so memory
// traffic overlaps the FFT; without it (one group per workgroup) the kernel
// holds no prefetch registers and more waves share each SIMD instead.
// Inputs: N / IN_DIV per line (the rest zero); outputs: a.n_out per line.
template <typename Real, int N, int IN_DIV, bool PIPE>
__device__ __forceinline__ void line_pipe_body(const LineArgs<Real> &a) {
    constexpr int T = N / 16;
    constexpr int G = T >= 64 ? 1 : 64 / T;
    constexpr int NIN = N / IN_DIV;
    constexpr int NI = NIN / T;
    static_assert(NIN % T == 0, "whole loads per lane");
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int j = (int)threadIdx.x / T, t = (int)threadIdx.x - j * T;
    cplx<Real> *line = lds + j * (N + kPad);
    const int64_t first = (int64_t)blockIdx.x * (PIPE ? a.iters : 1) * G + j;
    cplx<Real> buf[NI];
    auto fetch = [&](int64_t l) {
        // a line past the end loads line 0 instead: its transform is never
        // stored, and unconditional loads issue together (a load guarded by
        // a select becomes a branch, and the join waits for the loads
        // issued before it)
        const cplx<Real> *g = a.data + (l < a.lines ? l : 0) * a.line_pitch;
#pragma unroll
        for (int e = 0; e < NI; ++e) buf[e] = g[t + e * T];
    };
    fetch(first);
    const int iters = PIPE ? a.iters : 1;
    for (int it = 0; it < iters; ++it) {
        const int64_t l = first + (int64_t)it * G;
        if constexpr (p2_first_from_regs<N>()) {
            // the loaded elements are the first radix-16 butterfly's inputs
            cplx<Real> v[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = e < NI ? buf[e] : cplx<Real>{Real(0), Real(0)};
            if (PIPE && it + 1 < iters) fetch(l + G);
            p2_first_pass_regs<Real, N>(line, t, v, a.tw, a.inverse != 0, 0);
            p2_passes<Real, N, 1, 16>(line, t, a.tw, a.inverse != 0, 0);
        } else {
#pragma unroll
            for (int e = 0; e < NI; ++e) line[p2slot(t + e * T)] = buf[e];
            if (IN_DIV > 1)
#pragma unroll
                for (int e = NI; e < N / T; ++e) line[p2slot(t + e * T)] = cplx<Real>{Real(0), Real(0)};
            __syncthreads();
            if (PIPE && it + 1 < iters) fetch(l + G);
            p2_passes<Real, N, 0, 1>(line, t, a.tw, a.inverse != 0, 0);
        }
        if (l < a.lines) {
            cplx<Real> *g = a.data + l * a.line_pitch;
            if (sizeof(Real) == 4 && ((a.n_out | a.line_pitch) & 1) == 0) {
                // two consecutive outputs per lane in one 16-byte store (an
                // even line pitch keeps every line 16-byte aligned)
                for (int e = 2 * t; e < a.n_out; e += 2 * T) {
                    const cplx<Real> u = line[p2slot(e)], v = line[p2slot(e + 1)];
                    *reinterpret_cast<float4 *>(g + e) = float4{float(u.x), float(u.y), float(v.x), float(v.y)};
                }
            } else {
                for (int e = t; e < a.n_out; e += T) g[e] = line[p2slot(e)];
            }
        }
        if (PIPE) __syncthreads();
    }
}

template <typename Real, int N, int IN_DIV, bool PIPE>
__global__ void __launch_bounds__(256) line_pipe_kernel(LineArgs<Real> a) {
    line_pipe_body<Real, N, IN_DIV, PIPE>(a);
}
template <typename Real, int N, int IN_DIV, bool PIPE>
__global__ void __launch_bounds__(256) line_pipe_multi(const LineArgs<Real> *tab, int zdiv) {
    const LineArgs<Real> a = tab[blockIdx.z / zdiv];
    line_pipe_body<Real, N, IN_DIV, PIPE>(a);
}

// Thin film: per kx, the three components' ky lines (ny of 2ny inputs) are
// transformed, multiplied by the tensor at kz = 0, transformed back and the
// first ny outputs stored.
template <typename Real>
struct YConvArgs {
    cplx<Real> *data;
    const cplx<Real> *tw;
    LinePlan lp;  // length py
    int ny, hx, kpw;  // kx columns per workgroup
    SpecGeom g;
    KernelView<Real> K;
};

// MAXB: kMaxBlock, or 1024 for lines too long for three of them in
// kMaxBlock lanes.
template <typename Real>
__device__ __forceinline__ void yconv_body(const YConvArgs<Real> &a) {
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int n = a.lp.n, ls = n + kPad, T = a.lp.threads;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int kx0 = blockIdx.x * a.kpw;
    const int nl = 3 * a.kpw;  // slot s = c * kpw + j
    for (int idx = tid; idx < nl * n; idx += nthr) {
        const int s = idx / n, e = idx - s * n;
        const int c = s / a.kpw, kx = kx0 + (s - c * a.kpw);
        cplx<Real> v = {Real(0), Real(0)};
        if (e < a.ny && kx < a.hx) v = a.data[a.g.at(c, 0, kx, e)];
        lds[s * ls + lslot(a.lp, e)] = v;
    }
    __syncthreads();
    line_fft(lds + (tid / T) * ls, a.lp, tid % T, a.tw, false);
    for (int idx = tid; idx < a.kpw * n; idx += nthr) {
        const int j = idx / n, e = idx - j * n;
        const int kx = kx0 + j;
        if (kx >= a.hx) continue;
        Real t[6];
        tensor_at(a.K, kx, e, 0, t);
        const int se = lslot(a.lp, e);
        apply_tensor(t, lds[j * ls + se], lds[(a.kpw + j) * ls + se], lds[(2 * a.kpw + j) * ls + se]);
    }
    __syncthreads();
    line_fft(lds + (tid / T) * ls, a.lp, tid % T, a.tw, true);
    for (int idx = tid; idx < nl * a.ny; idx += nthr) {
        const int s = idx / a.ny, e = idx - s * a.ny;
        const int c = s / a.kpw, kx = kx0 + (s - c * a.kpw);
        if (kx < a.hx) a.data[a.g.at(c, 0, kx, e)] = lds[s * ls + lslot(a.lp, e)];
    }
}

template <typename Real, int MAXB>
__global__ void __launch_bounds__(MAXB) yconv_kernel(YConvArgs<Real> a) {
    yconv_body<Real>(a);
}
template <typename Real, int MAXB>
__global__ void __launch_bounds__(MAXB) yconv_multi(const YConvArgs<Real> *tab, int zdiv) {
    const YConvArgs<Real> a = tab[blockIdx.z / zdiv];
    yconv_body<Real>(a);
}

// Thin film, split form: the three components' ky spectra (all py entries,
// rows `rows` apart) are multiplied by the tensor at kz = 0 in place; the y
// transforms run before and after as line passes.  For films whose three
// lines do not fit one workgroup.
template <typename Real>
__global__ void __launch_bounds__(256) film_tensor_kernel(cplx<Real> *data, KernelView<Real> K, int hx, int py,
                                                          int64_t rows) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)hx * py) return;
    const int kx = (int)(i / py), ky = (int)(i - (int64_t)kx * py);
    Real t[6];
    tensor_at(K, kx, ky, 0, t);
    const int64_t cs = (int64_t)hx * rows, o = (int64_t)kx * rows + ky;
    apply_tensor(t, data[o], data[cs + o], data[2 * cs + o]);
}

// ---------------------------------------------------------------------------
// z passes: one lane per (kx, ky) column
// ---------------------------------------------------------------------------
template <typename Real>
struct ZArgs {
    cplx<Real> *data;
    const cplx<Real> *tw;  // roots of length pz
    int nz, hx, py;
    int64_t columns;       // hx * pitch (pad columns ky >= py are skipped)
    int64_t zstride;       // hx * pitch (3-D demag layout)
    int64_t comp_stride;   // nz * hx * pitch
    int pitch;             // distance between kx rows of the spectrum (>= py)
    int xcd;               // renumber workgroups per XCD
    KernelView<Real> K;
};

// DFT of length N on registers with roots w_N^t = w_P^(S t) taken from the
// length-P = S N table (power-of-two N uses exact constants).
template <typename Real, int N, int S>
__device__ __forceinline__ void strided_dft(cplx<Real> *v, const cplx<Real> *tw_p, bool inv) {
    if constexpr (N == 1) {
        (void)v; (void)tw_p; (void)inv;
    } else if constexpr (N == 2 || N == 4 || N == 8 || N == 16) {
        dft_any<Real, N>(v, tw_p, N, inv);
    } else {
        cplx<Real> out[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            cplx<Real> acc = v[0];
#pragma unroll
            for (int r = 1; r < N; ++r) {
                cplx<Real> w = tw_p[S * ((r * q) % N)];
                if (inv) w.y = -w.y;
                acc = cadd(acc, cmul(v[r], w));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = out[q];
    }
}

template <typename Real, int N>
__device__ __forceinline__ void half_dft(cplx<Real> *v, const cplx<Real> *tw_p, bool inv) {
    if constexpr (N == 1) {
        (void)v; (void)tw_p; (void)inv;
    } else if constexpr (N == 2 || N == 4 || N == 8 || N == 16) {
        dft_any<Real, N>(v, tw_p, N, inv);
    } else {
        cplx<Real> out[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            cplx<Real> acc = v[0];
#pragma unroll
            for (int r = 1; r < N; ++r) {
                cplx<Real> w = tw_p[2 * ((r * q) % N)];
                if (inv) w.y = -w.y;
                acc = cadd(acc, cmul(v[r], w));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = out[q];
    }
}

// z convolution with two lanes per (kx, ky) column (pz = 2 NZ): lane parity p
// owns kz = 2k + p.  Forward: X[2k+p] = DFT_NZ(x_z w_pz^(p z))[k] (only the
// NZ non-zero inputs exist); tensor at its own kz; inverse:
// y_z = IDFT_NZ(Y_even)[z] + w_pz^(-z) IDFT_NZ(Y_odd)[z], z < NZ, summed
// across the lane pair with a shuffle.  All three components in registers.
// Consecutive lane pairs take consecutive columns.  Used where the staged
// tile does not apply (the k-space projection fallback).
template <typename Real, int NZ>
__global__ void __launch_bounds__(256) zconv_pair_kernel(ZArgs<Real> a) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = (int)(gid & 1);
    // both lanes of a pair stay for the shuffle
    const int64_t col = gid >> 1;
    bool valid = col < a.columns;
    const int kx = valid ? (int)(col / a.pitch) : 0;
    const int ky = valid ? (int)(col - (int64_t)kx * a.pitch) : 0;
    valid = valid && ky < a.py;
    cplx<Real> v[3][NZ];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            cplx<Real> x = {Real(0), Real(0)};
            if (valid) x = a.data[c * a.comp_stride + z * a.zstride + col];
            if (p && z) x = cmul(x, a.tw[z]);
            v[c][z] = x;
        }
        half_dft<Real, NZ>(v[c], a.tw, false);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        Real t[6];
        tensor_at(a.K, kx, ky, 2 * k + p, t);
        apply_tensor(t, v[0][k], v[1][k], v[2][k]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        half_dft<Real, NZ>(v[c], a.tw, true);
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            cplx<Real> y = v[c][z];
            if (p && z) y = cmul(y, cconj(a.tw[z]));
            const cplx<Real> o = {__shfl_xor(y.x, 1), __shfl_xor(y.y, 1)};
            v[c][z] = cadd(y, o);
        }
    }
    if (!valid) return;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int z = 0; z < NZ; ++z)
            if ((z & 1) == p) a.data[c * a.comp_stride + z * a.zstride + col] = v[c][z];
}

// The two-lane z convolution staged through LDS: a workgroup of 256 lanes
// owns 128 consecutive columns.  It loads the tile's 3 NZ rows of 128
// complex values (one row = one (component, z) plane segment) with 16-byte
// loads, all issued before any arithmetic; lane pair j computes column j
// from LDS exactly as zconv_pair_kernel does; the results go back through
// LDS and out with 16-byte stores.  The tile's whole input is in flight at
// once and no lane keeps loaded data in registers across the transform.
constexpr int kZTile = 128;

// Tile of kZTile consecutive columns starting at col0: 3 NZ rows, each row
// one contiguous run of the spectrum.  Lane t moves 16-byte units t + 256 i.
template <typename Real, int NZ, int TW = kZTile>
struct ZTile {
    using C = cplx<Real>;
    static constexpr int ROWS = 3 * NZ;
    static constexpr int V = sizeof(Real) == 4 ? 2 : 1;  // complex values per 16-byte access
    static constexpr int UNITS = TW / V;                 // accesses per row
    static constexpr int PER = (ROWS * UNITS + 255) / 256;
    C v[PER][V];

    __device__ __forceinline__ void fetch(const ZArgs<Real> &a, int64_t col0, int t) {
        if (col0 + TW <= a.columns) {
            // the whole tile lies inside the spectrum (every tile but possibly
            // the last): unconditional 16-byte loads, all issued before any
            // wait; a lane's unit past the tile re-reads the tile's last unit
            // and put() drops it
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int idx = min(t + 256 * i, ROWS * UNITS - 1);
                const int r = idx / UNITS, u = idx - r * UNITS;
                const int c = r / NZ, z = r - c * NZ;
                const C *src = a.data + c * a.comp_stride + z * a.zstride + col0 + (int64_t)u * V;
                if constexpr (V == 2) {
                    const float4 q = *reinterpret_cast<const float4 *>(src);
                    v[i][0] = C{Real(q.x), Real(q.y)};
                    v[i][1] = C{Real(q.z), Real(q.w)};
                } else {
                    v[i][0] = *src;
                }
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = t + 256 * i;
#pragma unroll
            for (int e = 0; e < V; ++e) v[i][e] = C{Real(0), Real(0)};
            if (idx < ROWS * UNITS) {
                const int r = idx / UNITS, u = idx - r * UNITS;
                const int c = r / NZ, z = r - c * NZ;
                const int64_t col = col0 + (int64_t)u * V;
                const C *src = a.data + c * a.comp_stride + z * a.zstride + col;
                if (col + V <= a.columns) {
                    if constexpr (V == 2) {
                        const float4 q = *reinterpret_cast<const float4 *>(src);
                        v[i][0] = C{Real(q.x), Real(q.y)};
                        v[i][1] = C{Real(q.z), Real(q.w)};
                    } else {
                        v[i][0] = *src;
                    }
                } else {
                    for (int e = 0; e < V; ++e)
                        if (col + e < a.columns) v[i][e] = src[e];
                }
            }
        }
    }
    __device__ __forceinline__ void put(C (*tile)[TW + V], int t) const {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int idx = t + 256 * i;
            if (idx < ROWS * UNITS) {
                const int r = idx / UNITS, u = idx - r * UNITS;
#pragma unroll
                for (int e = 0; e < V; ++e) tile[r][u * V + e] = v[i][e];
            }
        }
    }
};

template <typename Real, int NZ, int TW = kZTile>
__device__ __forceinline__ void ztile_store(const ZArgs<Real> &a, cplx<Real> (*tile)[TW + ZTile<Real, NZ, TW>::V],
                                            int64_t col0, int t) {
    using C = cplx<Real>;
    using TL = ZTile<Real, NZ, TW>;
    constexpr int V = TL::V, UNITS = TL::UNITS;
#pragma unroll
    for (int i = 0; i < TL::PER; ++i) {
        const int idx = t + 256 * i;
        if (idx < TL::ROWS * UNITS) {
            const int r = idx / UNITS, u = idx - r * UNITS;
            const int c = r / NZ, z = r - c * NZ;
            const int64_t cl = col0 + (int64_t)u * V;
            C *dst = a.data + c * a.comp_stride + z * a.zstride + cl;
            if (cl + V <= a.columns) {
                if constexpr (V == 2) {
                    const C e0 = tile[r][u * 2], e1 = tile[r][u * 2 + 1];
                    *reinterpret_cast<float4 *>(dst) = float4{float(e0.x), float(e0.y), float(e1.x), float(e1.y)};
                } else {
                    *dst = tile[r][u];
                }
            } else {
                for (int e = 0; e < V; ++e)
                    if (cl + e < a.columns) dst[e] = tile[r][u * V + e];
            }
        }
    }
}

// The lane's half of its column's octant tensor block (the entries of its
// kz parity, six components each), in registers.  Issued at the start of the
// tile kernel together with the tile's loads, so one memory wait covers both.
// Pad columns and the projection fallback keep zeros.
// FP64 keeps only the block's address (30 doubles would double the
// registers) and reads each entry where it is applied.
template <typename Real, int NZ>
struct ZTensorRegs {
    static constexpr int NE = NZ / 2 + 1, NO = (NZ + 1) / 2;
    static constexpr bool REGS = sizeof(Real) == 4;
    static constexpr int NV = (6 * NE + 3) / 4;  // 16-byte loads of the longer (even) half
    Real k[REGS ? NV * 4 : 1];
    const Real *blk = nullptr;
    bool valid = false;
    __device__ __forceinline__ void load(const ZArgs<Real> &a, bool col_valid, int kx, int ky, int p) {
        valid = col_valid && a.K.k;
        const int oy = ky <= a.K.ny ? ky : a.py - ky;
        blk = valid ? a.K.k + ((int64_t)kx * (a.K.ny + 1) + oy) * a.K.kb + (p ? a.K.kodd : 0) : a.K.k;
        if constexpr (REGS) {
#pragma unroll
            for (int e = 0; e < NV * 4; ++e) k[e] = Real(0);
            if (!valid) return;
            const float4 *src = reinterpret_cast<const float4 *>(blk);
            const int nv = p ? (6 * NO + 3) / 4 : NV;
#pragma unroll
            for (int e = 0; e < NV; ++e) {
                if (e < nv) {
                    const float4 q = src[e];
                    k[4 * e] = q.x;
                    k[4 * e + 1] = q.y;
                    k[4 * e + 2] = q.z;
                    k[4 * e + 3] = q.w;
                }
            }
        }
    }
};

// The register convolution of one column of the tile (two lanes per column,
// as zconv_pair_kernel), in place in LDS: lane pair t/2 takes column t/2.  A
// column is read and written only by its own lane pair, which sits in one
// wave, so no barrier separates the reads from the writes.
template <typename Real, int NZ, int TW = kZTile>
__device__ __forceinline__ void ztile_convolve(const ZArgs<Real> &a, cplx<Real> (*tile)[TW + ZTile<Real, NZ, TW>::V],
                                               int j, int p, int ky, const ZTensorRegs<Real, NZ> &kr) {
    using C = cplx<Real>;
    C v[3][NZ];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            C x = tile[c * NZ + z][j];
            if (p && z) x = cmul(x, a.tw[z]);
            v[c][z] = x;
        }
        half_dft<Real, NZ>(v[c], a.tw, false);
    }
    const Real sy = ky <= a.K.ny ? Real(1) : Real(-1);
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        // lane parity p owns kz = 2k + p; its octant entry is (kz or pz - kz) / 2
        const int kz0 = 2 * k, kz1 = 2 * k + 1;
        const int i0 = (kz0 <= NZ ? kz0 : 2 * NZ - kz0) >> 1, i1 = (kz1 <= NZ ? kz1 : 2 * NZ - kz1) >> 1;
        const Real s0 = kz0 <= NZ ? Real(1) : Real(-1), s1 = kz1 <= NZ ? Real(1) : Real(-1);
        const Real sz = p ? s1 : s0;
        Real tt[6];
        if constexpr (sizeof(Real) == 4) {
            // select between two fixed registers with a bit mask (an index
            // chosen at run time would move the array to scratch)
            const uint32_t msk = 0u - (uint32_t)p;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const uint32_t b0 = __float_as_uint(kr.k[6 * i0 + q]), b1 = __float_as_uint(kr.k[6 * i1 + q]);
                tt[q] = __uint_as_float((b0 & ~msk) | (b1 & msk));
            }
        } else {
            const Real *e = kr.blk + 6 * (p ? i1 : i0);
#pragma unroll
            for (int q = 0; q < 6; ++q) tt[q] = kr.valid ? e[q] : Real(0);
        }
        tt[3] = sy * tt[3];
        tt[4] = sz * tt[4];
        tt[5] = sy * sz * tt[5];
        apply_tensor(tt, v[0][k], v[1][k], v[2][k]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        half_dft<Real, NZ>(v[c], a.tw, true);
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            C y = v[c][z];
            if (p && z) y = cmul(y, cconj(a.tw[z]));
            const C o = {__shfl_xor(y.x, 1), __shfl_xor(y.y, 1)};
            if ((z & 1) == p) tile[c * NZ + z][j] = cadd(y, o);
        }
    }
}

template <typename Real, int NZ, int TW>
__device__ __forceinline__ void zconv_tile_body(const ZArgs<Real> &a) {
    static_assert(TW == 128, "one column per lane pair of a 256-lane workgroup");
    using TL = ZTile<Real, NZ, TW>;
    __shared__ cplx<Real> tile[TL::ROWS][TW + TL::V];
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int64_t col0 = (int64_t)blk * TW;
    const int t = (int)threadIdx.x;
    const int j = t >> 1, p = t & 1;
    const int64_t col = col0 + j;
    bool valid = col < a.columns;
    const int kx = valid ? (int)(col / a.pitch) : 0;
    const int ky = valid ? (int)(col - (int64_t)kx * a.pitch) : 0;
    valid = valid && ky < a.py;  // pad columns: transformed with a zero tensor, never read
    ZTensorRegs<Real, NZ> kr;
    kr.load(a, valid, kx, ky, p);
    {
        TL in;
        in.fetch(a, col0, t);
        in.put(tile, t);
    }
    __syncthreads();
    ztile_convolve<Real, NZ, TW>(a, tile, j, p, ky, kr);
    __syncthreads();
    ztile_store<Real, NZ, TW>(a, tile, col0, t);
}

template <typename Real, int NZ, int TW = kZTile>
__global__ void __launch_bounds__(256) zconv_tile_kernel(ZArgs<Real> a) {
    zconv_tile_body<Real, NZ, TW>(a);
}
template <typename Real, int NZ, int TW = kZTile>
__global__ void __launch_bounds__(256) zconv_tile_multi(const ZArgs<Real> *tab, int zdiv) {
    const ZArgs<Real> a = tab[blockIdx.z / zdiv];
    zconv_tile_body<Real, NZ, TW>(a);
}

// z convolution with four lanes per (kx, ky) column (pz = 2 NZ = 4 M, NZ
// even): lane q owns kz = 4k + q.  Forward: X[4k+q] = DFT_M(u)[k] with the
// folded input u_z = x_z w^(q z) + x_(z+M) w^(q (z+M)), z < M; tensor at its
// own kz; inverse: y_z = sum_q w^(-q z) IDFT_M(Y_q)[z mod M], summed over the
// quad with two shuffles; lane q stores the z = q (mod 4).  Workgroups are
// renumbered per XCD so the columns ky and py - ky of one kx row, which read
// the same tensor octant entries, share an L2.
template <typename Real, int NZ>
__device__ __forceinline__ void zconv_quad_body(const ZArgs<Real> &a) {
    static_assert(NZ % 2 == 0, "four lanes per column need an even nz");
    constexpr int M = NZ / 2, PZ = 2 * NZ;
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int64_t gid = (int64_t)blk * blockDim.x + threadIdx.x;
    const int64_t col = gid >> 2;
    const int q = (int)(gid & 3);
    bool valid = col < a.columns;  // all four lanes of a quad stay for the shuffles
    const int kx = valid ? (int)(col / a.pitch) : 0;
    const int ky = valid ? (int)(col - (int64_t)kx * a.pitch) : 0;
    valid = valid && ky < a.py;
    // w^(q z), z < NZ: one gather per z, reused by all components and the inverse
    cplx<Real> wq[NZ];
#pragma unroll
    for (int z = 0; z < NZ; ++z) wq[z] = a.tw[(q * z) % PZ];
    cplx<Real> v[3][M];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int z = 0; z < M; ++z) {
            cplx<Real> lo = {Real(0), Real(0)}, hi = {Real(0), Real(0)};
            if (valid) {
                lo = a.data[c * a.comp_stride + z * a.zstride + col];
                hi = a.data[c * a.comp_stride + (z + M) * a.zstride + col];
            }
            v[c][z] = cadd(cmul(lo, wq[z]), cmul(hi, wq[z + M]));
        }
        strided_dft<Real, M, 4>(v[c], a.tw, false);
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
        Real t[6];
        tensor_at(a.K, kx, ky, 4 * k + q, t);
        apply_tensor(t, v[0][k], v[1][k], v[2][k]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        strided_dft<Real, M, 4>(v[c], a.tw, true);
        cplx<Real> y[NZ];
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            cplx<Real> s = cmul(v[c][z % M], cconj(wq[z]));
            s.x += __shfl_xor(s.x, 1);
            s.y += __shfl_xor(s.y, 1);
            s.x += __shfl_xor(s.x, 2);
            s.y += __shfl_xor(s.y, 2);
            y[z] = s;
        }
        if (valid)
#pragma unroll
            for (int z = 0; z < NZ; ++z)
                if ((z & 3) == q) a.data[c * a.comp_stride + z * a.zstride + col] = y[z];
    }
}

template <typename Real, int NZ>
__global__ void __launch_bounds__(256) zconv_quad_kernel(ZArgs<Real> a) {
    zconv_quad_body<Real, NZ>(a);
}
template <typename Real, int NZ>
__global__ void __launch_bounds__(256) zconv_quad_multi(const ZArgs<Real> *tab, int zdiv) {
    const ZArgs<Real> a = tab[blockIdx.z / zdiv];
    zconv_quad_body<Real, NZ>(a);
}

// Long z lines: a workgroup holds `tile` consecutive columns (contiguous in
// ky) of all three components in LDS.  conv = 0: plain forward (or, with
// `inverse`, inverse) transform of one component (blockIdx.y = component),
// used by the Newell setup and the multilayer stack.
template <typename Real>
struct ZLdsArgs {
    cplx<Real> *data;
    const cplx<Real> *tw;
    LinePlan lp;      // length pz
    int n_in, n_out;
    int tile, conv;
    int inverse;  // conv = 0: the inverse transform (n_in = pz, n_out <= pz)
    int64_t columns, zstride, comp_stride;
    int py, pitch;
    KernelView<Real> K;
};

template <typename Real, int N = 0>
__global__ void __launch_bounds__(kMaxBlock) zlds_kernel(ZLdsArgs<Real> a) {
    using L = LineOps<N>;
    extern __shared__ unsigned char smem_raw[];
    cplx<Real> *lds = reinterpret_cast<cplx<Real> *>(smem_raw);
    const int n = L::n(a.lp), ls = n + kPad, T = L::threads(a.lp);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int ncomp = a.conv ? 3 : 1;
    const int cbase = a.conv ? 0 : blockIdx.y;
    const int nslots = ncomp * a.tile;
    const int64_t col0 = (int64_t)blockIdx.x * a.tile;
    // batches of U loads per lane in flight before their LDS writes
    constexpr int U = 8;
    for (int base = tid; base < nslots * n; base += U * nthr) {
        cplx<Real> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = base + u * nthr;
            const int s = idx % nslots, e = idx / nslots;
            const int c = s / a.tile;
            const int64_t col = col0 + (s - c * a.tile);
            v[u] = {Real(0), Real(0)};
            if (idx < nslots * n && e < a.n_in && col < a.columns)
                v[u] = a.data[(cbase + c) * a.comp_stride + e * a.zstride + col];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int idx = base + u * nthr;
            if (idx < nslots * n) {
                const int s = idx % nslots, e = idx / nslots;
                lds[s * ls + L::slot(a.lp, e)] = v[u];
            }
        }
    }
    __syncthreads();
    L::template fft<Real>(lds + (tid / T) * ls, a.lp, tid % T, a.tw, !a.conv && a.inverse);
    __syncthreads();  // the tensor product and the stores cross lines
    if (a.conv) {
        // one pass per octant entry oz <= nz, applied at kz = oz and at its
        // mirror kz = pz - oz (xz and yz change sign): each tensor entry is
        // read once, as three 2-vector loads
        using V2 = typename std::conditional<sizeof(Real) == 4, float2, double2>::type;
        const int nzo = a.K.nz + 1;
        if (!a.K.k) {
            // k-space projection fallback: the tensor is computed per kz
            for (int idx = tid; idx < a.tile * n; idx += nthr) {
                const int j = idx % a.tile, e = idx / a.tile;
                const int64_t col = col0 + j;
                if (col >= a.columns) continue;
                const int kx = (int)(col / a.pitch), ky = (int)(col - (int64_t)kx * a.pitch);
                if (ky >= a.py) continue;
                Real t[6];
                tensor_at(a.K, kx, ky, e, t);
                const int se = L::slot(a.lp, e);
                apply_tensor(t, lds[j * ls + se], lds[(a.tile + j) * ls + se], lds[(2 * a.tile + j) * ls + se]);
            }
        }
        for (int idx = tid; a.K.k && idx < a.tile * nzo; idx += nthr) {
            const int j = idx % a.tile, oz = idx / a.tile;
            const int64_t col = col0 + j;
            if (col >= a.columns) continue;
            const int kx = (int)(col / a.pitch), ky = (int)(col - (int64_t)kx * a.pitch);
            if (ky >= a.py) continue;
            const bool ylo = ky <= a.K.ny;
            const int oy = ylo ? ky : a.py - ky;
            const Real sy = ylo ? Real(1) : Real(-1);
            const V2 *e = reinterpret_cast<const V2 *>(a.K.k +
                                                       kern_at(a.K.kb, a.K.kodd, a.K.kcomp, a.K.ny + 1, 0, oz, kx, oy));
            const V2 t01 = e[0], t23 = e[1], t45 = e[2];
            Real t[6] = {t01.x, t01.y, t23.x, sy * t23.y, t45.x, sy * t45.y};
            int se = L::slot(a.lp, oz);
            apply_tensor(t, lds[j * ls + se], lds[(a.tile + j) * ls + se], lds[(2 * a.tile + j) * ls + se]);
            if (oz > 0 && oz < a.K.pz - oz) {
                t[4] = -t[4];
                t[5] = -t[5];
                se = L::slot(a.lp, a.K.pz - oz);
                apply_tensor(t, lds[j * ls + se], lds[(a.tile + j) * ls + se], lds[(2 * a.tile + j) * ls + se]);
            }
        }
        __syncthreads();
        L::template fft<Real>(lds + (tid / T) * ls, a.lp, tid % T, a.tw, true);
        __syncthreads();
    }
    for (int idx = tid; idx < nslots * a.n_out; idx += nthr) {
        const int s = idx % nslots, e = idx / nslots;
        const int c = s / a.tile;
        const int64_t col = col0 + (s - c * a.tile);
        if (col < a.columns) a.data[(cbase + c) * a.comp_stride + e * a.zstride + col] = lds[s * ls + L::slot(a.lp, e)];
    }
}

// Octant extraction after the setup transform of layout spec[c][z][kx][ky]
// (z < pz, ky < py): entry (c, oz, kx, oy) of the tensor = Re spec[c][oz][kx][oy],
// stored at kern_at.  `total` counts kcomp * kz_n * hx * (ny + 1) entries.
template <typename Real>
__global__ void extract_octant_kernel(const cplx<double> *spec, Real *kern, int kz_n, int hx, int ny, int pz,
                                      int py, int64_t total, int kblock, int kodd, int kcomp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t ly = ny + 1;
    const int64_t per_comp = (int64_t)kz_n * hx * ly;
    const int c = (int)(i / per_comp);
    int64_t r = i - (int64_t)c * per_comp;
    const int oz = (int)(r / ((int64_t)hx * ly));
    r -= (int64_t)oz * hx * ly;
    const int kx = (int)(r / ly), oy = (int)(r - (int64_t)kx * ly);
    kern[kern_at(kblock, kodd, kcomp, ly, c, oz, kx, oy)] = Real(spec[(((int64_t)c * pz + oz) * hx + kx) * py + oy].x);
}

// ---------------------------------------------------------------------------
// Launch geometry
// ---------------------------------------------------------------------------
size_t lds_budget() { return 65536; }

// Rows per x workgroup: the largest power of two <= 16 whose packed lines fit
// in 64 KB of LDS (padded stride when that still fits) and 512 lanes; a line
// longer than 64 KB takes one packed line (two rows) per workgroup when it
// fits in `limit`, the device's LDS per workgroup (160 KB on gfx950).
bool pick_x(const LinePlan &lp, size_t ce, int &xr, int &ls, int max_rows = 16, size_t limit = 65536) {
    for (xr = max_rows; xr >= 2; xr >>= 1) {
        const int lines = xr / 2;
        if (lines * lp.threads > kMaxBlock) continue;
        for (int pad = kPad; pad >= 0; --pad)
            if ((size_t)lines * (lp.n + pad) * ce <= 65536) {
                ls = lp.n + pad;
                return true;
            }
    }
    xr = 2;
    if (lp.threads > kMaxBlock) return false;
    for (int pad = kPad; pad >= 0; --pad)
        if ((size_t)(lp.n + pad) * ce <= limit) {
            ls = lp.n + pad;
            return true;
        }
    return false;
}

int ilog2(int v) {
    int r = 0;
    while ((1 << r) < v) ++r;
    return r;
}

int pick_lpw(const LinePlan &lp, size_t ce, int nlines_per_unit = 1) {
    int lpw = std::max(1, 256 / std::max(1, lp.threads * nlines_per_unit));
    while (lpw > 1 && ((size_t)lpw * nlines_per_unit * (lp.n + kPad) * ce > lds_budget() ||
                       lpw * nlines_per_unit * lp.threads > kMaxBlock))
        --lpw;
    return lpw;
}

// Columns per long-z workgroup: the largest power of two whose lines fit in
// `budget` bytes of LDS and kMaxBlock lanes.
int pick_ztile(const LinePlan &lp, int ncomp, size_t ce, size_t budget = 65536) {
    int tile = 64;
    while (tile > 1 && ((size_t)ncomp * tile * (lp.n + kPad) * ce > budget || ncomp * tile * lp.threads > kMaxBlock))
        tile >>= 1;
    return tile;
}

// LDS bytes one workgroup may allocate on `device`: lines longer than 64 KB
// use the device's per-workgroup limit (160 KB on gfx950).
size_t device_lds_limit(int device) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
        (size_t)lds > 65536)
        return (size_t)lds;
    return 65536;
}

template <typename Real>
KernelView<Real> kview(const Context &c) {
    const DemagPlan &dp = c.dp;
    KernelView<Real> K;
    K.k = dp.projection ? nullptr : static_cast<const Real *>(dp.kern);
    K.kcomp = dp.kcomp;
    K.kb = dp.kblock;
    K.kodd = dp.kodd;
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

// Calls f(std::integral_constant<int, N>) with N the line length when `lp`
// has a compile-time schedule (power of two, 32..4096, default radices),
// else with N = 0 (runtime plan).
template <class F>
void with_line(const LinePlan &lp, bool allow_fixed, F &&f) {
    if (allow_fixed && p2_matches(lp, lp.n)) {
        switch (lp.n) {
            case 32: f(std::integral_constant<int, 32>{}); return;
            case 64: f(std::integral_constant<int, 64>{}); return;
            case 128: f(std::integral_constant<int, 128>{}); return;
            case 256: f(std::integral_constant<int, 256>{}); return;
            case 512: f(std::integral_constant<int, 512>{}); return;
            case 1024: f(std::integral_constant<int, 1024>{}); return;
            case 2048: f(std::integral_constant<int, 2048>{}); return;
            case 4096: f(std::integral_constant<int, 4096>{}); return;
            default: break;
        }
    }
    f(std::integral_constant<int, 0>{});
}

// One contiguous-line pass (the 3-D y transforms): the pipelined kernel for
// fixed power-of-two schedules with n_in = N or N / 2, else line_kernel.
template <typename Real>
void launch_lines(Context &c, const LinePlan &lp, LineArgs<Real> la, size_t ce) {
    const DemagPlan &dp = c.dp;
    with_line(lp, dp.tune.fixed != 0, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        if constexpr (N >= 32) {
            const int div = la.n_in == N ? 1 : (2 * la.n_in == N ? 2 : 0);
            if (dp.tune.y_pipe > 0 && div > 0) {
                constexpr int T = N / 16, G = T >= 64 ? 1 : 64 / T;
                la.iters = div == 2 && dp.tune.y_pipe_f > 0 ? dp.tune.y_pipe_f : dp.tune.y_pipe;
                const dim3 grid((unsigned)cdiv(la.lines, (int64_t)G * la.iters)), block((unsigned)(G * T));
                const size_t lds = (size_t)G * (N + kPad) * ce;
                if (la.iters == 1) {
                    if (div == 1)
                        launch_rec(c, reinterpret_cast<const void *>(&line_pipe_multi<Real, N, 1, false>), grid, block,
                                   lds, 1, la, [&] {
                                       hipLaunchKernelGGL((line_pipe_kernel<Real, N, 1, false>), grid, block, lds,
                                                          c.stream, la);
                                   });
                    else
                        launch_rec(c, reinterpret_cast<const void *>(&line_pipe_multi<Real, N, 2, false>), grid, block,
                                   lds, 1, la, [&] {
                                       hipLaunchKernelGGL((line_pipe_kernel<Real, N, 2, false>), grid, block, lds,
                                                          c.stream, la);
                                   });
                } else if (div == 1) {
                    launch_rec(c, nullptr, grid, block, lds, 1, la, [&] {
                        hipLaunchKernelGGL((line_pipe_kernel<Real, N, 1, true>), grid, block, lds, c.stream, la);
                    });
                } else {
                    launch_rec(c, nullptr, grid, block, lds, 1, la, [&] {
                        hipLaunchKernelGGL((line_pipe_kernel<Real, N, 2, true>), grid, block, lds, c.stream, la);
                    });
                }
                return;
            }
        }
        const dim3 grid((unsigned)cdiv(la.lines, la.lpw)), block((unsigned)(la.lpw * lp.threads));
        const size_t lds = (size_t)la.lpw * (lp.n + kPad) * ce;
        launch_rec(c, nullptr, grid, block, lds, 1, la,
                   [&] { hipLaunchKernelGGL((line_kernel<Real, N>), grid, block, lds, c.stream, la); });
    });
}

// The volume fraction of a tier backend in the field's precision (FP32
// under mixed precision).
template <typename Real>
const Real *phi_array(const Context &c) {
    if constexpr (sizeof(Real) == 4) return c.d_phi32;
    else return c.d_phi;
}

template <typename Real, typename S>
struct BoundaryCorrArgs {
    Real *out;
    const S *m;
    const Real *phi;  // volume fraction in the field's precision
    const double *ms_cell;
    double ms;
    int64_t n;
    const int32_t *cell;
    const uint32_t *off, *row;
    const int32_t *src;
    const Real *tensor;  // the field's precision (FP32 rows under mixed precision)
    uint32_t groups, stencil;
};

// Sparse near-field demag correction of the boundary tiers
// (native/backends/fdm/src/demag_boundary_fp64.cu:23-76): the field at a
// target cell gains -sum_s dN(t, s) . (phi_s Ms_s m_s) per row, dN symmetric
// with the components Nxx, Nxy, Nxz, Nyy, Nyz, Nzz; sources < 0 are skipped.
// One wavefront per target cell: its lanes share each row's stencil entries
// (entry lane + 64 u, u < 4, per pass), a butterfly reduction gives the row's
// sum in FP64, and the cell's rows are added in plan order, so a cell named
// by several rows gets all of them without atomics.  A pass issues all of a
// lane's loads together: its source indices, then the gathered sources and
// the tensors, with every address clamped to a valid entry; a skipped entry
// subtracts exactly zero.
template <typename Real, typename S>
__global__ void __launch_bounds__(256) boundary_corr_kernel(BoundaryCorrArgs<Real, S> a) {
    constexpr int U = 4;
    const uint32_t g = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (g >= a.groups) return;  // a whole wavefront
    const int64_t ti = a.cell[g], n = a.n;
    double h[3] = {double(a.out[ti]), double(a.out[n + ti]), double(a.out[2 * n + ti])};
    const double *msp = a.ms_cell;
    for (uint32_t e = a.off[g]; e < a.off[g + 1]; ++e) {
        const uint64_t r0 = (uint64_t)a.row[e] * a.stencil;
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (uint32_t s0 = 0; s0 < a.stencil; s0 += U * 64u) {
            int32_t si[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t s = s0 + lane + 64u * u;
                ok[u] = s < a.stencil;
                si[u] = a.src[r0 + (ok[u] ? s : 0u)];
            }
            double w[U], mv[U][3];
            Real nn[U][6];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = ok[u] && si[u] >= 0;
                const int64_t j = ok[u] ? si[u] : 0;
                const double pv = double(a.phi[j]);
                const double cv = *(msp ? msp + j : &kNoArrayF64);
                w[u] = pv * (msp ? cv : a.ms);
#pragma unroll
                for (int q = 0; q < 3; ++q) mv[u][q] = double(a.m[q * n + j]);
                const Real *t = a.tensor + (r0 + (ok[u] ? s0 + lane + 64u * u : 0u)) * 6;
#pragma unroll
                for (int q = 0; q < 6; ++q) nn[u][q] = t[q];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double mx = w[u] * mv[u][0], my = w[u] * mv[u][1], mz = w[u] * mv[u][2];
                const double n0 = nn[u][0], n1 = nn[u][1], n2 = nn[u][2];
                const double n3 = nn[u][3], n4 = nn[u][4], n5 = nn[u][5];
                cx -= ok[u] ? n0 * mx + n1 * my + n2 * mz : 0.0;
                cy -= ok[u] ? n1 * mx + n3 * my + n4 * mz : 0.0;
                cz -= ok[u] ? n2 * mx + n4 * my + n5 * mz : 0.0;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            cx += __shfl_xor(cx, off);
            cy += __shfl_xor(cy, off);
            cz += __shfl_xor(cz, off);
        }
        h[0] += cx;
        h[1] += cy;
        h[2] += cz;
    }
    if (lane == 0) {
        a.out[ti] = Real(h[0]);
        a.out[n + ti] = Real(h[1]);
        a.out[2 * n + ti] = Real(h[2]);
    }
}

// x R2C of Ms src, the y and z transforms with the tensor product and the
// inverse y and z transforms: everything up to the x C2R, in the spectrum.
template <typename Real, typename S>
void demag_transform_t(Context &c, const S *src, const XArgs<Real> &xa, dim3 xgrid, dim3 xblock, size_t xlds,
                       bool weight_ms) {
    DemagPlan &dp = c.dp;
    cplx<Real> *spec = static_cast<cplx<Real> *>(dp.spec);
    const SpecGeom g{dp.hx, dp.three_d ? dp.nz : 1, dp.rows};
    const size_t ce = sizeof(cplx<Real>);
    {
        const MagSource<Real, S> ms{src,    c.d_mask, weight_ms ? c.d_ms_cell : nullptr, weight_ms ? c.ms : 1.0,
                                    dp.nx,  dp.ny,    dp.nz,
                                    (int64_t)c.n, weight_ms ? phi_array<Real>(c) : nullptr};
        Timed t(c, "demag_x_r2c");
        auto launch = [&](const auto &src_v) {
            using Src = std::decay_t<decltype(src_v)>;
            with_line(dp.lx, dp.tune.fixed != 0, [&](auto NC) {
                constexpr int N = decltype(NC)::value;
                const XFwdPack<Real, Src> pk{xa, src_v, spec};
                launch_rec(c, N > 0 ? reinterpret_cast<const void *>(&xfwd_multi<Real, Src, N>) : nullptr, xgrid,
                           xblock, xlds, 3, pk, [&] {
                               hipLaunchKernelGGL((xfwd_kernel<Real, Src, N>), xgrid, xblock, xlds, c.stream, xa,
                                                  src_v, spec);
                           });
            });
        };
        if (ms.mask || ms.ms_cell || ms.phi)
            launch(MagSource<Real, S, true>{ms.m, ms.mask, ms.ms_cell, ms.ms, ms.nx, ms.ny, ms.nz, ms.n, ms.phi});
        else
            launch(ms);
    }
    const KernelView<Real> K = kview<Real>(c);
    if (dp.three_d) {
        LineArgs<Real> la{};
        la.data = spec;
        la.tw = static_cast<const cplx<Real> *>(dp.tw_y);
        la.lp = dp.ly;
        la.lines = 3ll * dp.nz * dp.hx;
        la.line_pitch = dp.rows;
        la.lpw = dp.y_lines_per_wg;
        {
            la.n_in = dp.ny;
            la.n_out = dp.py;
            la.inverse = 0;
            Timed t(c, "demag_y_fwd");
            launch_lines<Real>(c, dp.ly, la, ce);
        }
        {
            Timed t(c, "demag_z_conv");
            if (dp.pz <= z_reg_max<Real>()) {
                ZArgs<Real> za{};
                za.data = spec;
                za.tw = static_cast<const cplx<Real> *>(dp.tw_z);
                za.nz = dp.nz;
                za.hx = dp.hx;
                za.py = dp.py;
                za.pitch = dp.rows;
                za.columns = (int64_t)dp.hx * dp.rows;
                za.zstride = (int64_t)dp.hx * dp.rows;
                za.comp_stride = (int64_t)dp.nz * dp.hx * dp.rows;
                za.K = K;
                za.xcd = dp.tune.xcd;
                const dim3 zb(256);
                const bool quad = dp.tune.z_quad && (dp.nz == 4 || dp.nz == 6 || dp.nz == 8);
                // the staged tiles read the tensor blocks directly (no projection fallback)
                const bool staged = !quad && !dp.projection;
                auto by_nz = [&](auto launch) {
                    switch (dp.nz) {
                        case 1: launch(std::integral_constant<int, 1>{}); break;
                        case 2: launch(std::integral_constant<int, 2>{}); break;
                        case 3: launch(std::integral_constant<int, 3>{}); break;
                        case 4: launch(std::integral_constant<int, 4>{}); break;
                        case 5: launch(std::integral_constant<int, 5>{}); break;
                        case 6: launch(std::integral_constant<int, 6>{}); break;
                        case 7: launch(std::integral_constant<int, 7>{}); break;
                        default: launch(std::integral_constant<int, 8>{}); break;
                    }
                };
                if (staged) {
                    const dim3 zt((unsigned)cdiv(za.columns, kZTile));
                    by_nz([&](auto NZC) {
                        constexpr int NZ = decltype(NZC)::value;
                        launch_rec(c, reinterpret_cast<const void *>(&zconv_tile_multi<Real, NZ>), zt, zb, 0, 1, za,
                                   [&] { hipLaunchKernelGGL((zconv_tile_kernel<Real, NZ>), zt, zb, 0, c.stream, za); });
                    });
                } else if (quad) {
                    const dim3 zq((unsigned)cdiv(4 * za.columns, 256));
                    auto quad_launch = [&](auto NZC) {
                        constexpr int NZ = decltype(NZC)::value;
                        launch_rec(c, reinterpret_cast<const void *>(&zconv_quad_multi<Real, NZ>), zq, zb, 0, 1, za,
                                   [&] { hipLaunchKernelGGL((zconv_quad_kernel<Real, NZ>), zq, zb, 0, c.stream, za); });
                    };
                    switch (dp.nz) {
                        case 4: quad_launch(std::integral_constant<int, 4>{}); break;
                        case 6: quad_launch(std::integral_constant<int, 6>{}); break;
                        default: quad_launch(std::integral_constant<int, 8>{}); break;
                    }
                } else {
                    const dim3 zg((unsigned)cdiv(2 * za.columns, 256));
                    by_nz([&](auto NZC) {
                        launch_rec(c, nullptr, zg, zb, 0, 1, za, [&] {
                            hipLaunchKernelGGL((zconv_pair_kernel<Real, decltype(NZC)::value>), zg, zb, 0, c.stream,
                                               za);
                        });
                    });
                }
            } else {
                ZLdsArgs<Real> za{};
                za.data = spec;
                za.tw = static_cast<const cplx<Real> *>(dp.tw_z);
                za.lp = dp.lz;
                za.n_in = dp.nz;
                za.n_out = dp.nz;
                za.tile = dp.z_tile;
                za.conv = 1;
                za.columns = (int64_t)dp.hx * dp.rows;
                za.zstride = (int64_t)dp.hx * dp.rows;
                za.comp_stride = (int64_t)dp.nz * dp.hx * dp.rows;
                za.py = dp.py;
                za.pitch = dp.rows;
                za.K = K;
                with_line(dp.lz, dp.tune.fixed != 0, [&](auto NC) {
                    const dim3 g1((unsigned)cdiv(za.columns, za.tile)), b1((unsigned)(3 * za.tile * dp.lz.threads));
                    const size_t l1 = (size_t)3 * za.tile * (dp.lz.n + kPad) * ce;
                    launch_rec(c, nullptr, g1, b1, l1, 1, za, [&] {
                        hipLaunchKernelGGL((zlds_kernel<Real, decltype(NC)::value>), g1, b1, l1, c.stream, za);
                    });
                });
            }
        }
        {
            la.n_in = dp.py;
            la.n_out = dp.ny;
            la.inverse = 1;
            Timed t(c, "demag_y_inv");
            launch_lines<Real>(c, dp.ly, la, ce);
        }
    } else if (dp.y_split) {
        LineArgs<Real> la{};
        la.data = spec;
        la.tw = static_cast<const cplx<Real> *>(dp.tw_y);
        la.lp = dp.ly;
        la.lines = 3ll * dp.hx;
        la.line_pitch = dp.rows;
        la.lpw = dp.y_lines_per_wg;
        {
            la.n_in = dp.ny;
            la.n_out = dp.py;
            la.inverse = 0;
            Timed t(c, "demag_y_fwd");
            launch_lines<Real>(c, dp.ly, la, ce);
        }
        {
            Timed t(c, "demag_y_conv");
            const dim3 g1((unsigned)cdiv((int64_t)dp.hx * dp.py, 256));
            launch_rec(c, nullptr, g1, dim3(256), 0, 1, K, [&] {
                hipLaunchKernelGGL((film_tensor_kernel<Real>), g1, dim3(256), 0, c.stream, spec, K, dp.hx, dp.py,
                                   (int64_t)dp.rows);
            });
        }
        {
            la.n_in = dp.py;
            la.n_out = dp.ny;
            la.inverse = 1;
            Timed t(c, "demag_y_inv");
            launch_lines<Real>(c, dp.ly, la, ce);
        }
    } else {
        YConvArgs<Real> ya{};
        ya.data = spec;
        ya.tw = static_cast<const cplx<Real> *>(dp.tw_y);
        ya.lp = dp.ly;
        ya.ny = dp.ny;
        ya.hx = dp.hx;
        ya.kpw = dp.yconv_kpw;
        ya.g = g;
        ya.K = K;
        Timed t(c, "demag_y_conv");
        const dim3 grid((unsigned)cdiv(dp.hx, ya.kpw)), block((unsigned)(3 * ya.kpw * dp.ly.threads));
        const size_t lds = (size_t)3 * ya.kpw * (dp.ly.n + kPad) * ce;
        if (block.x > (unsigned)kMaxBlock)
            launch_rec(c, reinterpret_cast<const void *>(&yconv_multi<Real, 1024>), grid, block, lds, 1, ya,
                       [&] { hipLaunchKernelGGL((yconv_kernel<Real, 1024>), grid, block, lds, c.stream, ya); });
        else
            launch_rec(c, reinterpret_cast<const void *>(&yconv_multi<Real, kMaxBlock>), grid, block, lds, 1, ya,
                       [&] { hipLaunchKernelGGL((yconv_kernel<Real, kMaxBlock>), grid, block, lds, c.stream, ya); });
    }
}

template <typename Real, typename S>
void demag_eval_t(Context &c, const S *src, Real *out, bool weight_ms, int phase) {
    DemagPlan &dp = c.dp;
    cplx<Real> *spec = static_cast<cplx<Real> *>(dp.spec);
    const SpecGeom g{dp.hx, dp.three_d ? dp.nz : 1, dp.rows};
    const size_t ce = sizeof(cplx<Real>);
    XArgs<Real> xa{};
    xa.tw = static_cast<const cplx<Real> *>(dp.tw_x);
    xa.lp = dp.lx;
    xa.rows = dp.ny;
    xa.n_in = dp.nx;
    xa.n_out = dp.nx;
    xa.scale = Real(dp.inv_npad);
    xa.g = g;
    xa.xr = dp.xr;
    xa.xr_log2 = ilog2(dp.xr);
    xa.ls = dp.x_ls;
    const dim3 xgrid((unsigned)cdiv(dp.ny, dp.xr), (unsigned)dp.nz, 3);
    const dim3 xblock((unsigned)((dp.xr / 2) * dp.lx.threads));
    const size_t xlds = (size_t)(dp.xr / 2) * dp.x_ls * ce;
    if (phase != kDemagFinish) {
        c.spec_pre = false;  // the spectrum is overwritten
        demag_transform_t<Real, S>(c, src, xa, xgrid, xblock, xlds, weight_ms);
    }
    if (phase == kDemagSpectrum) return;
    {
        FieldSink<Real> sink{out, c.d_mask, dp.nx, dp.ny, (int64_t)c.n};
        Timed t(c, "demag_x_c2r");
        with_line(dp.lx, dp.tune.fixed != 0, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            const XInvPack<Real, FieldSink<Real>> pk{xa, spec, sink};
            launch_rec(c, N > 0 ? reinterpret_cast<const void *>(&xinv_multi<Real, FieldSink<Real>, N>) : nullptr,
                       xgrid, xblock, xlds, 3, pk, [&] {
                           hipLaunchKernelGGL((xinv_kernel<Real, FieldSink<Real>, N>), xgrid, xblock, xlds, c.stream,
                                              xa, spec, sink);
                       });
        });
    }
    if (weight_ms && c.corr_groups > 0) {
        BoundaryCorrArgs<Real, S> ba{};
        ba.out = out;
        ba.m = src;
        ba.phi = phi_array<Real>(c);
        ba.ms_cell = c.d_ms_cell;
        ba.ms = c.ms;
        ba.n = (int64_t)c.n;
        ba.cell = c.d_corr_cell;
        ba.off = c.d_corr_off;
        ba.row = c.d_corr_row;
        ba.src = c.d_corr_src;
        ba.tensor = static_cast<const Real *>(c.d_corr_n);
        ba.groups = c.corr_groups;
        ba.stencil = c.corr_stencil;
        Timed t(c, "demag_boundary");
        const dim3 g1((unsigned)cdiv((int64_t)c.corr_groups, 4));  // one wavefront per target cell
        launch_rec(c, nullptr, g1, dim3(256), 0, 1, ba, [&] {
            hipLaunchKernelGGL((boundary_corr_kernel<Real, S>), g1, dim3(256), 0, c.stream, ba);
        });
    }
}

template <typename Real>
bool upload_twiddles(Context &c, int n, void **dst) {
    std::vector<cplx<double>> w = twiddle_table(n);
    const size_t bytes = w.size() * sizeof(cplx<Real>);
    if (!ctx_alloc(c, dst, bytes, "twiddle table")) return false;
    if (sizeof(Real) == sizeof(double)) {
        return ctx_upload(c, *dst, w.data(), bytes, "hipMemcpy(twiddles)", c.last_error);
    }
    std::vector<cplx<float>> wf(w.size());
    for (size_t i = 0; i < w.size(); ++i) wf[i] = {(float)w[i].x, (float)w[i].y};
    return ctx_upload(c, *dst, wf.data(), bytes, "hipMemcpy(twiddles)", c.last_error);
}

bool plan_lines(DemagPlan &dp, std::string &err) {
    const DemagTuning &t = dp.tune;
    if (!plan_line(dp.px, dp.lx, err, std::max(1, t.x_lanes), t.x_radix)) return false;
    if (!plan_line(dp.py, dp.ly, err, std::max(1, t.y_lanes), t.y_radix)) return false;
    if (dp.three_d && !plan_line(dp.pz, dp.lz, err)) return false;
    return true;
}

}  // namespace

void parse_tuning_string(DemagTuning &t, const char *text) {
    if (!text) return;
    std::string s(text);
    size_t pos = 0;
    while (pos < s.size()) {
        size_t end = s.find(',', pos);
        if (end == std::string::npos) end = s.size();
        const std::string kv = s.substr(pos, end - pos);
        const size_t eq = kv.find('=');
        if (eq != std::string::npos) {
            const std::string k = kv.substr(0, eq);
            const int v = std::atoi(kv.c_str() + eq + 1);
            if (k == "x_radix") t.x_radix = v;
            else if (k == "x_lanes") t.x_lanes = v;
            else if (k == "x_rows") t.x_rows = v;
            else if (k == "y_radix") t.y_radix = v;
            else if (k == "y_lanes") t.y_lanes = v;
            else if (k == "y_lpw") t.y_lpw = v;
            else if (k == "fixed") t.fixed = v;
            else if (k == "z_quad") t.z_quad = v;
            else if (k == "xcd") t.xcd = v;
            else if (k == "y_pipe") t.y_pipe = v;
            else if (k == "y_pipe_f") t.y_pipe_f = v;
            else if (k == "z_tile") t.z_tile = v;
            else if (k == "zc") t.zc = v;
            else if (k == "y_split") t.y_split = v;
        }
        pos = end + 1;
    }
}

namespace {

void parse_tuning(DemagTuning &t) { parse_tuning_string(t, std::getenv("FULLMAG_FDM_TUNE")); }

// Padded forward transform of the Newell kernel, FP64, into a temporary
// spectrum spec[6][pz][hx][py].
bool newell_forward(Context &c, const DemagPlan &dp, const std::vector<double> &oct, cplx<double> **spec_out,
                    std::string &err, bool z_even = false) {
    void *d_oct = nullptr, *d_spec = nullptr, *twx = nullptr, *twy = nullptr, *twz = nullptr;
    const int pz = dp.three_d ? dp.pz : 1;
    const SpecGeom g{dp.hx, pz, dp.py};
    const size_t spec_bytes = 6 * (size_t)pz * dp.hx * dp.py * sizeof(cplx<double>);
    const size_t ce = sizeof(cplx<double>);
    bool ok = hip_ok(hipMalloc(&d_oct, oct.size() * sizeof(double)), "hipMalloc(newell octant)", err) &&
              hip_ok(hipMalloc(&d_spec, spec_bytes), "hipMalloc(newell spectrum)", err) &&
              ctx_upload(c, d_oct, oct.data(), oct.size() * sizeof(double), "hipMemcpy(newell octant)", err);
    auto up = [&](int n, void **dst) {
        std::vector<cplx<double>> w = twiddle_table(n);
        return hip_ok(hipMalloc(dst, w.size() * sizeof(cplx<double>)), "hipMalloc(twiddles)", err) &&
               ctx_upload(c, *dst, w.data(), w.size() * sizeof(cplx<double>), "hipMemcpy(twiddles)", err);
    };
    ok = ok && up(dp.px, &twx) && up(dp.py, &twy) && (!dp.three_d || up(dp.pz, &twz));
    int xr = 2, xls = dp.px;
    if (ok && !pick_x(dp.lx, ce, xr, xls, 16, dp.lds_limit)) {
        err = "Newell setup: x line does not fit in LDS";
        ok = false;
    }
    if (ok) {
        cplx<double> *spec = static_cast<cplx<double> *>(d_spec);
        NewellSource src{static_cast<const double *>(d_oct), dp.nx, dp.ny, dp.three_d ? dp.nz : 1, dp.px, dp.py, pz,
                         z_even ? 1 : 0};
        XArgs<double> xa{};
        xa.tw = static_cast<const cplx<double> *>(twx);
        xa.lp = dp.lx;
        xa.rows = dp.py;
        xa.n_in = dp.px;
        xa.n_out = 0;
        xa.scale = 1.0;
        xa.g = g;
        xa.xr = xr;
        xa.xr_log2 = ilog2(xr);
        xa.ls = xls;
        hipLaunchKernelGGL((xfwd_kernel<double, NewellSource>), dim3((unsigned)cdiv(dp.py, xr), pz, 6),
                           dim3((unsigned)((xr / 2) * dp.lx.threads)), (size_t)(xr / 2) * xls * ce, c.stream, xa, src,
                           spec);
        LineArgs<double> la{};
        la.data = spec;
        la.tw = static_cast<const cplx<double> *>(twy);
        la.lp = dp.ly;
        la.lines = 6ll * pz * dp.hx;
        la.line_pitch = dp.py;
        la.lpw = pick_lpw(dp.ly, ce);
        la.n_in = dp.py;
        la.n_out = dp.py;
        la.inverse = 0;
        if ((size_t)la.lpw * (dp.ly.n + kPad) * ce > dp.lds_limit || la.lpw * dp.ly.threads > kMaxBlock) {
            err = "Newell setup: y line does not fit in LDS";
            ok = false;
        }
        if (ok)
            hipLaunchKernelGGL((line_kernel<double>), dim3((unsigned)cdiv(la.lines, la.lpw)),
                               dim3((unsigned)(la.lpw * dp.ly.threads)), (size_t)la.lpw * (dp.ly.n + kPad) * ce,
                               c.stream, la);
        if (ok && dp.three_d) {
            ZLdsArgs<double> za{};
            za.data = spec;
            za.tw = static_cast<const cplx<double> *>(twz);
            za.lp = dp.lz;
            za.n_in = dp.pz;
            za.n_out = dp.pz;
            za.tile = pick_ztile(dp.lz, 1, ce);
            za.conv = 0;
            za.columns = (int64_t)dp.hx * dp.py;
            za.zstride = (int64_t)dp.hx * dp.py;
            za.comp_stride = (int64_t)dp.pz * dp.hx * dp.py;
            za.py = dp.py;
            za.pitch = dp.py;
            if ((size_t)za.tile * (dp.lz.n + kPad) * ce > dp.lds_limit || za.tile * dp.lz.threads > kMaxBlock) {
                err = "Newell setup: z line does not fit in LDS";
                ok = false;
            } else {
                hipLaunchKernelGGL((zlds_kernel<double>), dim3((unsigned)cdiv(za.columns, za.tile), 6),
                                   dim3((unsigned)(za.tile * dp.lz.threads)),
                                   (size_t)za.tile * (dp.lz.n + kPad) * ce, c.stream, za);
            }
        }
        ok = ok && hip_ok(hipStreamSynchronize(c.stream), "Newell setup transform", err);
        ok = ok && hip_ok(hipGetLastError(), "Newell setup launch", err);
    }
    if (d_oct) (void)hipFree(d_oct);
    if (twx) (void)hipFree(twx);
    if (twy) (void)hipFree(twy);
    if (twz) (void)hipFree(twz);
    if (!ok) {
        if (d_spec) (void)hipFree(d_spec);
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
    const int64_t ly = dp.ny + 1;
    kern.assign((size_t)dp.hx * ly * dp.kblock, 0.0);
    // parity of the six components along (x, y, z)
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
                        kern[kern_at(dp.kblock, dp.kodd, dp.kcomp, ly, c, z, x, y)] = re;
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
        return ctx_upload(c, dp.kern, kern.data(), bytes, "hipMemcpy(tensor)", c.last_error);
    std::vector<float> f(kern.size());
    for (size_t i = 0; i < kern.size(); ++i) f[i] = (float)kern[i];
    return ctx_upload(c, dp.kern, f.data(), bytes, "hipMemcpy(tensor)", c.last_error);
}

// Launch geometry of the planned lines (no allocation): x rows per
// workgroup, y lines per workgroup, long-z tile, thin-film y convolution
// form.  `keep_split` >= 0 requires the thin-film form to stay as it is (a
// re-plan may not change the spectrum's row count).
template <typename Real>
bool plan_geometry(DemagPlan &dp, std::string &err, int keep_split = -1) {
    const size_t ce = sizeof(cplx<Real>);
    {
        // tensor blocks: even-oz half, then odd-oz half, each padded to 16 bytes
        const int vec = 16 / (int)sizeof(Real);
        const int ne = dp.three_d ? dp.nz / 2 + 1 : 1, no = dp.three_d ? (dp.nz + 1) / 2 : 0;
        dp.kodd = (int)cdiv(ne * dp.kcomp, vec) * vec;
        dp.kblock = dp.kodd + (int)cdiv(no * dp.kcomp, vec) * vec;
    }
    if (!pick_x(dp.lx, ce, dp.xr, dp.x_ls, dp.tune.x_rows > 0 ? dp.tune.x_rows : 16, dp.lds_limit)) {
        err = "x line FFT does not fit in one workgroup";
        return false;
    }
    if (dp.three_d) {
        dp.y_lines_per_wg = pick_lpw(dp.ly, ce);
        if (dp.tune.y_lpw > 0) dp.y_lines_per_wg = std::min(dp.y_lines_per_wg, dp.tune.y_lpw);
        if ((size_t)dp.y_lines_per_wg * (dp.ly.n + kPad) * ce > dp.lds_limit ||
            dp.y_lines_per_wg * dp.ly.threads > kMaxBlock) {
            err = "y line FFT does not fit in LDS or 512 lanes";
            return false;
        }
        if (dp.pz > z_reg_max<Real>()) {
            // at most 32 KB per workgroup: five or more workgroups share a CU
            // (measured at pz = 128: 209 us with 8 columns against 265 us
            // with 16 at 48 KB)
            dp.z_tile = pick_ztile(dp.lz, 3, ce, 32768);
            if (dp.tune.z_tile > 0) dp.z_tile = std::min(dp.z_tile, dp.tune.z_tile);
            if ((size_t)3 * dp.z_tile * (dp.lz.n + kPad) * ce > dp.lds_limit ||
                3 * dp.z_tile * dp.lz.threads > kMaxBlock) {
                err = "z line FFT does not fit in LDS or 512 lanes";
                return false;
            }
        }
    } else {
        dp.yconv_kpw = pick_lpw(dp.ly, ce, 3);
        const bool split = dp.tune.y_split != 0 || (size_t)3 * dp.yconv_kpw * (dp.ly.n + kPad) * ce > dp.lds_limit ||
                           3 * dp.yconv_kpw * dp.ly.threads > 1024;
        if (keep_split >= 0 && split != (keep_split != 0)) {
            err = "schedule changes the thin-film y convolution form";
            return false;
        }
        dp.y_split = split;
        if (dp.y_split) {
            // the spectrum keeps all py ky entries of every kx row
            dp.rows = dp.py;
            dp.y_lines_per_wg = pick_lpw(dp.ly, ce);
            if ((size_t)dp.y_lines_per_wg * (dp.ly.n + kPad) * ce > dp.lds_limit ||
                dp.y_lines_per_wg * dp.ly.threads > kMaxBlock) {
                err = "thin-film y line FFT does not fit in LDS or 512 lanes";
                return false;
            }
        }
    }
    return true;
}

template <typename Real>
bool setup_t(Context &c, const fullmag_fdm_plan_desc &plan) {
    DemagPlan &dp = c.dp;
    const size_t ce = sizeof(cplx<Real>);
    if (!plan_geometry<Real>(dp, c.last_error)) return false;
    if (!upload_twiddles<Real>(c, dp.px, &dp.tw_x) || !upload_twiddles<Real>(c, dp.py, &dp.tw_y)) return false;
    if (dp.three_d && !upload_twiddles<Real>(c, dp.pz, &dp.tw_z)) return false;
    dp.spec_elems = 3ull * (dp.three_d ? dp.nz : 1) * dp.hx * dp.rows;
    if (!ctx_alloc(c, &dp.spec, dp.spec_elems * ce, "demag spectrum")) return false;
    if (dp.projection) return true;
    if (plan.demag_kernel_spectrum_len != 0) {
        std::vector<double> kern;
        if (!host_octant_from_spectra(dp, plan, kern, c.last_error)) return false;
        return upload_kern<Real>(c, kern);
    }
    // In-library Newell build: octant on the host, padded transform on the device.
    std::vector<double> oct;
    newell_octant(dp.nx, dp.ny, dp.three_d ? dp.nz : 1, c.dx, c.dy, c.dz, oct);
    cplx<double> *spec = nullptr;
    if (!newell_forward(c, dp, oct, &spec, c.last_error)) return false;
    dp.kern_elems = (uint64_t)dp.hx * (dp.ny + 1) * dp.kblock;
    const int64_t entries = (int64_t)dp.kcomp * dp.kz_n * dp.hx * (dp.ny + 1);
    bool ok = ctx_alloc(c, &dp.kern, dp.kern_elems * sizeof(Real), "demag tensor") &&
              hip_ok(hipMemsetAsync(dp.kern, 0, dp.kern_elems * sizeof(Real), c.stream), "hipMemsetAsync(tensor)",
                     c.last_error);
    if (ok) {
        const int pz = dp.three_d ? dp.pz : 1;
        hipLaunchKernelGGL((extract_octant_kernel<Real>), dim3((unsigned)cdiv(entries, 256)), dim3(256), 0,
                           c.stream, spec, static_cast<Real *>(dp.kern), dp.kz_n, dp.hx, dp.ny, pz, dp.py, entries,
                           dp.kblock, dp.kodd, dp.kcomp);
        ok = hip_ok(hipStreamSynchronize(c.stream), "extract octant", c.last_error);
    }
    (void)hipFree(spec);
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
    dp.lds_limit = device_lds_limit(c.device);
    parse_tuning(dp.tune);
    // measured at 512x512x8: four lanes per column are faster in FP64 (half
    // the registers per lane), the LDS-staged two-lane tile in FP32
    if (dp.tune.z_quad < 0) dp.tune.z_quad = c.mode == MODE_DOUBLE ? 1 : 0;
    if (!plan_lines(dp, c.last_error)) return false;
    if (c.mode == MODE_DOUBLE) return setup_t<double>(c, plan);
    return setup_t<float>(c, plan);
}

bool demag_replan(Context &c, const DemagTuning &t, std::string &err) {
    if (!c.demag) {
        c.dp.tune = t;
        return true;
    }
    DemagPlan dp = c.dp;
    dp.tune = t;
    if (dp.tune.z_quad < 0) dp.tune.z_quad = c.mode == MODE_DOUBLE ? 1 : 0;
    if (!plan_lines(dp, err)) return false;
    const bool ok = c.mode == MODE_DOUBLE ? plan_geometry<double>(dp, err, c.dp.y_split ? 1 : 0)
                                          : plan_geometry<float>(dp, err, c.dp.y_split ? 1 : 0);
    if (!ok) return false;
    c.dp = dp;  // same buffers: their sizes do not depend on the schedule
    return true;
}

void demag_eval(Context &c, const void *src, bool src_f64, void *out, bool weight_ms, int phase) {
    if (c.mode == MODE_DOUBLE) {
        demag_eval_t<double, double>(c, static_cast<const double *>(src), static_cast<double *>(out), weight_ms, phase);
    } else if (src_f64) {
        demag_eval_t<float, double>(c, static_cast<const double *>(src), static_cast<float *>(out), weight_ms, phase);
    } else {
        demag_eval_t<float, float>(c, static_cast<const float *>(src), static_cast<float *>(out), weight_ms, phase);
    }
}

// ---------------------------------------------------------------------------
// Multilayer stack demag (crates/fullmag-engine/src/multilayer.rs:159-262):
// every layer's Ms m on the common convolution grid is transformed (x R2C,
// y lines, z lines for a 3-D grid; a thin common grid uses the 2-D
// transform, whose padded z plane would only repeat the first), the pair
// kernels are applied point by point in k-space,
//   H_d(k) = -sum_s K(d, s)(k) . M_s(k),
// and every layer's field is transformed back.  Layers on another grid than
// the common one go through the push/pull transfers of
// crates/fullmag-fdm-demag/src/transfer.rs.
// ---------------------------------------------------------------------------
namespace {

template <typename Real>
struct StackTensorArgs {
    const cplx<Real> *m;
    cplx<Real> *h;
    const Real *kern;       // per kernel 6 components of `points` reals
    const int *pair;        // kernel of (dst, src)
    const uint32_t *kimag;  // per kernel: bit c set when component c is imaginary
    int layers;
    int64_t points;        // planes * hx * py, the same linear index in spectra and kernels
    int64_t layer_stride;  // complex elements per layer spectrum
    int64_t comp_stride;   // complex elements per component of a spectrum (= points)
    int64_t kern_stride;   // reals per pair kernel (6 * points)
};

// acc += k m for a purely real (im = false) or purely imaginary (k i) kernel
// entry.
template <typename T>
__device__ __forceinline__ void kmad(cplx<T> &acc, T k, bool im, const cplx<T> &m) {
    if (im) {
        acc.x -= k * m.y;
        acc.y += k * m.x;
    } else {
        acc.x += k * m.x;
        acc.y += k * m.y;
    }
}

// One lane per k-point; for every destination layer the sources are summed
// in layer order (the order of the reference's kernel pairs for that
// destination).  Kernel components: xx, yy, zz, xy, xz, yz.  Every padded
// Newell component has a definite parity along each axis, so its transform
// is purely real (an even number of odd axes) or purely imaginary (odd): one
// real per component is stored.  LT > 0: that many layers, their sources held
// in registers; 0: any count, sources reloaded per destination.
template <typename Real, int LT>
__global__ void __launch_bounds__(256) stack_tensor_kernel(StackTensorArgs<Real> a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.points) return;
    const int64_t cs = a.comp_stride;
    const int L = LT > 0 ? LT : a.layers;
    constexpr int NR = LT > 0 ? LT : 1;
    cplx<Real> mr[NR][3];
    if (LT > 0) {
#pragma unroll
        for (int s = 0; s < NR; ++s) {
            const cplx<Real> *M = a.m + (int64_t)s * a.layer_stride + i;
            mr[s][0] = M[0];
            mr[s][1] = M[cs];
            mr[s][2] = M[2 * cs];
        }
    }
    for (int d = 0; d < L; ++d) {
        cplx<Real> hx{Real(0), Real(0)}, hy{Real(0), Real(0)}, hz{Real(0), Real(0)};
#pragma unroll
        for (int s = 0; s < (LT > 0 ? LT : 1); ++s) {
            for (int s2 = s; s2 < (LT > 0 ? s + 1 : L); ++s2) {
                const int u = a.pair[d * L + s2];
                const uint32_t im = a.kimag[u];
                const Real *K = a.kern + (int64_t)u * a.kern_stride + i;
                cplx<Real> mx, my, mz;
                if (LT > 0) {
                    mx = mr[s][0];
                    my = mr[s][1];
                    mz = mr[s][2];
                } else {
                    const cplx<Real> *M = a.m + (int64_t)s2 * a.layer_stride + i;
                    mx = M[0];
                    my = M[cs];
                    mz = M[2 * cs];
                }
                const Real kxx = K[0], kyy = K[cs], kzz = K[2 * cs], kxy = K[3 * cs], kxz = K[4 * cs], kyz = K[5 * cs];
                const bool ixx = im & 1u, iyy = im & 2u, izz = im & 4u, ixy = im & 8u, ixz = im & 16u, iyz = im & 32u;
                kmad(hx, kxx, ixx, mx); kmad(hx, kxy, ixy, my); kmad(hx, kxz, ixz, mz);
                kmad(hy, kxy, ixy, mx); kmad(hy, kyy, iyy, my); kmad(hy, kyz, iyz, mz);
                kmad(hz, kxz, ixz, mx); kmad(hz, kyz, iyz, my); kmad(hz, kzz, izz, mz);
            }
        }
        cplx<Real> *H = a.h + (int64_t)d * a.layer_stride + i;
        H[0] = {-hx.x, -hx.y};
        H[cs] = {-hy.x, -hy.y};
        H[2 * cs] = {-hz.x, -hz.y};
    }
}

struct GridXfer {
    int nx, ny, nz;     // native grid
    double ndx, ndy, ndz;
    int cx, cy, cz;     // common grid
    double cdx, cdy, cdz;
};

// push_m (transfer.rs:18-121): every common-grid cell is the overlap-volume
// weighted mean of the native cells it overlaps (FP64); SoA in, SoA out.
template <typename S>
__global__ void __launch_bounds__(256) stack_push_kernel(GridXfer g, const S *m, double *out) {
    const int64_t nc = (int64_t)g.cx * g.cy * g.cz, nn = (int64_t)g.nx * g.ny * g.nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= nc) return;
    const int x = (int)(p % g.cx), y = (int)((p / g.cx) % g.cy), z = (int)(p / ((int64_t)g.cx * g.cy));
    const double lo[3] = {x * g.cdx, y * g.cdy, z * g.cdz};
    const double hi[3] = {lo[0] + g.cdx, lo[1] + g.cdy, lo[2] + g.cdz};
    const double nd[3] = {g.ndx, g.ndy, g.ndz};
    const int nmax[3] = {g.nx, g.ny, g.nz};
    int a0[3], a1[3];
    for (int d = 0; d < 3; ++d) {
        a0[d] = max((int)floor(lo[d] / nd[d]), 0);
        a1[d] = min((int)ceil(hi[d] / nd[d]), nmax[d]);
    }
    double tot = 0.0, acc[3] = {0.0, 0.0, 0.0};
    for (int k = a0[2]; k < a1[2]; ++k)
        for (int j = a0[1]; j < a1[1]; ++j)
            for (int i = a0[0]; i < a1[0]; ++i) {
                const double o[3] = {i * g.ndx, j * g.ndy, k * g.ndz};
                double ov = 1.0;
                for (int d = 0; d < 3; ++d) ov *= fmax(fmin(hi[d], o[d] + nd[d]) - fmax(lo[d], o[d]), 0.0);
                if (ov > 0.0) {
                    const int64_t q = ((int64_t)k * g.ny + j) * g.nx + i;
                    for (int c = 0; c < 3; ++c) acc[c] += double(m[c * nn + q]) * ov;
                    tot += ov;
                }
            }
    for (int c = 0; c < 3; ++c) out[c * nc + p] = tot > 0.0 ? acc[c] / tot : 0.0;
}

// pull_h (transfer.rs:124-198): trilinear interpolation of the common grid at
// every native cell centre, neighbour indices clamped; inactive cells zero.
template <typename Real>
__global__ void __launch_bounds__(256) stack_pull_kernel(GridXfer g, const Real *h, const uint8_t *mask, Real *out) {
    const int64_t nc = (int64_t)g.cx * g.cy * g.cz, nn = (int64_t)g.nx * g.ny * g.nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= nn) return;
    const int x = (int)(p % g.nx), y = (int)((p / g.nx) % g.ny), z = (int)(p / ((int64_t)g.nx * g.ny));
    const double f[3] = {(x + 0.5) * g.ndx / g.cdx - 0.5, (y + 0.5) * g.ndy / g.cdy - 0.5,
                         (z + 0.5) * g.ndz / g.cdz - 0.5};
    const int cmax[3] = {g.cx, g.cy, g.cz};
    int b[3];
    double w[3];
    for (int d = 0; d < 3; ++d) {
        b[d] = (int)floor(f[d]);
        w[d] = f[d] - floor(f[d]);
    }
    double r[3] = {0.0, 0.0, 0.0};
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int ix = min(max(b[0] + dx, 0), cmax[0] - 1);
                const int iy = min(max(b[1] + dy, 0), cmax[1] - 1);
                const int iz = min(max(b[2] + dz, 0), cmax[2] - 1);
                const double ww = (dx == 0 ? 1.0 - w[0] : w[0]) * (dy == 0 ? 1.0 - w[1] : w[1]) *
                                  (dz == 0 ? 1.0 - w[2] : w[2]);
                const int64_t q = ((int64_t)iz * g.cy + iy) * g.cx + ix;
                for (int c = 0; c < 3; ++c) r[c] += double(h[c * nc + q]) * ww;
            }
    const bool act = !mask || mask[p];
    for (int c = 0; c < 3; ++c) out[c * nn + p] = act ? Real(r[c]) : Real(0);
}

// One real per kernel component: the imaginary part where bit c of `imag`
// is set, else the real part (the other part is zero up to rounding).
template <typename Real>
__global__ void __launch_bounds__(256) spec_part_kernel(const cplx<double> *in, Real *out, int64_t points,
                                                        uint32_t imag) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 6 * points) return;
    const int c = (int)(i / points);
    out[i] = Real(((imag >> c) & 1u) ? in[i].y : in[i].x);
}

GridXfer xfer_of(const Stack &s, const Context &l) {
    const Context &c = s.conv;
    return GridXfer{(int)l.nx, (int)l.ny, (int)l.nz, l.dx, l.dy, l.dz,
                    (int)c.nx, (int)c.ny, (int)c.nz, c.dx, c.dy, c.dz};
}

template <typename Real>
bool stack_setup_t(Stack &s, std::string &err) {
    Context &c = s.conv;
    DemagPlan &dp = c.dp;
    const size_t ce = sizeof(cplx<Real>);
    const int L = (int)s.layers.size();
    dp = DemagPlan{};
    dp.enabled = true;
    dp.nx = c.nx;
    dp.ny = c.ny;
    dp.nz = c.nz;
    dp.three_d = c.nz > 1;
    dp.px = 2 * c.nx;
    dp.py = 2 * c.ny;
    dp.pz = dp.three_d ? 2 * c.nz : 1;
    dp.hx = c.nx + 1;
    dp.rows = dp.py;
    dp.inv_npad = 1.0 / ((double)dp.px * dp.py * dp.pz);
    dp.lds_limit = device_lds_limit(c.device);
    if (!plan_lines(dp, err)) return false;
    if (!pick_x(dp.lx, ce, dp.xr, dp.x_ls, 16, dp.lds_limit)) {
        err = "stack: x line FFT does not fit in one workgroup";
        return false;
    }
    dp.y_lines_per_wg = pick_lpw(dp.ly, ce);
    if ((size_t)dp.y_lines_per_wg * (dp.ly.n + kPad) * ce > dp.lds_limit ||
        dp.y_lines_per_wg * dp.ly.threads > kMaxBlock) {
        err = "stack: y line FFT does not fit in LDS or 512 lanes";
        return false;
    }
    if (dp.three_d) {
        dp.z_tile = pick_ztile(dp.lz, 1, ce);
        if ((size_t)dp.z_tile * (dp.lz.n + kPad) * ce > dp.lds_limit || dp.z_tile * dp.lz.threads > kMaxBlock) {
            err = "stack: z line FFT does not fit in LDS or 512 lanes";
            return false;
        }
    }
    if (!upload_twiddles<Real>(c, dp.px, &dp.tw_x) || !upload_twiddles<Real>(c, dp.py, &dp.tw_y)) {
        err = c.last_error;
        return false;
    }
    if (dp.three_d && !upload_twiddles<Real>(c, dp.pz, &dp.tw_z)) {
        err = c.last_error;
        return false;
    }
    s.planes = dp.pz;
    const int64_t points = (int64_t)s.planes * dp.hx * dp.py;
    s.layer_spec = 3 * (uint64_t)points;
    if (!ctx_alloc(c, &s.spec_m, (size_t)L * s.layer_spec * ce, "stack spectra") ||
        !ctx_alloc(c, &s.spec_h, (size_t)L * s.layer_spec * ce, "stack field spectra")) {
        err = c.last_error;
        return false;
    }
    // pair kernels: the self kernel once, one shifted kernel per distinct
    // z distance dst - src (compared exactly)
    std::vector<double> shifts;
    s.pair.assign((size_t)L * L, 0);
    for (int d = 0; d < L; ++d)
        for (int q = 0; q < L; ++q) {
            if (d == q) continue;
            const double sh = s.oz[d] - s.oz[q];
            auto it = std::find(shifts.begin(), shifts.end(), sh);
            if (it == shifts.end()) {
                shifts.push_back(sh);
                it = shifts.end() - 1;
            }
            s.pair[(size_t)d * L + q] = 1 + (int)(it - shifts.begin());
        }
    s.nkern = 1 + (int)shifts.size();
    s.kern_stride = 6 * (uint64_t)points;
    // components with an odd number of odd axes transform to imaginary
    // values: xz (odd in x) and yz (odd in y) of the shifted kernels, whose
    // z mirror keeps the sign, and of the thin self kernel (zero there); in
    // the 3-D self kernel they are also odd in z, so every component is real
    std::vector<uint32_t> kimag(s.nkern, (1u << 4) | (1u << 5));
    if (dp.three_d) kimag[0] = 0;
    if (!ctx_alloc(c, &s.kern, (size_t)s.nkern * s.kern_stride * sizeof(Real), "stack pair kernels") ||
        !ctx_alloc(c, (void **)&s.d_pair, s.pair.size() * sizeof(int), "stack pair table") ||
        !ctx_alloc(c, (void **)&s.d_kimag, kimag.size() * sizeof(uint32_t), "stack kernel parities") ||
        !ctx_upload(c, s.d_pair, s.pair.data(), s.pair.size() * sizeof(int), "hipMemcpy(pair table)", err) ||
        !ctx_upload(c, s.d_kimag, kimag.data(), kimag.size() * sizeof(uint32_t), "hipMemcpy(parities)", err))
        return false;
    for (int u = 0; u < s.nkern; ++u) {
        std::vector<double> oct;
        newell_octant(dp.nx, dp.ny, dp.three_d ? dp.nz : 1, c.dx, c.dy, c.dz, oct, u == 0 ? 0.0 : shifts[u - 1]);
        cplx<double> *spec = nullptr;
        if (!newell_forward(c, dp, oct, &spec, err, u > 0)) return false;
        Real *dst = static_cast<Real *>(s.kern) + (uint64_t)u * s.kern_stride;
        hipLaunchKernelGGL((spec_part_kernel<Real>), dim3((unsigned)cdiv(s.kern_stride, 256)), dim3(256), 0,
                           c.stream, spec, dst, (int64_t)points, kimag[u]);
        const bool ok = hip_ok(hipStreamSynchronize(c.stream), "stack kernel", err);
        (void)hipFree(spec);
        if (!ok) return false;
    }
    s.conv_m.assign(L, nullptr);
    s.conv_h.assign(L, nullptr);
    const uint64_t ncv = (uint64_t)c.nx * c.ny * c.nz;
    for (int l = 0; l < L; ++l) {
        const Context &lc = *s.layers[l];
        if (lc.nx == c.nx && lc.ny == c.ny && lc.nz == c.nz) continue;  // transfer.rs identity fast path
        if (!ctx_alloc(c, &s.conv_m[l], 3 * ncv * sizeof(double), "stack transfer m") ||
            !ctx_alloc(c, &s.conv_h[l], 3 * ncv * sizeof(Real), "stack transfer H")) {
            err = c.last_error;
            return false;
        }
    }
    return hip_ok(hipGetLastError(), "stack setup", err);
}

template <typename Real, typename Src>
void stack_xfwd(Context &c, const XArgs<Real> &xa, const Src &src, cplx<Real> *spec) {
    const DemagPlan &dp = c.dp;
    const dim3 grid((unsigned)cdiv(dp.ny, dp.xr), (unsigned)dp.nz, 3);
    const dim3 block((unsigned)((dp.xr / 2) * dp.lx.threads));
    const size_t lds = (size_t)(dp.xr / 2) * dp.x_ls * sizeof(cplx<Real>);
    with_line(dp.lx, true, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        hipLaunchKernelGGL((xfwd_kernel<Real, Src, N>), grid, block, lds, c.stream, xa, src, spec);
    });
}

// y lines (3-D: per component, over the nz stored planes) and z lines of one
// spectrum, forward or inverse.
template <typename Real>
void stack_yz(Stack &s, cplx<Real> *spec, bool inverse) {
    Context &c = s.conv;
    const DemagPlan &dp = c.dp;
    const size_t ce = sizeof(cplx<Real>);
    const int64_t comp = (int64_t)s.planes * dp.hx * dp.py;
    auto zpass = [&] {
        ZLdsArgs<Real> za{};
        za.data = spec;
        za.tw = static_cast<const cplx<Real> *>(dp.tw_z);
        za.lp = dp.lz;
        za.n_in = inverse ? dp.pz : dp.nz;
        za.n_out = inverse ? dp.nz : dp.pz;
        za.tile = dp.z_tile;
        za.conv = 0;
        za.inverse = inverse ? 1 : 0;
        za.columns = (int64_t)dp.hx * dp.py;
        za.zstride = (int64_t)dp.hx * dp.py;
        za.comp_stride = comp;
        za.py = dp.py;
        za.pitch = dp.py;
        with_line(dp.lz, true, [&](auto NC) {
            hipLaunchKernelGGL((zlds_kernel<Real, decltype(NC)::value>), dim3((unsigned)cdiv(za.columns, za.tile), 3),
                               dim3((unsigned)(za.tile * dp.lz.threads)),
                               (size_t)za.tile * (dp.lz.n + kPad) * ce, c.stream, za);
        });
    };
    if (dp.three_d && inverse) zpass();
    LineArgs<Real> la{};
    la.tw = static_cast<const cplx<Real> *>(dp.tw_y);
    la.lp = dp.ly;
    la.line_pitch = dp.py;
    la.lpw = dp.y_lines_per_wg;
    la.n_in = inverse ? dp.py : dp.ny;
    la.n_out = inverse ? dp.ny : dp.py;
    la.inverse = inverse ? 1 : 0;
    if (dp.three_d) {
        la.lines = (int64_t)dp.nz * dp.hx;
        for (int k = 0; k < 3; ++k) {
            la.data = spec + k * comp;
            launch_lines<Real>(c, dp.ly, la, ce);
        }
    } else {
        la.data = spec;
        la.lines = 3ll * dp.hx;
        launch_lines<Real>(c, dp.ly, la, ce);
    }
    if (dp.three_d && !inverse) zpass();
}

template <typename Real, typename S>
void stack_eval_t(Stack &s, const S *const *src) {
    Context &c = s.conv;
    DemagPlan &dp = c.dp;
    const int L = (int)s.layers.size();
    XArgs<Real> xa{};
    xa.tw = static_cast<const cplx<Real> *>(dp.tw_x);
    xa.lp = dp.lx;
    xa.rows = dp.ny;
    xa.n_in = dp.nx;
    xa.n_out = dp.nx;
    xa.scale = Real(dp.inv_npad);
    xa.g = SpecGeom{dp.hx, s.planes, dp.py};
    xa.xr = dp.xr;
    xa.xr_log2 = ilog2(dp.xr);
    xa.ls = dp.x_ls;
    const int64_t ncv = (int64_t)c.nx * c.ny * c.nz;
    cplx<Real> *sm = static_cast<cplx<Real> *>(s.spec_m), *sh = static_cast<cplx<Real> *>(s.spec_h);
    {
        Timed t(c, "stack_forward");
        for (int l = 0; l < L; ++l) {
            Context &lc = *s.layers[l];
            cplx<Real> *spec = sm + (uint64_t)l * s.layer_spec;
            if (s.conv_m[l]) {
                double *cm = static_cast<double *>(s.conv_m[l]);
                hipLaunchKernelGGL((stack_push_kernel<S>), dim3((unsigned)cdiv(ncv, 256)), dim3(256), 0, c.stream,
                                   xfer_of(s, lc), src[l], cm);
                MagSource<Real, double> ms{cm, nullptr, nullptr, lc.ms, dp.nx, dp.ny, dp.nz, ncv, nullptr};
                stack_xfwd<Real>(c, xa, ms, spec);
            } else {
                if (lc.d_mask) {
                    MagSource<Real, S, true> ms{src[l], lc.d_mask, nullptr, lc.ms, dp.nx, dp.ny, dp.nz, ncv, nullptr};
                    stack_xfwd<Real>(c, xa, ms, spec);
                } else {
                    MagSource<Real, S> ms{src[l], nullptr, nullptr, lc.ms, dp.nx, dp.ny, dp.nz, ncv, nullptr};
                    stack_xfwd<Real>(c, xa, ms, spec);
                }
            }
            stack_yz<Real>(s, spec, false);
        }
    }
    {
        Timed t(c, "stack_tensor");
        StackTensorArgs<Real> ta{};
        ta.m = sm;
        ta.h = sh;
        ta.kern = static_cast<const Real *>(s.kern);
        ta.pair = s.d_pair;
        ta.kimag = s.d_kimag;
        ta.layers = L;
        ta.points = (int64_t)s.planes * dp.hx * dp.py;
        ta.layer_stride = (int64_t)s.layer_spec;
        ta.comp_stride = ta.points;
        ta.kern_stride = (int64_t)s.kern_stride;
        const dim3 tg((unsigned)cdiv(ta.points, 256)), tb(256);
        switch (L) {
            case 1: hipLaunchKernelGGL((stack_tensor_kernel<Real, 1>), tg, tb, 0, c.stream, ta); break;
            case 2: hipLaunchKernelGGL((stack_tensor_kernel<Real, 2>), tg, tb, 0, c.stream, ta); break;
            case 3: hipLaunchKernelGGL((stack_tensor_kernel<Real, 3>), tg, tb, 0, c.stream, ta); break;
            case 4: hipLaunchKernelGGL((stack_tensor_kernel<Real, 4>), tg, tb, 0, c.stream, ta); break;
            default: hipLaunchKernelGGL((stack_tensor_kernel<Real, 0>), tg, tb, 0, c.stream, ta); break;
        }
    }
    {
        Timed t(c, "stack_inverse");
        const dim3 xgrid((unsigned)cdiv(dp.ny, dp.xr), (unsigned)dp.nz, 3);
        const dim3 xblock((unsigned)((dp.xr / 2) * dp.lx.threads));
        const size_t xlds = (size_t)(dp.xr / 2) * dp.x_ls * sizeof(cplx<Real>);
        for (int l = 0; l < L; ++l) {
            Context &lc = *s.layers[l];
            cplx<Real> *spec = sh + (uint64_t)l * s.layer_spec;
            stack_yz<Real>(s, spec, true);
            const bool xfer = s.conv_h[l] != nullptr;
            Real *out = xfer ? static_cast<Real *>(s.conv_h[l]) : static_cast<Real *>(lc.hd);
            FieldSink<Real> sink{out, xfer ? nullptr : lc.d_mask, dp.nx, dp.ny, ncv};
            with_line(dp.lx, true, [&](auto NC) {
                constexpr int N = decltype(NC)::value;
                hipLaunchKernelGGL((xinv_kernel<Real, FieldSink<Real>, N>), xgrid, xblock, xlds, c.stream, xa, spec,
                                   sink);
            });
            if (xfer)
                hipLaunchKernelGGL((stack_pull_kernel<Real>), dim3((unsigned)cdiv(lc.n, 256)), dim3(256), 0,
                                   c.stream, xfer_of(s, lc), static_cast<const Real *>(s.conv_h[l]), lc.d_mask,
                                   static_cast<Real *>(lc.hd));
        }
    }
}

}  // namespace

bool stack_demag_setup(Stack &s, std::string &err) {
    if (!s.demag) return true;
    if (s.conv.mode == MODE_DOUBLE) return stack_setup_t<double>(s, err);
    return stack_setup_t<float>(s, err);
}

void stack_demag_eval(Stack &s, const void *const *src, bool src_state) {
    if (!s.demag) return;
    const int mode = s.conv.mode;
    const size_t L = s.layers.size();
    if (mode == MODE_DOUBLE) {
        std::vector<const double *> v(L);
        for (size_t l = 0; l < L; ++l) v[l] = static_cast<const double *>(src[l]);
        stack_eval_t<double, double>(s, v.data());
    } else if (mode == MODE_MIXED && src_state) {
        std::vector<const double *> v(L);
        for (size_t l = 0; l < L; ++l) v[l] = static_cast<const double *>(src[l]);
        stack_eval_t<float, double>(s, v.data());
    } else {
        std::vector<const float *> v(L);
        for (size_t l = 0; l < L; ++l) v[l] = static_cast<const float *>(src[l]);
        stack_eval_t<float, float>(s, v.data());
    }
}

bool newell_spectra_full(uint32_t nx, uint32_t ny, uint32_t nz, double dx, double dy, double dz, bool thin_film_2d,
                         double *out, std::string &err) {
    Context c;
    c.nx = nx; c.ny = ny; c.nz = nz; c.dx = dx; c.dy = dy; c.dz = dz;
    c.mode = MODE_DOUBLE;
    if (!hip_ok(hipGetDevice(&c.device), "hipGetDevice", err)) return false;
    if (!hip_ok(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate", err)) return false;
    DemagPlan dp;
    dp.lds_limit = device_lds_limit(c.device);
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
        const SpecGeom g{dp.hx, pz, dp.py};
        std::vector<cplx<double>> h((size_t)6 * pz * dp.hx * dp.py);
        ok = hip_ok(hipMemcpy(h.data(), spec, h.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost),
                    "hipMemcpy(newell spectrum)", err);
        (void)hipFree(spec);
        if (ok) {
            const int64_t npad = (int64_t)dp.px * dp.py * pz;
            for (int comp = 0; comp < 6; ++comp)
                for (int z = 0; z < pz; ++z)
                    for (int y = 0; y < dp.py; ++y)
                        for (int x = 0; x < dp.px; ++x) {
                            cplx<double> v;
                            if (x < dp.hx) {
                                v = h[g.at(comp, z, x, y)];
                            } else {
                                // Hermitian continuation of a real kernel
                                v = h[g.at(comp, (pz - z) % pz, dp.px - x, (dp.py - y) % dp.py)];
                                v.y = -v.y;
                            }
                            const int64_t i = ((int64_t)z * dp.py + y) * dp.px + x;
                            out[comp * 2 * npad + 2 * i] = v.x;
                            out[comp * 2 * npad + 2 * i + 1] = v.y;
                        }
        }
    }
    (void)hipStreamDestroy(c.stream);
    return ok;
}

}  // namespace fm
