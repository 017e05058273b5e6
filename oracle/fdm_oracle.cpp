// fdm_oracle.cpp -- CPU FP64 oracle for the fullmag FDM LLG time step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
// compared against; it is never linked into, loaded by, or called from the
// product library (fullmag_amd/).  Only tests/, __graft_entry__.smoke() and
// the cpu_baseline leg of bench.py load it.
//
// It restates, in plain C++17, the CPU reference engine of fullmag:
//   crates/fullmag-engine/src/lib.rs        (fields, energies, integrators)
//   crates/fullmag-engine/src/vector.rs     (vector helpers, normalisation)
//   crates/fullmag-fdm-demag/src/newell.rs  (Newell-Williams-Dunlop tensor)
//   crates/fullmag-plan/src/lib.rs          (seeded random m0 generator)
// The reference engine is Rust and uses rustfft 6.4.1 (not vendored, no Rust
// toolchain here), so the FFT below is this file's own mixed-radix Stockham
// FFT with Bluestein for large prime factors; its output is pinned against
// the DFT definition (brute force), numpy.fft fixtures and Parseval in tests.
//
// Pinning status of each part (how a plausible mistake would be caught):
//   newell_f/newell_g/ldia/asymptotic/placement  pinned by the reference's own
//       KATs (newell.rs:676-907): trace=1, N_xx=1/3 for a cube, N_xy(1,1,0),
//       N_xy(2,2,0), N_xy(5,5,0), odd/even symmetry, dipole asymptotics.
//   fft / fft3                                   pinned by brute-force DFT and
//       numpy.fft golden vectors, inverse(forward)=N*identity.
//   exchange field                               pinned by the KATs at
//       lib.rs:3001-3028 ([2,-2,0] and [1,-1,0]) and summation by parts.
//   demag field                                  pinned by the closed form for
//       a single cube (H = -M/3), extended thin film (H_z -> -M_z), E_d >= 0.
//   LLG rhs / Heun / RK4 / RK23 / DP45 / ABM3    pinned by the analytic
//       single-spin Zeeman precession solution (closed form), the formal
//       order of each method on that solution, |m| = 1 preservation.
//   random unit vectors                          bit-exact integer recurrence,
//       checked against hand-computed xorshift64 states.
//   per-cell Ms                                  extension with no reference
//       counterpart: reduces to the uniform case when all Ms_i are equal.
//   Heun adjoint (extension)                     pinned by central finite
//       differences of the oracle's own heun_step (state, per-cell Ms and
//       H_ext directions) and by the dot-product identity.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
using V3 = std::array<double, 3>;

constexpr double PI = 3.14159265358979323846;
// lib.rs:25
constexpr double MU0 = 4.0 * PI * 1e-7;

// ---------------------------------------------------------------------------
// vector.rs:11-76
// ---------------------------------------------------------------------------
inline V3 add(const V3 &a, const V3 &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 sub(const V3 &a, const V3 &b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 scale(const V3 &a, double f) { return {a[0] * f, a[1] * f, a[2] * f}; }
inline double dot(const V3 &a, const V3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(const V3 &a, const V3 &b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double norm(const V3 &a) { return std::sqrt(dot(a, a)); }
// vector.rs:11-19: zero vectors (inactive cells) stay zero.
inline V3 normalized(const V3 &v) {
    double n = norm(v);
    if (n <= 0.0) return {0.0, 0.0, 0.0};
    return scale(v, 1.0 / n);
}
double max_norm(const std::vector<V3> &v) {
    double m = 0.0;
    for (const auto &x : v) m = std::max(m, norm(x));
    return m;
}

// ---------------------------------------------------------------------------
// FFT: forward is unnormalised exp(-2 pi i jk/N), inverse unnormalised
// exp(+2 pi i jk/N), matching rustfft's conventions used at lib.rs:343-386.
// Stockham autosort, mixed radix 2/3/4/5 plus generic small primes; prime
// factors above 61 go through Bluestein on a power-of-two length.
// ---------------------------------------------------------------------------
struct FftPlan {
    int n = 0;
    std::vector<int> radices;
    std::vector<cd> tw;          // tw[t] = exp(-2 pi i t / n), long-double accurate
    std::vector<std::vector<cd>> pass_tw;  // per pass: [k][r-1] = tw[r k n/(ns R)]
    bool bluestein = false;
    int m = 0;                   // Bluestein convolution length
    std::vector<cd> chirp;       // exp(-pi i j^2 / n)
    std::vector<cd> bfft;        // forward FFT of the conjugate chirp (length m)
    std::shared_ptr<FftPlan> sub;
};

std::shared_ptr<FftPlan> make_plan(int n);
void fft_exec(const FftPlan &p, cd *data, bool inverse, cd *scratch);

std::vector<int> factorize(int n, bool &has_large) {
    std::vector<int> r;
    has_large = false;
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    for (int p = 3; p * p <= n; p += 2) {
        while (n % p == 0) { r.push_back(p); n /= p; if (p > 61) has_large = true; }
    }
    if (n > 1) { r.push_back(n); if (n > 61) has_large = true; }
    return r;
}

std::vector<cd> roots(int n) {
    std::vector<cd> w(n);
    for (int t = 0; t < n; ++t) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
        w[t] = cd((double)cosl(a), (double)sinl(a));
    }
    return w;
}

std::shared_ptr<FftPlan> make_plan(int n) {
    auto p = std::make_shared<FftPlan>();
    p->n = n;
    bool large = false;
    p->radices = factorize(n, large);
    p->tw = roots(n);
    {
        int ns = 1;
        for (int R : p->radices) {
            std::vector<cd> t((size_t)ns * (R - 1));
            const int stride = n / (ns * R);
            for (int k = 0; k < ns; ++k)
                for (int r = 1; r < R; ++r) t[(size_t)k * (R - 1) + (r - 1)] = p->tw[(size_t)r * k * stride];
            p->pass_tw.push_back(std::move(t));
            ns *= R;
        }
    }
    if (large) {
        p->bluestein = true;
        int m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p->m = m;
        p->chirp.resize(n);
        for (int j = 0; j < n; ++j) {
            // exact reduction of j^2 modulo 2n keeps the chirp angle small
            long long q = ((long long)j * j) % (2LL * n);
            long double a = -3.141592653589793238462643383279502884L * (long double)q / (long double)n;
            p->chirp[j] = cd((double)cosl(a), (double)sinl(a));
        }
        p->sub = make_plan(m);
        std::vector<cd> b(m, cd(0, 0));
        b[0] = std::conj(p->chirp[0]);
        for (int j = 1; j < n; ++j) b[j] = b[m - j] = std::conj(p->chirp[j]);
        std::vector<cd> scratch(m);
        p->bfft = b;
        // forward transform of b
        fft_exec(*p->sub, p->bfft.data(), false, scratch.data());
    }
    return p;
}

std::shared_ptr<FftPlan> get_plan(int n) {
    static std::mutex mu;
    static std::map<int, std::shared_ptr<FftPlan>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    auto p = make_plan(n);
    cache[n] = p;
    return p;
}

// Stockham DIT pass structure: for radix R and span Ns (product of radices
// already applied), butterfly j reads x[j + r*N/R], twiddles by w^(r*k*N/(Ns R))
// with k = j mod Ns, applies the R-point DFT and writes
// y[(j/Ns)*Ns*R + k + r*Ns].  Output ends in natural order.
void fft_exec(const FftPlan &p, cd *data, bool inverse, cd *scratch) {
    const int n = p.n;
    if (n <= 1) return;
    if (p.bluestein) {
        const int m = p.m;
        std::vector<cd> a(m, cd(0, 0)), s(m);
        for (int j = 0; j < n; ++j) {
            cd c = inverse ? std::conj(p.chirp[j]) : p.chirp[j];
            a[j] = data[j] * c;
        }
        fft_exec(*p.sub, a.data(), false, s.data());
        // conv with conj chirp (forward) or chirp (inverse): FFT(b) for the
        // inverse direction is the conjugate-reversed spectrum of the forward one
        for (int k = 0; k < m; ++k) {
            cd b = inverse ? std::conj(p.bfft[(m - k) % m]) : p.bfft[k];
            a[k] *= b;
        }
        fft_exec(*p.sub, a.data(), true, s.data());
        for (int k = 0; k < n; ++k) {
            cd c = inverse ? std::conj(p.chirp[k]) : p.chirp[k];
            data[k] = a[k] / (double)m * c;
        }
        return;
    }
    cd *x = data, *y = scratch;
    int ns = 1;
    for (size_t pi = 0; pi < p.radices.size(); ++pi) {
        const int R = p.radices[pi];
        const int nr = n / R;
        const cd *ptw = p.pass_tw[pi].data();
        const int nblk = nr / ns;
        if (R == 2) {
            for (int b = 0; b < nblk; ++b)
                for (int k = 0; k < ns; ++k) {
                    const int j = b * ns + k;
                    cd a = x[j], c = x[j + nr];
                    if (k != 0) c *= inverse ? std::conj(ptw[k]) : ptw[k];
                    const int base = b * ns * 2 + k;
                    y[base] = a + c;
                    y[base + ns] = a - c;
                }
        } else if (R == 4) {
            for (int b = 0; b < nblk; ++b)
                for (int k = 0; k < ns; ++k) {
                    const int j = b * ns + k;
                    cd v0 = x[j], v1 = x[j + nr], v2 = x[j + 2 * nr], v3 = x[j + 3 * nr];
                    if (k != 0) {
                        const cd *w = ptw + (size_t)k * 3;
                        if (inverse) {
                            v1 *= std::conj(w[0]); v2 *= std::conj(w[1]); v3 *= std::conj(w[2]);
                        } else {
                            v1 *= w[0]; v2 *= w[1]; v3 *= w[2];
                        }
                    }
                    cd a0 = v0 + v2, a1 = v0 - v2, a2 = v1 + v3, a3 = v1 - v3;
                    cd ja3 = inverse ? cd(-a3.imag(), a3.real()) : cd(a3.imag(), -a3.real());
                    const int base = b * ns * 4 + k;
                    y[base] = a0 + a2;
                    y[base + ns] = a1 + ja3;
                    y[base + 2 * ns] = a0 - a2;
                    y[base + 3 * ns] = a1 - ja3;
                }
        } else {
            for (int j = 0; j < nr; ++j) {
                const int k = j % ns;
                cd v[64], vv[64];
                for (int r = 0; r < R; ++r) {
                    cd a = x[j + r * nr];
                    if (r != 0 && k != 0) {
                        cd w = ptw[(size_t)k * (R - 1) + (r - 1)];
                        a *= inverse ? std::conj(w) : w;
                    }
                    v[r] = a;
                }
                for (int q = 0; q < R; ++q) {
                    cd acc = v[0];
                    for (int r = 1; r < R; ++r) {
                        int t = (int)(((long long)r * q % R) * (n / R));
                        cd w = inverse ? std::conj(p.tw[t]) : p.tw[t];
                        acc += v[r] * w;
                    }
                    vv[q] = acc;
                }
                const int base = (j / ns) * ns * R + k;
                for (int q = 0; q < R; ++q) y[base + q * ns] = vv[q];
            }
        }
        std::swap(x, y);
        ns *= R;
    }
    if (x != data) std::memcpy(data, x, sizeof(cd) * n);
}


void fft1(cd *data, int n, bool inverse) {
    auto p = get_plan(n);
    std::vector<cd> s(n);
    fft_exec(*p, data, inverse, s.data());
}

// lib.rs:2841-2843
inline size_t padded_index(size_t nx, size_t ny, size_t x, size_t y, size_t z) {
    return x + nx * (y + ny * z);
}

// lib.rs:2768-2812: x lines, then y lines, then z lines.
void fft3_core(cd *data, int nx, int ny, int nz, bool inverse) {
    auto px = get_plan(nx), py = get_plan(ny), pz = get_plan(nz);
    std::vector<cd> s(std::max({nx, ny, nz}));
    std::vector<cd> line(std::max(ny, nz));
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) fft_exec(*px, data + padded_index(nx, ny, 0, y, z), inverse, s.data());
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            for (int y = 0; y < ny; ++y) line[y] = data[padded_index(nx, ny, x, y, z)];
            fft_exec(*py, line.data(), inverse, s.data());
            for (int y = 0; y < ny; ++y) data[padded_index(nx, ny, x, y, z)] = line[y];
        }
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            for (int z = 0; z < nz; ++z) line[z] = data[padded_index(nx, ny, x, y, z)];
            fft_exec(*pz, line.data(), inverse, s.data());
            for (int z = 0; z < nz; ++z) data[padded_index(nx, ny, x, y, z)] = line[z];
        }
}

// ---------------------------------------------------------------------------
// Newell-Williams-Dunlop tensor, newell.rs:21-519
// ---------------------------------------------------------------------------

// newell.rs:21-61
double newell_f(double x, double y, double z) {
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    double r = std::sqrt(r2);
    double result = (2.0 * x2 - y2 - z2) * r / 6.0;
    double rxz2 = x2 + z2;
    if (rxz2 > 1e-300) {
        double arg = 2.0 * y * (y + r) / rxz2;
        if (arg > -1.0) result += y * (z2 - x2) / 4.0 * std::log1p(arg);
    }
    double rxy2 = x2 + y2;
    if (rxy2 > 1e-300) {
        double arg = 2.0 * z * (z + r) / rxy2;
        if (arg > -1.0) result += z * (y2 - x2) / 4.0 * std::log1p(arg);
    }
    if (std::fabs(x) > 1e-300) result -= x * y * z * std::atan(y * z / (x * r));
    return result;
}

// newell.rs:63-123
double newell_g(double x, double y, double z) {
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    double r = std::sqrt(r2);
    double result = -x * y * r / 3.0;
    double rxy2 = x2 + y2;
    if (rxy2 > 1e-300) {
        double arg = 2.0 * z * (z + r) / rxy2;
        if (arg > -1.0) result += x * y * z * std::log1p(arg) / 2.0;
    }
    double ryz2 = y2 + z2;
    if (ryz2 > 1e-300) {
        double arg = 2.0 * x * (x + r) / ryz2;
        if (arg > -1.0) result += y * (3.0 * z2 - y2) * std::log1p(arg) / 12.0;
    }
    double rxz2 = x2 + z2;
    if (rxz2 > 1e-300) {
        double arg = 2.0 * y * (y + r) / rxz2;
        if (arg > -1.0) result += x * (3.0 * z2 - x2) * std::log1p(arg) / 12.0;
    }
    if (std::fabs(z) > 1e-300) result -= z2 * z / 6.0 * std::atan(x * y / (z * r));
    if (std::fabs(y) > 1e-300) result -= y2 * z / 2.0 * std::atan(x * z / (y * r));
    if (std::fabs(x) > 1e-300) result -= x2 * z / 2.0 * std::atan(y * z / (x * r));
    return result;
}

// newell.rs:133-146 (Kahan-Neumaier)
double kahan_sum(const double *terms, int n) {
    double sum = 0.0, comp = 0.0;
    for (int i = 0; i < n; ++i) {
        double val = terms[i];
        double t = sum + val;
        if (std::fabs(sum) >= std::fabs(val)) comp += (sum - t) + val;
        else comp += (val - t) + sum;
        sum = t;
    }
    return sum + comp;
}

// newell.rs:150-203: 27-point stencil with weights 8/-4/2/-1.
double ldia(size_t i, size_t j, size_t k, const std::vector<double> &f, size_t sx, size_t sy,
            double hx, double hy, double hz) {
    i += 1; j += 1; k += 1;
    auto id = [&](size_t a, size_t b, size_t c) { return c * sy * sx + b * sx + a; };
    const double t[27] = {
        8.0 * f[id(i, j, k)],
        -4.0 * f[id(i + 1, j, k)], -4.0 * f[id(i - 1, j, k)],
        -4.0 * f[id(i, j + 1, k)], -4.0 * f[id(i, j - 1, k)],
        -4.0 * f[id(i, j, k + 1)], -4.0 * f[id(i, j, k - 1)],
        2.0 * f[id(i - 1, j - 1, k)], 2.0 * f[id(i - 1, j + 1, k)],
        2.0 * f[id(i + 1, j - 1, k)], 2.0 * f[id(i + 1, j + 1, k)],
        2.0 * f[id(i - 1, j, k - 1)], 2.0 * f[id(i - 1, j, k + 1)],
        2.0 * f[id(i + 1, j, k - 1)], 2.0 * f[id(i + 1, j, k + 1)],
        2.0 * f[id(i, j - 1, k - 1)], 2.0 * f[id(i, j - 1, k + 1)],
        2.0 * f[id(i, j + 1, k - 1)], 2.0 * f[id(i, j + 1, k + 1)],
        -f[id(i - 1, j - 1, k - 1)], -f[id(i - 1, j - 1, k + 1)],
        -f[id(i - 1, j + 1, k - 1)], -f[id(i + 1, j - 1, k - 1)],
        -f[id(i - 1, j + 1, k + 1)], -f[id(i + 1, j - 1, k + 1)],
        -f[id(i + 1, j + 1, k - 1)], -f[id(i + 1, j + 1, k + 1)],
    };
    return kahan_sum(t, 27) / (4.0 * PI * hx * hy * hz);
}

// newell.rs:212-226
double asymptotic_nxx(double x, double y, double z, double vol) {
    double r2 = x * x + y * y + z * z;
    double r = std::sqrt(r2);
    double r3 = r2 * r;
    return (1.0 / r3 - 3.0 * x * x / (r3 * r2)) / (4.0 * PI) * vol;
}
double asymptotic_nxy(double x, double y, double z, double vol) {
    double r2 = x * x + y * y + z * z;
    double r = std::sqrt(r2);
    double r5 = r2 * r2 * r;
    return -3.0 * x * y / (4.0 * PI * r5) * vol;
}

constexpr size_t ASYMPTOTIC_DISTANCE = 40;  // newell.rs:229

// newell.rs:254-519: six real-space components on the (2nx,2ny,2nz) grid.
// `shifted` is compute_newell_kernels_shifted (newell.rs:528-651): every z
// coordinate is offset by z_offset, and the mirrored z slots pz - k keep the
// value and the sign of slot k for all six components (the reference
// places them so; only the x and y parities are applied).
void compute_newell_kernels(size_t nx, size_t ny, size_t nz, double dx, double dy, double dz,
                            std::vector<double> out[6], double z_offset = 0.0, bool shifted = false) {
    const size_t px = 2 * nx, py = 2 * ny, pz = 2 * nz;
    const size_t plen = px * py * pz;
    const size_t fsx = std::min(nx, ASYMPTOTIC_DISTANCE) + 2;
    const size_t fsy = std::min(ny, ASYMPTOTIC_DISTANCE) + 2;
    const size_t fsz = std::min(nz, ASYMPTOTIC_DISTANCE) + 2;
    const size_t flen = fsx * fsy * fsz;
    std::vector<double> fxx(flen), fyy(flen), fzz(flen), gxy(flen), gxz(flen), gyz(flen);
#pragma omp parallel for schedule(dynamic)
    for (long kk = 0; kk < (long)fsz; ++kk) {
        for (size_t j = 0; j < fsy; ++j)
            for (size_t i = 0; i < fsx; ++i) {
                double x = (double)((long)i - 1) * dx;
                double y = (double)((long)j - 1) * dy;
                double z = (double)(kk - 1) * dz + z_offset;
                size_t id = (size_t)kk * fsx * fsy + j * fsx + i;
                fxx[id] = newell_f(x, y, z);
                fyy[id] = newell_f(y, x, z);
                fzz[id] = newell_f(z, y, x);
                gxy[id] = newell_g(x, y, z);
                gxz[id] = newell_g(x, z, y);
                gyz[id] = newell_g(y, z, x);
            }
    }
    for (int c = 0; c < 6; ++c) out[c].assign(plen, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (long kl = 0; kl < (long)nz; ++kl) {
        size_t k = (size_t)kl;
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                double nxx, nyy, nzz, nxy, nxz, nyz;
                size_t dist2 = i * i + j * j + k * k;
                bool asym = i >= ASYMPTOTIC_DISTANCE || j >= ASYMPTOTIC_DISTANCE || k >= ASYMPTOTIC_DISTANCE ||
                            dist2 >= ASYMPTOTIC_DISTANCE * ASYMPTOTIC_DISTANCE;
                if (asym) {
                    double x = (double)i * dx, y = (double)j * dy, z = (double)k * dz + z_offset;
                    double vol = dx * dy * dz;
                    nxx = asymptotic_nxx(x, y, z, vol);
                    nyy = asymptotic_nxx(y, x, z, vol);
                    nzz = asymptotic_nxx(z, y, x, vol);
                    nxy = asymptotic_nxy(x, y, z, vol);
                    nxz = asymptotic_nxy(x, z, y, vol);
                    nyz = asymptotic_nxy(y, z, x, vol);
                } else {
                    nxx = ldia(i, j, k, fxx, fsx, fsy, dx, dy, dz);
                    nyy = ldia(i, j, k, fyy, fsx, fsy, dy, dx, dz);
                    nzz = ldia(i, j, k, fzz, fsx, fsy, dz, dy, dx);
                    nxy = ldia(i, j, k, gxy, fsx, fsy, dx, dy, dz);
                    nxz = ldia(i, j, k, gxz, fsx, fsy, dx, dz, dy);
                    nyz = ldia(i, j, k, gyz, fsx, fsy, dy, dz, dx);
                }
                size_t xs_i[2] = {i, px - i}, ys_i[2] = {j, py - j}, zs_i[2] = {k, pz - k};
                double xs_s[2] = {1.0, -1.0}, ys_s[2] = {1.0, -1.0}, zs_s[2] = {1.0, shifted ? 1.0 : -1.0};
                int nxs = i == 0 ? 1 : 2, nys = j == 0 ? 1 : 2, nzs = k == 0 ? 1 : 2;
                for (int a = 0; a < nxs; ++a)
                    for (int b = 0; b < nys; ++b)
                        for (int c = 0; c < nzs; ++c) {
                            size_t p = zs_i[c] * py * px + ys_i[b] * px + xs_i[a];
                            double sx = xs_s[a], sy = ys_s[b], sz = zs_s[c];
                            out[0][p] = nxx;
                            out[1][p] = nyy;
                            out[2][p] = nzz;
                            out[3][p] = nxy * sx * sy;
                            out[4][p] = nxz * sx * sz;
                            out[5][p] = nyz * sy * sz;
                        }
            }
    }
}

// ---------------------------------------------------------------------------
// Engine: ExchangeLlgProblem + state (lib.rs:869-2700)
// ---------------------------------------------------------------------------
enum Integrator { HEUN = 1, DP45 = 2, ABM3 = 3, RK4 = 4, RK23 = 5 };

struct Report {
    double time_seconds, dt_used, suggested_next_dt;
    double e_ex, e_demag, e_ext, e_total;
    double max_h_eff, max_h_demag, max_rhs;
};

struct Problem {
    size_t nx, ny, nz, n;
    double dx, dy, dz;
    double ms, a, alpha, gamma;
    std::vector<double> ms_cell;  // extension: per-cell Ms (empty = uniform)
    int integrator;
    bool precession, exchange, demag, has_ext;
    V3 hext;
    std::vector<uint8_t> mask;  // empty = all active
    // per-pair exchange table (native/backends/fdm/src/exchange_fp64.cu:80-114):
    // region id per cell and a 256 x 256 table of A_ij; empty = uniform A
    std::vector<uint32_t> region;
    std::vector<double> lut;
    double max_error, dt_min, dt_max, headroom;
    // sub-cell boundary tiers (plan_desc boundary_correction; the reference's
    // FP64 kernels native/backends/fdm/src/exchange_t0_fp64.cu,
    // exchange_t1_fp64.cu, demag_boundary_fp64.cu and the volume-fraction
    // packing of demag_fp64.cu:54-100): tier 1 = T0 face-link exchange, tier
    // 2 = T1 intersection-distance exchange; empty phi = no tier
    int tier = 0;
    double phi_floor = 0.05, delta_min = 0.0;
    std::vector<double> phi;
    std::vector<double> face[6];   // xp, xm, yp, ym, zp, zm
    std::vector<double> delta[6];  // xp, xm, yp, ym, zp, zm
    std::vector<int32_t> corr_t, corr_s;
    std::vector<double> corr_n;    // per (target, stencil entry): Nxx, Nxy, Nxz, Nyy, Nyz, Nzz
    uint32_t corr_stencil = 0;
    bool tiers() const { return tier > 0 && !phi.empty(); }
    // state
    std::vector<V3> m;
    double time = 0.0;
    std::vector<V3> k_fsal;
    bool has_fsal = false;
    // ABM3 history (lib.rs:620-680)
    std::vector<V3> f_n, f_n1, f_n2;
    int abm_count = 0;  // number of stored entries (0..3)
    int startup_steps = 0;
    double abm_last_dt = 0.0;
    // FFT workspace (lib.rs:337-405)
    size_t px, py, pz;
    std::vector<cd> kern[6];
    std::vector<cd> bm[3], bh[3];

    bool active(size_t i) const { return mask.empty() || mask[i] != 0; }
    bool use_lut() const { return !region.empty(); }
    double a_pair(size_t i, size_t j) const { return lut[(size_t)region[i] * kRegions + region[j]]; }
    static constexpr size_t kRegions = 256;  // FULLMAG_FDM_MAX_EXCHANGE_REGIONS
    double ms_at(size_t i) const { return ms_cell.empty() ? ms : ms_cell[i]; }
};

void build_workspace(Problem &p) {
    p.px = 2 * p.nx; p.py = 2 * p.ny; p.pz = 2 * p.nz;
    size_t plen = p.px * p.py * p.pz;
    for (int c = 0; c < 3; ++c) { p.bm[c].assign(plen, cd(0, 0)); p.bh[c].assign(plen, cd(0, 0)); }
    if (!p.demag) return;
    std::vector<double> nk[6];
    compute_newell_kernels(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, nk);
    for (int c = 0; c < 6; ++c) {
        p.kern[c].resize(plen);
        for (size_t i = 0; i < plen; ++i) p.kern[c][i] = cd(nk[c][i], 0.0);
        fft3_core(p.kern[c].data(), (int)p.px, (int)p.py, (int)p.pz, false);
    }
}

// Per-pair exchange sum at cell f with the region table,
// native/backends/fdm/src/exchange_fp64.cu:80-114:
//   sum over the six faces of A(r_f, r_j) (q_j - q_f) / d^2,
// where a face at the grid edge or towards an inactive cell takes j = f (its
// difference is zero whatever A is), and A is read from row r_f of the table.
// The engine (crates/fullmag-engine) has no region table; this follows the
// reference GPU kernel that the runner drives whenever region_mask is set.
V3 pair_sum(const Problem &p, const std::vector<V3> &q, size_t f) {
    const size_t x = f % p.nx, y = (f / p.nx) % p.ny, z = f / (p.nx * p.ny);
    auto nbi = [&](size_t a, size_t b, size_t cc) -> size_t {
        const size_t id = a + p.nx * (b + p.ny * cc);
        return p.active(id) ? id : f;
    };
    const size_t xm = nbi(x > 0 ? x - 1 : x, y, z), xp = nbi(x + 1 < p.nx ? x + 1 : x, y, z);
    const size_t ym = nbi(x, y > 0 ? y - 1 : y, z), yp = nbi(x, y + 1 < p.ny ? y + 1 : y, z);
    const size_t zm = nbi(x, y, z > 0 ? z - 1 : z), zp = nbi(x, y, z + 1 < p.nz ? z + 1 : z);
    const double ix = 1.0 / (p.dx * p.dx), iy = 1.0 / (p.dy * p.dy), iz = 1.0 / (p.dz * p.dz);
    const V3 c = q[f];
    V3 s;
    for (int k = 0; k < 3; ++k)
        s[k] = p.a_pair(f, xp) * (q[xp][k] - c[k]) * ix + p.a_pair(f, xm) * (q[xm][k] - c[k]) * ix +
               p.a_pair(f, yp) * (q[yp][k] - c[k]) * iy + p.a_pair(f, ym) * (q[ym][k] - c[k]) * iy +
               p.a_pair(f, zp) * (q[zp][k] - c[k]) * iz + p.a_pair(f, zm) * (q[zm][k] - c[k]) * iz;
    return s;
}

// T0 exchange, native/backends/fdm/src/exchange_t0_fp64.cu:19-135: empty
// cells (phi <= 0) and zero m give 0; otherwise
//   H = 2A / (mu0 Ms) / max(phi, phi_floor) * sum_faces f (m_j - m) / d^2
// over the faces with link f > 0 and a neighbour inside the grid (z faces
// only when nz > 1), in the order -x, +x, -y, +y, -z, +z.
V3 exchange_t0(const Problem &p, const std::vector<V3> &m, size_t f) {
    const double phi0 = p.phi[f];
    const V3 c = m[f];
    if (phi0 <= 0.0 || (c[0] == 0.0 && c[1] == 0.0 && c[2] == 0.0)) return {0, 0, 0};
    const size_t x = f % p.nx, y = (f / p.nx) % p.ny, z = f / (p.nx * p.ny);
    const size_t nxy = p.nx * p.ny;
    const double inv_phi = 1.0 / (phi0 > p.phi_floor ? phi0 : p.phi_floor);
    const double ix = 1.0 / (p.dx * p.dx), iy = 1.0 / (p.dy * p.dy), iz = 1.0 / (p.dz * p.dz);
    V3 b{0, 0, 0};
    auto face = [&](int k, bool inside, size_t j, double inv2) {
        const double fl = p.face[k][f];
        if (fl > 0.0 && inside)
            for (int c3 = 0; c3 < 3; ++c3) b[c3] += fl * inv2 * (m[j][c3] - c[c3]);
    };
    face(1, x > 0, f - 1, ix);
    face(0, x + 1 < p.nx, f + 1, ix);
    face(3, y > 0, f - p.nx, iy);
    face(2, y + 1 < p.ny, f + p.nx, iy);
    if (p.nz > 1) {
        face(5, z > 0, f - nxy, iz);
        face(4, z + 1 < p.nz, f + nxy, iz);
    }
    return scale(b, 2.0 * p.a / (MU0 * p.ms_at(f)) * inv_phi);
}

// T1 exchange, native/backends/fdm/src/exchange_t1_fp64.cu:26-177: empty
// cells and zero m give 0; with delta = max(delta_side, delta_min) per side,
// a neighbour inside the grid with nonzero m and delta > 0 on its side adds
//   2 (m_j - m) / (d (d + delta_opposite))
// (the opposite side's distance sets the non-uniform spacing); z faces only
// when nz > 1; H = 2A / (mu0 Ms) * sum.
V3 exchange_t1(const Problem &p, const std::vector<V3> &m, size_t f) {
    const double phi0 = p.phi[f];
    const V3 c = m[f];
    if (phi0 <= 0.0 || (c[0] == 0.0 && c[1] == 0.0 && c[2] == 0.0)) return {0, 0, 0};
    const size_t x = f % p.nx, y = (f / p.nx) % p.ny, z = f / (p.nx * p.ny);
    const size_t nxy = p.nx * p.ny;
    V3 b{0, 0, 0};
    auto side = [&](bool inside, double dnear, double dfar, size_t j, double d) {
        if (!(inside && dnear > 0.0)) return;
        const V3 mn = m[j];
        if (mn[0] == 0.0 && mn[1] == 0.0 && mn[2] == 0.0) return;
        const double coeff = 2.0 / (d * (d + dfar));
        for (int c3 = 0; c3 < 3; ++c3) b[c3] += coeff * (mn[c3] - c[c3]);
    };
    auto dl = [&](int k) { return std::max(p.delta[k][f], p.delta_min); };
    const double dxp = dl(0), dxm = dl(1), dyp = dl(2), dym = dl(3);
    side(x > 0, dxm, dxp, f - 1, p.dx);
    side(x + 1 < p.nx, dxp, dxm, f + 1, p.dx);
    side(y > 0, dym, dyp, f - p.nx, p.dy);
    side(y + 1 < p.ny, dyp, dym, f + p.nx, p.dy);
    if (p.nz > 1) {
        const double dzp = dl(4), dzm = dl(5);
        side(z > 0, dzm, dzp, f - nxy, p.dz);
        side(z + 1 < p.nz, dzp, dzm, f + nxy, p.dz);
    }
    return scale(b, 2.0 * p.a / (MU0 * p.ms_at(f)));
}

// Which exchange the reference launches (native/backends/fdm/src/
// exchange_fp64.cu:158-214): T0 for tier 1 with face links, T1 for tier >= 2
// with intersection distances, else the standard stencil.
int exchange_tier(const Problem &p) {
    if (!p.tiers()) return 0;
    if (p.tier == 1 && !p.face[0].empty()) return 1;
    if (p.tier >= 2 && !p.delta[0].empty()) return 2;
    return 0;
}

// lib.rs:2247-2299 (per-cell Ms generalises the prefactor 2A/(mu0 Ms)); with a
// region table, H_f = 2 / (mu0 Ms_f) pair_sum(m, f).  `standard`: the
// masked stencil even when a boundary tier replaces it for the field.
std::vector<V3> exchange_field(const Problem &p, const std::vector<V3> &m, bool standard = false) {
    std::vector<V3> h(p.n);
    if (const int t = standard ? 0 : exchange_tier(p)) {
#pragma omp parallel for schedule(static)
        for (long fl = 0; fl < (long)p.n; ++fl)
            h[fl] = t == 1 ? exchange_t0(p, m, (size_t)fl) : exchange_t1(p, m, (size_t)fl);
        return h;
    }
    const double dx2 = p.dx * p.dx, dy2 = p.dy * p.dy, dz2 = p.dz * p.dz;
#pragma omp parallel for schedule(static)
    for (long fl = 0; fl < (long)p.n; ++fl) {
        size_t f = (size_t)fl;
        if (!p.active(f)) { h[f] = {0, 0, 0}; continue; }
        const double prefactor = 2.0 * p.a / (MU0 * p.ms_at(f));
        size_t x = f % p.nx, y = (f / p.nx) % p.ny, z = f / (p.nx * p.ny);
        const V3 c = m[f];
        auto nb = [&](size_t a, size_t b, size_t cc) -> V3 {
            size_t id = a + p.nx * (b + p.ny * cc);
            return p.active(id) ? m[id] : c;
        };
        V3 xm = nb(x > 0 ? x - 1 : 0, y, z), xp = nb(std::min(x + 1, p.nx - 1), y, z);
        V3 ym = nb(x, y > 0 ? y - 1 : 0, z), yp = nb(x, std::min(y + 1, p.ny - 1), z);
        V3 zm = nb(x, y, z > 0 ? z - 1 : 0), zp = nb(x, y, std::min(z + 1, p.nz - 1));
        if (p.use_lut()) {
            h[f] = scale(pair_sum(p, m, f), 2.0 / (MU0 * p.ms_at(f)));
            continue;
        }
        V3 lap;
        for (int k = 0; k < 3; ++k)
            lap[k] = (xp[k] - 2.0 * c[k] + xm[k]) / dx2 + (yp[k] - 2.0 * c[k] + ym[k]) / dy2 +
                     (zp[k] - 2.0 * c[k] + zm[k]) / dz2;
        h[f] = scale(lap, prefactor);
    }
    return h;
}

// lib.rs:2306-2400.  weight_ms = false applies the bare convolution -N * x
// (the linear operator with Ms = 1), which the adjoint needs.
std::vector<V3> demag_apply(Problem &p, const std::vector<V3> &m, bool weight_ms) {
    const size_t px = p.px, py = p.py, pz = p.pz, plen = px * py * pz;
    for (int c = 0; c < 3; ++c) {
        std::fill(p.bm[c].begin(), p.bm[c].end(), cd(0, 0));
        std::fill(p.bh[c].begin(), p.bh[c].end(), cd(0, 0));
    }
    for (size_t z = 0; z < p.nz; ++z)
        for (size_t y = 0; y < p.ny; ++y)
            for (size_t x = 0; x < p.nx; ++x) {
                size_t s = x + p.nx * (y + p.ny * z), d = padded_index(px, py, x, y, z);
                // with tiers every cell packs phi Ms m (demag_fp64.cu:80-85)
                V3 mo = p.tiers() ? scale(m[s], p.phi[s] * (weight_ms ? p.ms_at(s) : 1.0))
                        : p.active(s) ? scale(m[s], weight_ms ? p.ms_at(s) : 1.0)
                                      : V3{0, 0, 0};
                for (int c = 0; c < 3; ++c) p.bm[c][d] = cd(mo[c], 0.0);
            }
    for (int c = 0; c < 3; ++c) fft3_core(p.bm[c].data(), (int)px, (int)py, (int)pz, false);
    const auto &kxx = p.kern[0], &kyy = p.kern[1], &kzz = p.kern[2];
    const auto &kxy = p.kern[3], &kxz = p.kern[4], &kyz = p.kern[5];
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)plen; ++il) {
        size_t i = (size_t)il;
        cd mx = p.bm[0][i], my = p.bm[1][i], mz = p.bm[2][i];
        p.bh[0][i] = -(kxx[i] * mx + kxy[i] * my + kxz[i] * mz);
        p.bh[1][i] = -(kxy[i] * mx + kyy[i] * my + kyz[i] * mz);
        p.bh[2][i] = -(kxz[i] * mx + kyz[i] * my + kzz[i] * mz);
    }
    for (int c = 0; c < 3; ++c) fft3_core(p.bh[c].data(), (int)px, (int)py, (int)pz, true);
    const double norm_ = 1.0 / (double)plen;
    std::vector<V3> h(p.n);
    for (size_t z = 0; z < p.nz; ++z)
        for (size_t y = 0; y < p.ny; ++y)
            for (size_t x = 0; x < p.nx; ++x) {
                size_t s = padded_index(px, py, x, y, z), d = x + p.nx * (y + p.ny * z);
                h[d] = p.active(d) ? V3{p.bh[0][s].real() * norm_, p.bh[1][s].real() * norm_,
                                        p.bh[2][s].real() * norm_}
                                   : V3{0, 0, 0};
            }
    // sparse near-field correction, native/backends/fdm/src/
    // demag_boundary_fp64.cu:20-76: H_t += -dN(t, s) . (phi_s Ms m_s)
    if (p.tiers() && p.corr_stencil > 0) {
        for (size_t t = 0; t < p.corr_t.size(); ++t) {
            const int32_t ti = p.corr_t[t];
            if (ti < 0) continue;
            V3 corr{0, 0, 0};
            for (uint32_t e = 0; e < p.corr_stencil; ++e) {
                const int32_t si = p.corr_s[t * p.corr_stencil + e];
                if (si < 0) continue;
                const double *nn = &p.corr_n[(t * p.corr_stencil + e) * 6];
                const double w = p.phi[si] * (weight_ms ? p.ms_at(si) : 1.0);
                const V3 mm = scale(m[si], w);
                corr[0] -= nn[0] * mm[0] + nn[1] * mm[1] + nn[2] * mm[2];
                corr[1] -= nn[1] * mm[0] + nn[3] * mm[1] + nn[4] * mm[2];
                corr[2] -= nn[2] * mm[0] + nn[4] * mm[1] + nn[5] * mm[2];
            }
            h[ti] = add(h[ti], corr);
        }
    }
    return h;
}

std::vector<V3> demag_field(Problem &p, const std::vector<V3> &m) { return demag_apply(p, m, true); }

// Discrete Laplacian sum_j w_ij (q_j - q_i) with the exchange stencil's
// clamped and masked neighbours (exchange_field = (2A / (mu0 Ms)) lap(m)), or
// with w_ij = A_ij from the region table (exchange_field = 2 / (mu0 Ms) lap(m)).
// Symmetric in i, j (for a symmetric table), so it is its own transpose.
std::vector<V3> lap_field(const Problem &p, const std::vector<V3> &q) {
    std::vector<V3> h(p.n);
    const double dx2 = p.dx * p.dx, dy2 = p.dy * p.dy, dz2 = p.dz * p.dz;
#pragma omp parallel for schedule(static)
    for (long fl = 0; fl < (long)p.n; ++fl) {
        size_t f = (size_t)fl;
        if (!p.active(f)) { h[f] = {0, 0, 0}; continue; }
        if (p.use_lut()) {
            // weighted by A_ij (symmetric tables give a symmetric operator)
            h[f] = pair_sum(p, q, f);
            continue;
        }
        size_t x = f % p.nx, y = (f / p.nx) % p.ny, z = f / (p.nx * p.ny);
        const V3 c = q[f];
        auto nb = [&](size_t a, size_t b, size_t cc) -> V3 {
            size_t id = a + p.nx * (b + p.ny * cc);
            return p.active(id) ? q[id] : c;
        };
        V3 xm = nb(x > 0 ? x - 1 : 0, y, z), xp = nb(std::min(x + 1, p.nx - 1), y, z);
        V3 ym = nb(x, y > 0 ? y - 1 : 0, z), yp = nb(x, std::min(y + 1, p.ny - 1), z);
        V3 zm = nb(x, y, z > 0 ? z - 1 : 0), zp = nb(x, y, std::min(z + 1, p.nz - 1));
        for (int k = 0; k < 3; ++k)
            h[f][k] = (xp[k] - 2.0 * c[k] + xm[k]) / dx2 + (yp[k] - 2.0 * c[k] + ym[k]) / dy2 +
                      (zp[k] - 2.0 * c[k] + zm[k]) / dz2;
    }
    return h;
}

// lib.rs:2584-2595
inline V3 llg_rhs_from_field(const Problem &p, const V3 &m, const V3 &h) {
    double gamma_bar = p.gamma / (1.0 + p.alpha * p.alpha);
    V3 prec = cross(m, h);
    V3 damp = cross(m, prec);
    V3 pt = p.precession ? prec : V3{0, 0, 0};
    return scale(add(pt, scale(damp, p.alpha)), -gamma_bar);
}

struct Eval {
    std::vector<V3> h_ex, h_d, h_ext, h_eff, rhs;
    Report r;
};

// lib.rs:2529-2582 (llg_rhs_full_ws) and energies lib.rs:2653-2703.
Eval full_eval(Problem &p, const std::vector<V3> &m, bool want_stats) {
    Eval e;
    e.h_ex = p.exchange ? exchange_field(p, m) : std::vector<V3>(p.n, V3{0, 0, 0});
    e.h_d = p.demag ? demag_field(p, m) : std::vector<V3>(p.n, V3{0, 0, 0});
    e.h_ext.resize(p.n);
    const V3 ext = p.has_ext ? p.hext : V3{0, 0, 0};
    for (size_t i = 0; i < p.n; ++i) e.h_ext[i] = p.active(i) ? ext : V3{0, 0, 0};
    e.h_eff.resize(p.n);
    e.rhs.resize(p.n);
#pragma omp parallel for schedule(static)
    for (long il = 0; il < (long)p.n; ++il) {
        size_t i = (size_t)il;
        e.h_eff[i] = add(add(e.h_ex[i], e.h_d[i]), e.h_ext[i]);  // lib.rs:2849-2869
        e.rhs[i] = llg_rhs_from_field(p, m[i], e.h_eff[i]);
    }
    if (!want_stats) return e;
    const double v = p.dx * p.dy * p.dz;
    double ex = 0, ed = 0, eext = 0;
    // with a boundary tier the exchange energy is still that of the standard
    // masked stencil (native/backends/fdm/src/reductions_fp64.cu:123-200 sums
    // pair differences without face links or distances)
    std::vector<V3> h_std;
    if (p.exchange && exchange_tier(p)) h_std = exchange_field(p, m, true);
    const std::vector<V3> &h_energy = h_std.empty() ? e.h_ex : h_std;
    for (size_t i = 0; i < p.n; ++i) {
        double ms = p.ms_at(i);
        if (p.exchange && p.active(i)) ex += -0.5 * MU0 * ms * dot(m[i], h_energy[i]) * v;
        if (p.demag) ed += -0.5 * MU0 * ms * dot(m[i], e.h_d[i]) * v;
        if (p.has_ext) eext += -MU0 * ms * dot(m[i], e.h_ext[i]) * v;
    }
    e.r.e_ex = ex; e.r.e_demag = ed; e.r.e_ext = eext;
    e.r.e_total = ex + ed + eext;
    e.r.max_h_eff = max_norm(e.h_eff);
    e.r.max_h_demag = max_norm(e.h_d);
    e.r.max_rhs = max_norm(e.rhs);
    return e;
}

std::vector<V3> rhs(Problem &p, const std::vector<V3> &m) { return full_eval(p, m, false).rhs; }

Report finish(Problem &p, double dt, double suggested) {
    Eval e = full_eval(p, p.m, true);
    Report r = e.r;
    r.time_seconds = p.time;
    r.dt_used = dt;
    r.suggested_next_dt = suggested;
    return r;
}

// lib.rs:1217-1249
Report heun_step(Problem &p, double dt) {
    const size_t n = p.n;
    std::vector<V3> m0 = p.m;
    std::vector<V3> k1 = rhs(p, m0);
    std::vector<V3> ms(n);
    for (size_t i = 0; i < n; ++i) ms[i] = normalized(add(m0[i], scale(k1[i], dt)));
    std::vector<V3> k2 = rhs(p, ms);
    for (size_t i = 0; i < n; ++i) p.m[i] = normalized(add(m0[i], scale(add(k1[i], k2[i]), 0.5 * dt)));
    p.time += dt;
    return finish(p, dt, 0.0);
}

// lib.rs:1254-1302
Report rk4_step(Problem &p, double dt) {
    const size_t n = p.n;
    std::vector<V3> m0 = p.m, st(n);
    auto k1 = rhs(p, m0);
    for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k1[i], 0.5 * dt)));
    auto k2 = rhs(p, st);
    for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k2[i], 0.5 * dt)));
    auto k3 = rhs(p, st);
    for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k3[i], dt)));
    auto k4 = rhs(p, st);
    for (size_t i = 0; i < n; ++i)
        p.m[i] = normalized(add(m0[i], scale(add(add(k1[i], scale(k2[i], 2.0)), add(scale(k3[i], 2.0), k4[i])),
                                             dt / 6.0)));
    p.time += dt;
    return finish(p, dt, 0.0);
}

// lib.rs:1653-1674
double max_error_norm(const std::vector<std::pair<const std::vector<V3> *, double>> &ws, double dt, size_t n) {
    double mx = 0.0;
    for (size_t i = 0; i < n; ++i) {
        V3 e = {0, 0, 0};
        for (auto &w : ws) {
            const V3 &k = (*w.first)[i];
            e[0] += w.second * k[0];
            e[1] += w.second * k[1];
            e[2] += w.second * k[2];
        }
        e[0] *= dt; e[1] *= dt; e[2] *= dt;
        mx = std::max(mx, norm(e));
    }
    return mx;
}

// lib.rs:1332-1407, one attempt of size dt from m0: the candidate state in
// st and the returned error norm (Bogacki-Shampine 3(2), no FSAL).
double rk23_attempt(Problem &p, const std::vector<V3> &m0, double dt, std::vector<V3> &st) {
    const size_t n = p.n;
    st.assign(n, V3{0, 0, 0});
    auto k1 = rhs(p, m0);
    for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k1[i], 0.5 * dt)));
    auto k2 = rhs(p, st);
    for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k2[i], 0.75 * dt)));
    auto k3 = rhs(p, st);
    for (size_t i = 0; i < n; ++i) {
        V3 d = scale(add(add(scale(k1[i], 2.0 / 9.0), scale(k2[i], 1.0 / 3.0)), scale(k3[i], 4.0 / 9.0)), dt);
        st[i] = normalized(add(m0[i], d));
    }
    auto k4 = rhs(p, st);
    return max_error_norm({{&k1, -5.0 / 72.0}, {&k2, 1.0 / 12.0}, {&k3, 1.0 / 9.0}, {&k4, -1.0 / 8.0}}, dt, n);
}

// lib.rs:1395-1404: dt for the next step (or the retry) from the error norm.
double next_dt(const Problem &p, double dt, double err, double order_inv) {
    return std::min(std::max(p.headroom * dt * std::pow(p.max_error / std::max(err, 1e-30), order_inv), p.dt_min),
                    p.dt_max);
}

Report rk23_step(Problem &p, double dt) {
    dt = std::max(std::min(dt, p.dt_max), p.dt_min);
    std::vector<V3> m0 = p.m, st;
    for (;;) {
        const double err = rk23_attempt(p, m0, dt, st);
        if (err <= p.max_error || dt <= p.dt_min) {
            p.m = st;
            p.time += dt;
            return finish(p, dt, next_dt(p, dt, err, 1.0 / 3.0));
        }
        double dt_new = p.headroom * dt * std::pow(p.max_error / err, 1.0 / 3.0);
        dt = std::min(std::max(dt_new, p.dt_min), p.dt_max);
    }
}

// lib.rs:1412-1568, one attempt of size dt from m0 with first slope k1: the
// candidate in st, the last slope (the next step's FSAL slope) in k7 and the
// returned error norm (Dormand-Prince 5(4)).
double dp45_attempt(Problem &p, const std::vector<V3> &m0, const std::vector<V3> &k1, double dt,
                    std::vector<V3> &st, std::vector<V3> &k7) {
    const size_t n = p.n;
    const double A21 = 1.0 / 5.0;
    const double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
    const double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
    const double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0, A54 = -212.0 / 729.0;
    const double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0, A64 = 49.0 / 176.0,
                 A65 = -5103.0 / 18656.0;
    const double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0, B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
    const double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0, E5 = -17253.0 / 339200.0,
                 E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;
    st.assign(n, V3{0, 0, 0});
    for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k1[i], A21 * dt)));
    auto k2 = rhs(p, st);
    for (size_t i = 0; i < n; ++i)
        st[i] = normalized(add(m0[i], scale(add(scale(k1[i], A31), scale(k2[i], A32)), dt)));
    auto k3 = rhs(p, st);
    for (size_t i = 0; i < n; ++i)
        st[i] = normalized(add(m0[i], scale(add(add(scale(k1[i], A41), scale(k2[i], A42)), scale(k3[i], A43)), dt)));
    auto k4 = rhs(p, st);
    for (size_t i = 0; i < n; ++i)
        st[i] = normalized(add(m0[i], scale(add(add(scale(k1[i], A51), scale(k2[i], A52)),
                                                add(scale(k3[i], A53), scale(k4[i], A54))),
                                            dt)));
    auto k5 = rhs(p, st);
    for (size_t i = 0; i < n; ++i)
        st[i] = normalized(add(m0[i], scale(add(add(add(scale(k1[i], A61), scale(k2[i], A62)), scale(k3[i], A63)),
                                                add(scale(k4[i], A64), scale(k5[i], A65))),
                                            dt)));
    auto k6 = rhs(p, st);
    for (size_t i = 0; i < n; ++i)
        st[i] = normalized(add(m0[i], scale(add(add(add(scale(k1[i], B1), scale(k3[i], B3)), scale(k4[i], B4)),
                                                add(scale(k5[i], B5), scale(k6[i], B6))),
                                            dt)));
    k7 = rhs(p, st);
    return max_error_norm({{&k1, E1}, {&k3, E3}, {&k4, E4}, {&k5, E5}, {&k6, E6}, {&k7, E7}}, dt, n);
}

Report dp45_step(Problem &p, double dt) {
    dt = std::max(std::min(dt, p.dt_max), p.dt_min);
    std::vector<V3> m0 = p.m, st, k7;
    for (;;) {
        std::vector<V3> k1;
        if (p.has_fsal) { k1 = p.k_fsal; p.has_fsal = false; p.k_fsal.clear(); }
        else k1 = rhs(p, m0);
        const double err = dp45_attempt(p, m0, k1, dt, st, k7);
        if (err <= p.max_error || dt <= p.dt_min) {
            p.m = st;
            p.time += dt;
            p.k_fsal = k7;
            p.has_fsal = true;
            return finish(p, dt, next_dt(p, dt, err, 0.2));
        }
        double dt_new = p.headroom * dt * std::pow(p.max_error / err, 0.2);
        dt = std::min(std::max(dt_new, p.dt_min), p.dt_max);
    }
}

// lib.rs:663-674 AbmHistory::push (restart when dt changes by more than 10%)
void abm_push(Problem &p, std::vector<V3> f, double dt) {
    if (p.abm_last_dt > 0.0 && std::fabs(dt - p.abm_last_dt) / p.abm_last_dt > 0.1) {
        p.f_n.clear(); p.f_n1.clear(); p.f_n2.clear();
        p.abm_count = 0; p.startup_steps = 0; p.abm_last_dt = 0.0;
    }
    p.f_n2 = std::move(p.f_n1);
    p.f_n1 = std::move(p.f_n);
    p.f_n = std::move(f);
    p.abm_count = std::min(p.abm_count + 1, 3);
    p.startup_steps = std::min(p.startup_steps + 1, 3);
    p.abm_last_dt = dt;
}

// lib.rs:1573-1648
Report abm3_step(Problem &p, double dt) {
    const size_t n = p.n;
    bool ready = p.startup_steps >= 3 && !p.f_n.empty() && !p.f_n1.empty() && !p.f_n2.empty();
    std::vector<V3> m0 = p.m, st(n);
    if (!ready) {
        auto k1 = rhs(p, m0);
        for (size_t i = 0; i < n; ++i) st[i] = normalized(add(m0[i], scale(k1[i], dt)));
        auto k2 = rhs(p, st);
        for (size_t i = 0; i < n; ++i) p.m[i] = normalized(add(m0[i], scale(add(k1[i], k2[i]), 0.5 * dt)));
        p.time += dt;
        abm_push(p, rhs(p, p.m), dt);
        return finish(p, dt, 0.0);
    }
    for (size_t i = 0; i < n; ++i) {
        V3 pred = add(add(scale(p.f_n[i], 23.0 / 12.0), scale(p.f_n1[i], -16.0 / 12.0)), scale(p.f_n2[i], 5.0 / 12.0));
        st[i] = normalized(add(m0[i], scale(pred, dt)));
    }
    auto fs = rhs(p, st);
    for (size_t i = 0; i < n; ++i) {
        V3 corr = add(add(scale(fs[i], 5.0 / 12.0), scale(p.f_n[i], 8.0 / 12.0)), scale(p.f_n1[i], -1.0 / 12.0));
        p.m[i] = normalized(add(m0[i], scale(corr, dt)));
    }
    p.time += dt;
    abm_push(p, fs, dt);
    return finish(p, dt, 0.0);
}

void report_to(const Report &r, double *out) {
    out[0] = r.time_seconds; out[1] = r.dt_used; out[2] = r.suggested_next_dt;
    out[3] = r.e_ex; out[4] = r.e_demag; out[5] = r.e_ext; out[6] = r.e_total;
    out[7] = r.max_h_eff; out[8] = r.max_h_demag; out[9] = r.max_rhs;
}

// ---------------------------------------------------------------------------
// Multilayer stack: layers on their own grids coupled by demag on a common
// convolution grid (crates/fullmag-engine/src/multilayer.rs:159-262,
// crates/fullmag-fdm-demag/src/{shifted_kernel.rs, transfer.rs,
// multiply.rs:25-112}), advanced by the runner's Heun
// (crates/fullmag-runner/src/multilayer_reference.rs:398-700).
// ---------------------------------------------------------------------------

// push_m, transfer.rs:18-121: each convolution cell averages the native cells
// it overlaps, weighted by the overlap volume; grids of equal size are copied.
std::vector<V3> push_m(const std::vector<V3> &m, const size_t nc[3], const double nd[3], const size_t cc[3],
                       const double cd[3]) {
    if (nc[0] == cc[0] && nc[1] == cc[1] && nc[2] == cc[2]) return m;
    std::vector<V3> out(cc[0] * cc[1] * cc[2], V3{0, 0, 0});
    for (size_t cz = 0; cz < cc[2]; ++cz)
        for (size_t cy = 0; cy < cc[1]; ++cy)
            for (size_t cx = 0; cx < cc[0]; ++cx) {
                const double lo[3] = {cx * cd[0], cy * cd[1], cz * cd[2]};
                const double hi[3] = {lo[0] + cd[0], lo[1] + cd[1], lo[2] + cd[2]};
                long a0[3], a1[3];
                for (int d = 0; d < 3; ++d) {
                    a0[d] = (long)std::floor(lo[d] / nd[d]);
                    a1[d] = (long)std::ceil(hi[d] / nd[d]);
                }
                double tot = 0.0;
                V3 acc{0, 0, 0};
                for (long z = a0[2]; z < a1[2]; ++z)
                    for (long y = a0[1]; y < a1[1]; ++y)
                        for (long x = a0[0]; x < a1[0]; ++x) {
                            if (x < 0 || y < 0 || z < 0 || x >= (long)nc[0] || y >= (long)nc[1] || z >= (long)nc[2])
                                continue;
                            const double n0[3] = {x * nd[0], y * nd[1], z * nd[2]};
                            double ov = 1.0;
                            for (int d = 0; d < 3; ++d)
                                ov *= std::max(std::min(hi[d], n0[d] + nd[d]) - std::max(lo[d], n0[d]), 0.0);
                            if (ov > 0.0) {
                                const V3 v = m[(size_t)x + nc[0] * ((size_t)y + nc[1] * (size_t)z)];
                                for (int k = 0; k < 3; ++k) acc[k] += v[k] * ov;
                                tot += ov;
                            }
                        }
                if (tot > 0.0) out[cx + cc[0] * (cy + cc[1] * cz)] = {acc[0] / tot, acc[1] / tot, acc[2] / tot};
            }
    return out;
}

// pull_h, transfer.rs:124-198: trilinear interpolation of the convolution
// grid at each native cell centre, indices clamped at the edges.
std::vector<V3> pull_h(const std::vector<V3> &h, const size_t cc[3], const double cd[3], const size_t nc[3],
                       const double nd[3]) {
    if (nc[0] == cc[0] && nc[1] == cc[1] && nc[2] == cc[2]) return h;
    std::vector<V3> out(nc[0] * nc[1] * nc[2]);
    for (size_t z = 0; z < nc[2]; ++z)
        for (size_t y = 0; y < nc[1]; ++y)
            for (size_t x = 0; x < nc[0]; ++x) {
                const double f[3] = {(x + 0.5) * nd[0] / cd[0] - 0.5, (y + 0.5) * nd[1] / cd[1] - 0.5,
                                     (z + 0.5) * nd[2] / cd[2] - 0.5};
                long b[3];
                double w[3];
                for (int d = 0; d < 3; ++d) {
                    b[d] = (long)std::floor(f[d]);
                    w[d] = f[d] - std::floor(f[d]);
                }
                V3 r{0, 0, 0};
                for (int dz = 0; dz < 2; ++dz)
                    for (int dy = 0; dy < 2; ++dy)
                        for (int dx = 0; dx < 2; ++dx) {
                            const long ix = std::min(std::max(b[0] + dx, 0L), (long)cc[0] - 1);
                            const long iy = std::min(std::max(b[1] + dy, 0L), (long)cc[1] - 1);
                            const long iz = std::min(std::max(b[2] + dz, 0L), (long)cc[2] - 1);
                            const double ww = (dx == 0 ? 1.0 - w[0] : w[0]) * (dy == 0 ? 1.0 - w[1] : w[1]) *
                                              (dz == 0 ? 1.0 - w[2] : w[2]);
                            const V3 v = h[(size_t)ix + cc[0] * ((size_t)iy + cc[1] * (size_t)iz)];
                            for (int k = 0; k < 3; ++k) r[k] += v[k] * ww;
                        }
                out[x + nc[0] * (y + nc[1] * z)] = r;
            }
    return out;
}

struct Stack {
    std::vector<Problem *> layer;  // exchange, Zeeman, mask and material of each layer
    std::vector<double> oz;        // native origin z per layer
    size_t cc[3];                  // common convolution grid
    double csz[3];                 // its cell size
    bool demag = true;
    size_t px, py, pz;
    std::vector<std::array<std::vector<cd>, 6>> kern;  // pair src * L + dst
    double time = 0.0;
};

// Kernel pairs in the runner's order (multilayer_reference.rs:398-439): self
// kernel for src == dst, shifted kernel with dst.z - src.z otherwise.
void stack_build(Stack &s) {
    s.px = 2 * s.cc[0];
    s.py = 2 * s.cc[1];
    s.pz = 2 * s.cc[2];
    const size_t L = s.layer.size(), plen = s.px * s.py * s.pz;
    s.kern.assign(L * L, {});
    if (!s.demag) return;
    for (size_t a = 0; a < L; ++a)
        for (size_t b = 0; b < L; ++b) {
            std::vector<double> nk[6];
            if (a == b) compute_newell_kernels(s.cc[0], s.cc[1], s.cc[2], s.csz[0], s.csz[1], s.csz[2], nk);
            else
                compute_newell_kernels(s.cc[0], s.cc[1], s.cc[2], s.csz[0], s.csz[1], s.csz[2], nk, s.oz[b] - s.oz[a],
                                       true);
            for (int c = 0; c < 6; ++c) {
                auto &k = s.kern[a * L + b][c];
                k.resize(plen);
                for (size_t i = 0; i < plen; ++i) k[i] = cd(nk[c][i], 0.0);
                fft3_core(k.data(), (int)s.px, (int)s.py, (int)s.pz, false);
            }
        }
}

// MultilayerDemagRuntime::compute_demag_fields (multilayer.rs:159-262): the
// fields of all layers, not yet zeroed outside their masks.
std::vector<std::vector<V3>> stack_demag(Stack &s, const std::vector<std::vector<V3>> &ms) {
    const size_t L = s.layer.size(), px = s.px, py = s.py, pz = s.pz, plen = px * py * pz;
    std::vector<std::vector<V3>> out(L);
    if (!s.demag) {
        for (size_t l = 0; l < L; ++l) out[l].assign(s.layer[l]->n, V3{0, 0, 0});
        return out;
    }
    std::vector<std::array<std::vector<cd>, 3>> mf(L), hf(L);
    for (size_t l = 0; l < L; ++l) {
        const Problem &p = *s.layer[l];
        const size_t nc[3] = {p.nx, p.ny, p.nz};
        const double nd[3] = {p.dx, p.dy, p.dz};
        const std::vector<V3> cm = push_m(ms[l], nc, nd, s.cc, s.csz);
        for (int c = 0; c < 3; ++c) {
            mf[l][c].assign(plen, cd(0, 0));
            hf[l][c].assign(plen, cd(0, 0));
        }
        for (size_t z = 0; z < s.cc[2]; ++z)
            for (size_t y = 0; y < s.cc[1]; ++y)
                for (size_t x = 0; x < s.cc[0]; ++x) {
                    const V3 v = cm[x + s.cc[0] * (y + s.cc[1] * z)];
                    const size_t d = padded_index(px, py, x, y, z);
                    for (int c = 0; c < 3; ++c) mf[l][c][d] = cd(v[c] * p.ms, 0.0);
                }
        for (int c = 0; c < 3; ++c) fft3_core(mf[l][c].data(), (int)px, (int)py, (int)pz, false);
    }
    // accumulate_tensor_convolution (multiply.rs:25-112) over the pairs
    for (size_t a = 0; a < L; ++a)
        for (size_t b = 0; b < L; ++b) {
            const auto &K = s.kern[a * L + b];
            auto &H = hf[b];
            const auto &M = mf[a];
#pragma omp parallel for schedule(static)
            for (long il = 0; il < (long)plen; ++il) {
                const size_t i = (size_t)il;
                const cd mx = M[0][i], my = M[1][i], mz = M[2][i];
                H[0][i] += K[0][i] * mx + K[3][i] * my + K[4][i] * mz;
                H[1][i] += K[3][i] * mx + K[1][i] * my + K[5][i] * mz;
                H[2][i] += K[4][i] * mx + K[5][i] * my + K[2][i] * mz;
            }
        }
    const double norm_ = 1.0 / (double)plen;
    for (size_t l = 0; l < L; ++l) {
        for (int c = 0; c < 3; ++c) {
            for (auto &v : hf[l][c]) v = -v;
            fft3_core(hf[l][c].data(), (int)px, (int)py, (int)pz, true);
        }
        std::vector<V3> ch(s.cc[0] * s.cc[1] * s.cc[2]);
        for (size_t z = 0; z < s.cc[2]; ++z)
            for (size_t y = 0; y < s.cc[1]; ++y)
                for (size_t x = 0; x < s.cc[0]; ++x) {
                    const size_t i = padded_index(px, py, x, y, z);
                    ch[x + s.cc[0] * (y + s.cc[1] * z)] = {hf[l][0][i].real() * norm_, hf[l][1][i].real() * norm_,
                                                         hf[l][2][i].real() * norm_};
                }
        const Problem &p = *s.layer[l];
        const size_t nc[3] = {p.nx, p.ny, p.nz};
        const double nd[3] = {p.dx, p.dy, p.dz};
        out[l] = pull_h(ch, s.cc, s.csz, nc, nd);
    }
    return out;
}

struct StackEval {
    std::vector<std::vector<V3>> h_ex, h_d, h_ext, h_eff, rhs;
    Report r{};
};

// llg_rhs_multilayer / observe_multilayer (multilayer_reference.rs:441-651):
// per layer H_eff = (H_ex + H_d) + H_ext with H_d and H_ext zeroed outside the
// layer's mask, the layer's LLG right-hand side, and the stack's energies
// (sums) and maxima.
StackEval stack_eval(Stack &s, const std::vector<std::vector<V3>> &ms, bool want_stats) {
    const size_t L = s.layer.size();
    StackEval e;
    e.h_d = stack_demag(s, ms);
    e.h_ex.resize(L);
    e.h_ext.resize(L);
    e.h_eff.resize(L);
    e.rhs.resize(L);
    double ex = 0, ed = 0, eext = 0, mh = 0, md = 0, mr = 0;
    for (size_t l = 0; l < L; ++l) {
        Problem &p = *s.layer[l];
        const auto &m = ms[l];
        e.h_ex[l] = p.exchange ? exchange_field(p, m) : std::vector<V3>(p.n, V3{0, 0, 0});
        e.h_ext[l].resize(p.n);
        e.h_eff[l].resize(p.n);
        e.rhs[l].resize(p.n);
        const V3 ext = p.has_ext ? p.hext : V3{0, 0, 0};
        for (size_t i = 0; i < p.n; ++i) {
            if (!p.active(i)) e.h_d[l][i] = {0, 0, 0};
            e.h_ext[l][i] = p.active(i) ? ext : V3{0, 0, 0};
            e.h_eff[l][i] = p.active(i) ? add(add(e.h_ex[l][i], e.h_d[l][i]), e.h_ext[l][i]) : V3{0, 0, 0};
            e.rhs[l][i] = llg_rhs_from_field(p, m[i], e.h_eff[l][i]);
        }
        if (!want_stats) continue;
        const double v = p.dx * p.dy * p.dz;
        for (size_t i = 0; i < p.n; ++i) {
            if (p.exchange && p.active(i)) ex += -0.5 * MU0 * p.ms * dot(m[i], e.h_ex[l][i]) * v;
            if (s.demag) ed += -0.5 * MU0 * p.ms * dot(m[i], e.h_d[l][i]) * v;
            if (p.has_ext) eext += -MU0 * p.ms * dot(m[i], e.h_ext[l][i]) * v;
        }
        mh = std::max(mh, max_norm(e.h_eff[l]));
        md = std::max(md, max_norm(e.h_d[l]));
        mr = std::max(mr, max_norm(e.rhs[l]));
    }
    e.r.e_ex = ex;
    e.r.e_demag = ed;
    e.r.e_ext = eext;
    e.r.e_total = ex + ed + eext;
    e.r.max_h_eff = mh;
    e.r.max_h_demag = md;
    e.r.max_rhs = mr;
    e.r.time_seconds = s.time;
    return e;
}

// step_multilayer (multilayer_reference.rs:543-589): Heun over all layers.
Report stack_heun(Stack &s, double dt) {
    const size_t L = s.layer.size();
    std::vector<std::vector<V3>> m0(L), pred(L);
    for (size_t l = 0; l < L; ++l) m0[l] = s.layer[l]->m;
    const StackEval e1 = stack_eval(s, m0, false);
    for (size_t l = 0; l < L; ++l) {
        pred[l].resize(m0[l].size());
        for (size_t i = 0; i < m0[l].size(); ++i) pred[l][i] = normalized(add(m0[l][i], scale(e1.rhs[l][i], dt)));
    }
    const StackEval e2 = stack_eval(s, pred, false);
    for (size_t l = 0; l < L; ++l) {
        Problem &p = *s.layer[l];
        for (size_t i = 0; i < p.n; ++i)
            p.m[i] = normalized(add(m0[l][i], scale(add(e1.rhs[l][i], e2.rhs[l][i]), 0.5 * dt)));
        p.time += dt;
    }
    s.time += dt;
    std::vector<std::vector<V3>> m1(L);
    for (size_t l = 0; l < L; ++l) m1[l] = s.layer[l]->m;
    Report r = stack_eval(s, m1, true).r;
    r.dt_used = dt;
    return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points (ctypes from tests/ and bench.py's cpu_baseline leg)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Discrete adjoint of one Heun step (extension; no counterpart in the
// reference).  With f(m, H) = -gb (p m x H + alpha m x (m x H)):
//   (df/dm)^T mu = -gb (p H x mu + alpha ((m.H) mu + H (m.mu) - 2 m (H.mu)))
//   (df/dH)^T mu = -gb (p mu x m + alpha (m (m.mu) - |m|^2 mu))
// normalisation n(v) = v/|v|: (dn/dv)^T l = (l - n (n.l)) / |v| (0 at v = 0);
// H(m) = (2A/(mu0 Ms)) lap(m) - N * (Ms m) + H_ext on active cells, so
// (dH/dm)^T g = lap(2A/(mu0 Ms) g) + Ms (-N * g)  (lap and the convolution
// are symmetric), dH_i/dMs_i = -H_ex,i/Ms_i for exchange and
// d(-N*(Ms m))/dMs_j = -N * (e_j m_j) for demag, dH/dH_ext = I.
// ---------------------------------------------------------------------------
inline V3 torque_adj_m(const Problem &p, const V3 &m, const V3 &h, const V3 &mu) {
    const double gb = p.gamma / (1.0 + p.alpha * p.alpha);
    V3 t = p.precession ? cross(h, mu) : V3{0, 0, 0};
    const double mh = dot(m, h), mm = dot(m, mu), hm = dot(h, mu);
    for (int k = 0; k < 3; ++k) t[k] += p.alpha * (mh * mu[k] + h[k] * mm - 2.0 * m[k] * hm);
    return scale(t, -gb);
}

inline V3 torque_adj_h(const Problem &p, const V3 &m, const V3 &mu) {
    const double gb = p.gamma / (1.0 + p.alpha * p.alpha);
    V3 t = p.precession ? cross(mu, m) : V3{0, 0, 0};
    const double mm = dot(m, mu), m2 = dot(m, m);
    for (int k = 0; k < 3; ++k) t[k] += p.alpha * (m[k] * mm - m2 * mu[k]);
    return scale(t, -gb);
}

inline V3 norm_adj(const V3 &v, const V3 &l) {
    const double nv = norm(v);
    if (nv <= 0.0) return {0, 0, 0};
    const V3 y = scale(v, 1.0 / nv);
    return scale(sub(l, scale(y, dot(y, l))), 1.0 / nv);
}

// Adds (dH/dm)^T g at state m to lam, and the parameter gradients.
void field_adjoint(Problem &p, const std::vector<V3> &m, const std::vector<V3> &h_ex, const std::vector<V3> &g,
                   std::vector<V3> &lam, std::vector<double> &grad_ms, V3 &grad_hext) {
    const size_t n = p.n;
    if (p.exchange) {
        std::vector<V3> q(n);
        for (size_t i = 0; i < n; ++i)
            q[i] = p.active(i) ? scale(g[i], (p.use_lut() ? 2.0 : 2.0 * p.a) / (MU0 * p.ms_at(i))) : V3{0, 0, 0};
        std::vector<V3> lq = lap_field(p, q);
        for (size_t i = 0; i < n; ++i) {
            lam[i] = add(lam[i], lq[i]);
            if (p.active(i)) grad_ms[i] += -dot(g[i], h_ex[i]) / p.ms_at(i);
        }
    }
    if (p.demag) {
        std::vector<V3> w = demag_apply(p, g, false);
        for (size_t i = 0; i < n; ++i) {
            if (!p.active(i)) continue;
            lam[i] = add(lam[i], scale(w[i], p.ms_at(i)));
            grad_ms[i] += dot(m[i], w[i]);
        }
    }
    if (p.has_ext)
        for (size_t i = 0; i < n; ++i)
            if (p.active(i)) grad_hext = add(grad_hext, g[i]);
}

// lam0 = (d m1 / d m0)^T lam1 for m1 = heun(m0); grad_ms (n) and grad_hext
// (3) accumulate d<lam1, m1>/dMs_i and d<lam1, m1>/dH_ext.
void heun_adjoint(Problem &p, double dt, const std::vector<V3> &m0, const std::vector<V3> &lam1,
                  std::vector<V3> &lam0, std::vector<double> &grad_ms, V3 &grad_hext) {
    const size_t n = p.n;
    Eval e0 = full_eval(p, m0, false);
    std::vector<V3> v(n), mt(n);
    for (size_t i = 0; i < n; ++i) {
        v[i] = add(m0[i], scale(e0.rhs[i], dt));
        mt[i] = normalized(v[i]);
    }
    Eval et = full_eval(p, mt, false);
    std::vector<V3> u(n);
    for (size_t i = 0; i < n; ++i) u[i] = add(m0[i], scale(add(e0.rhs[i], et.rhs[i]), 0.5 * dt));
    // backward through m1 = n(u), u = m0 + dt/2 (k1 + k2)
    std::vector<V3> lam_u(n), lam_k1(n), lam_mt(n), g(n);
    for (size_t i = 0; i < n; ++i) {
        lam_u[i] = norm_adj(u[i], lam1[i]);
        const V3 mu = scale(lam_u[i], 0.5 * dt);  // lam_k2
        lam_mt[i] = torque_adj_m(p, mt[i], et.h_eff[i], mu);
        g[i] = p.active(i) ? torque_adj_h(p, mt[i], mu) : V3{0, 0, 0};
    }
    field_adjoint(p, mt, et.h_ex, g, lam_mt, grad_ms, grad_hext);
    // mt = n(v), v = m0 + dt k1
    lam0.assign(n, V3{0, 0, 0});
    for (size_t i = 0; i < n; ++i) {
        const V3 lv = norm_adj(v[i], lam_mt[i]);
        lam0[i] = add(lam_u[i], lv);
        const V3 mu = add(scale(lam_u[i], 0.5 * dt), scale(lv, dt));  // lam_k1
        lam0[i] = add(lam0[i], torque_adj_m(p, m0[i], e0.h_eff[i], mu));
        g[i] = p.active(i) ? torque_adj_h(p, m0[i], mu) : V3{0, 0, 0};
    }
    field_adjoint(p, m0, e0.h_ex, g, lam0, grad_ms, grad_hext);
    for (size_t i = 0; i < n; ++i)
        if (!p.active(i)) lam0[i] = {0, 0, 0};
}

extern "C" {

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

double oracle_newell_f(double x, double y, double z) { return newell_f(x, y, z); }
double oracle_newell_g(double x, double y, double z) { return newell_g(x, y, z); }

// out: 6 arrays of (2nx)(2ny)(2nz) doubles, order xx,yy,zz,xy,xz,yz
void oracle_newell_kernels(uint64_t nx, uint64_t ny, uint64_t nz, double dx, double dy, double dz, double *out) {
    std::vector<double> k[6];
    compute_newell_kernels(nx, ny, nz, dx, dy, dz, k);
    size_t plen = 8 * nx * ny * nz;
    for (int c = 0; c < 6; ++c) std::memcpy(out + c * plen, k[c].data(), plen * sizeof(double));
}

// lib.rs:511-610.  thin_film_2d != 0 requires nz == 1 and produces the 2-D
// (2nx)(2ny) spectra; otherwise the full 3-D (2nx)(2ny)(2nz) spectra.
// out: 6 interleaved complex arrays [re,im,...] back to back.
int oracle_newell_spectra(uint64_t nx, uint64_t ny, uint64_t nz, double dx, double dy, double dz,
                          int thin_film_2d, double *out) {
    if (thin_film_2d && nz != 1) return -1;
    std::vector<double> k[6];
    compute_newell_kernels(nx, ny, nz, dx, dy, dz, k);
    size_t px = 2 * nx, py = 2 * ny, pz = 2 * nz;
    if (thin_film_2d) {
        size_t plane = px * py;
        for (int c = 0; c < 6; ++c) {
            std::vector<cd> b(plane);
            for (size_t y = 0; y < py; ++y)
                for (size_t x = 0; x < px; ++x) b[x + px * y] = cd(k[c][padded_index(px, py, x, y, 0)], 0.0);
            fft3_core(b.data(), (int)px, (int)py, 1, false);
            std::memcpy(out + c * 2 * plane, b.data(), plane * sizeof(cd));
        }
    } else {
        size_t plen = px * py * pz;
        for (int c = 0; c < 6; ++c) {
            std::vector<cd> b(plen);
            for (size_t i = 0; i < plen; ++i) b[i] = cd(k[c][i], 0.0);
            fft3_core(b.data(), (int)px, (int)py, (int)pz, false);
            std::memcpy(out + c * 2 * plen, b.data(), plen * sizeof(cd));
        }
    }
    return 0;
}

void oracle_fft(double *interleaved, uint64_t n, int inverse) {
    fft1(reinterpret_cast<cd *>(interleaved), (int)n, inverse != 0);
}

void oracle_fft3(double *interleaved, uint64_t nx, uint64_t ny, uint64_t nz, int inverse) {
    fft3_core(reinterpret_cast<cd *>(interleaved), (int)nx, (int)ny, (int)nz, inverse != 0);
}

// fullmag-plan/src/lib.rs:1927-1957
void oracle_random_unit_vectors(uint64_t seed, uint64_t count, double *out) {
    uint64_t state = seed;
    for (uint64_t i = 0; i < count; ++i) {
        double c[3];
        for (;;) {
            for (int k = 0; k < 3; ++k) {
                state ^= state << 13;
                state ^= state >> 7;
                state ^= state << 17;
                c[k] = ((double)state / (double)UINT64_MAX) * 2.0 - 1.0;
            }
            double nrm = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            if (nrm > 1e-10) {
                c[0] /= nrm; c[1] /= nrm; c[2] /= nrm;
                break;
            }
        }
        out[3 * i + 0] = c[0]; out[3 * i + 1] = c[1]; out[3 * i + 2] = c[2];
    }
}

// params: [dx,dy,dz, Ms, A, alpha, gamma, hx,hy,hz, max_error, dt_min, dt_max, headroom]
// flags:  [integrator, precession, exchange, demag, has_ext]
void *oracle_problem_create(uint64_t nx, uint64_t ny, uint64_t nz, const double *params, const int *flags,
                            const uint8_t *active_mask, const double *ms_cell, const double *m0_aos) {
    auto *p = new Problem();
    p->nx = nx; p->ny = ny; p->nz = nz; p->n = nx * ny * nz;
    p->dx = params[0]; p->dy = params[1]; p->dz = params[2];
    p->ms = params[3]; p->a = params[4]; p->alpha = params[5]; p->gamma = params[6];
    p->hext = {params[7], params[8], params[9]};
    // AdaptiveStepConfig defaults lib.rs:217-226
    p->max_error = params[10] > 0 ? params[10] : 1e-5;
    p->dt_min = params[11] > 0 ? params[11] : 1e-18;
    p->dt_max = params[12] > 0 ? params[12] : 1e-10;
    p->headroom = params[13] > 0 ? params[13] : 0.8;
    p->integrator = flags[0];
    p->precession = flags[1] != 0;
    p->exchange = flags[2] != 0;
    p->demag = flags[3] != 0;
    p->has_ext = flags[4] != 0;
    if (active_mask) p->mask.assign(active_mask, active_mask + p->n);
    if (ms_cell) p->ms_cell.assign(ms_cell, ms_cell + p->n);
    p->m.resize(p->n);
    // ExchangeLlgState::new normalises; new_state zeroes inactive cells (lib.rs:1009-1019)
    for (size_t i = 0; i < p->n; ++i) {
        V3 v = {m0_aos[3 * i], m0_aos[3 * i + 1], m0_aos[3 * i + 2]};
        p->m[i] = p->active(i) ? normalized(v) : V3{0, 0, 0};
    }
    build_workspace(*p);
    return p;
}

void oracle_problem_destroy(void *h) { delete static_cast<Problem *>(h); }

// report (10 doubles): time, dt_used, suggested_next_dt (0 = none),
// E_ex, E_demag, E_ext, E_total, max|H_eff|, max|H_demag|, max|dm/dt|
int oracle_problem_step(void *h, double dt, double *report) {
    auto *p = static_cast<Problem *>(h);
    if (!(dt > 0.0)) return -1;
    Report r{};
    switch (p->integrator) {
        case RK4: r = rk4_step(*p, dt); break;
        case RK23: r = rk23_step(*p, dt); break;
        case DP45: r = dp45_step(*p, dt); break;
        case ABM3: r = abm3_step(*p, dt); break;
        case HEUN:
        default: r = heun_step(*p, dt); break;
    }
    report_to(r, report);
    return 0;
}

// Region ids (one per cell, < 256) and the 256 x 256 exchange table A_ij.
// lut NULL: the default table of native/backends/fdm/src/api.cpp:215-231,
// A on the diagonal and 0 elsewhere.  region NULL: back to uniform exchange.
int oracle_problem_set_regions(void *h, const uint32_t *region, const double *lut) {
    auto *p = static_cast<Problem *>(h);
    const size_t R = Problem::kRegions;
    if (!region) {
        p->region.clear();
        p->lut.clear();
        return 0;
    }
    for (size_t i = 0; i < p->n; ++i)
        if (region[i] >= R) return -1;
    p->region.assign(region, region + p->n);
    p->lut.assign(R * R, 0.0);
    if (lut) std::memcpy(p->lut.data(), lut, R * R * sizeof(double));
    else
        for (size_t r = 0; r < R; ++r) p->lut[r * R + r] = p->a;
    return 0;
}

// Error norm (lib.rs:1653-1674) of one attempted RK23 or DP45 step of size
// dt from the current state, which is left unchanged.  The first slope is
// evaluated at the state.  Returns -1 for other integrators.
// Sub-cell boundary tiers: face (6 n: xp, xm, yp, ym, zp, zm) and delta
// (6 n) may be null; the correction arrays are target (count), source
// (count * stencil) and tensor (count * stencil * 6).
int oracle_problem_set_tiers(void *h, int tier, double phi_floor, double delta_min, const double *phi,
                             const double *face, const double *delta, uint32_t count, uint32_t stencil,
                             const int32_t *target, const int32_t *source, const double *tensor) {
    Problem *p = static_cast<Problem *>(h);
    const size_t n = p->n;
    p->tier = tier;
    p->phi_floor = phi_floor > 0.0 ? phi_floor : 0.05;
    p->delta_min = delta_min > 0.0 ? delta_min : 0.1 * std::min({p->dx, p->dy, p->dz});
    p->phi.assign(phi, phi + n);
    for (int k = 0; k < 6; ++k) {
        p->face[k].clear();
        p->delta[k].clear();
        if (face) p->face[k].assign(face + k * n, face + (k + 1) * n);
        if (delta) p->delta[k].assign(delta + k * n, delta + (k + 1) * n);
    }
    p->has_fsal = false;
    p->k_fsal.clear();
    p->abm_count = 0;
    p->corr_stencil = 0;
    p->corr_t.clear();
    p->corr_s.clear();
    p->corr_n.clear();
    if (count > 0 && stencil > 0 && target && source && tensor) {
        p->corr_stencil = stencil;
        p->corr_t.assign(target, target + count);
        p->corr_s.assign(source, source + (size_t)count * stencil);
        p->corr_n.assign(tensor, tensor + (size_t)count * stencil * 6);
    }
    return 0;
}

double oracle_problem_trial_error(void *h, double dt) {
    auto *p = static_cast<Problem *>(h);
    std::vector<V3> st, k7;
    if (p->integrator == RK23) return rk23_attempt(*p, p->m, dt, st);
    if (p->integrator == DP45) return dp45_attempt(*p, p->m, rhs(*p, p->m), dt, st, k7);
    return -1.0;
}

void oracle_problem_get_m(void *h, double *out) {
    auto *p = static_cast<Problem *>(h);
    for (size_t i = 0; i < p->n; ++i)
        for (int k = 0; k < 3; ++k) out[3 * i + k] = p->m[i][k];
}

// set_magnetization (lib.rs:739-752) normalises and keeps FSAL / ABM history.
void oracle_problem_set_m(void *h, const double *in) {
    auto *p = static_cast<Problem *>(h);
    for (size_t i = 0; i < p->n; ++i) {
        V3 v = {in[3 * i], in[3 * i + 1], in[3 * i + 2]};
        p->m[i] = p->active(i) ? normalized(v) : V3{0, 0, 0};
    }
}

// Raw state (no normalisation; inactive cells zero): finite-difference
// checks of the adjoint perturb m0 off the unit sphere.
void oracle_problem_set_m_raw(void *h, const double *in) {
    auto *p = static_cast<Problem *>(h);
    for (size_t i = 0; i < p->n; ++i)
        p->m[i] = p->active(i) ? V3{in[3 * i], in[3 * i + 1], in[3 * i + 2]} : V3{0, 0, 0};
}

double oracle_problem_time(void *h) { return static_cast<Problem *>(h)->time; }

// fields: 5 AoS arrays of 3N doubles back to back: H_ex, H_demag, H_ext, H_eff, rhs.
// scalars (10 doubles, same layout as the step report with time/dt fields zero).
void oracle_problem_observe(void *h, double *fields, double *scalars) {
    auto *p = static_cast<Problem *>(h);
    Eval e = full_eval(*p, p->m, true);
    const std::vector<V3> *arr[5] = {&e.h_ex, &e.h_d, &e.h_ext, &e.h_eff, &e.rhs};
    if (fields)
        for (int a = 0; a < 5; ++a)
            for (size_t i = 0; i < p->n; ++i)
                for (int k = 0; k < 3; ++k) fields[a * 3 * p->n + 3 * i + k] = (*arr[a])[i][k];
    e.r.time_seconds = p->time;
    e.r.dt_used = 0; e.r.suggested_next_dt = 0;
    if (scalars) report_to(e.r, scalars);
}

// Adjoint of one Heun step from state m0 (AoS 3N, not the problem's state).
// lam1: AoS 3N cotangent of the state after the step; lam0_out: AoS 3N;
// grad_ms (N) and grad_hext (3) are accumulated into.
int oracle_problem_heun_adjoint(void *h, double dt, const double *m0, const double *lam1, double *lam0_out,
                                double *grad_ms, double *grad_hext) {
    auto *p = static_cast<Problem *>(h);
    const size_t n = p->n;
    std::vector<V3> vm0(n), vl1(n), vl0;
    for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            vm0[i][k] = m0[3 * i + k];
            vl1[i][k] = lam1[3 * i + k];
        }
    std::vector<double> gms(n, 0.0);
    V3 gh = {0, 0, 0};
    heun_adjoint(*p, dt, vm0, vl1, vl0, gms, gh);
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) lam0_out[3 * i + k] = vl0[i][k];
        if (grad_ms) grad_ms[i] += gms[i];
    }
    if (grad_hext)
        for (int k = 0; k < 3; ++k) grad_hext[k] += gh[k];
    return 0;
}

// Shifted Newell kernel (newell.rs:528-651) on the padded (2nx, 2ny, 2nz)
// grid, out[6][2nz][2ny][2nx].
void oracle_newell_kernels_shifted(uint64_t nx, uint64_t ny, uint64_t nz, double dx, double dy, double dz,
                                   double z_offset, double *out) {
    std::vector<double> k[6];
    compute_newell_kernels(nx, ny, nz, dx, dy, dz, k, z_offset, true);
    const size_t plen = 8 * nx * ny * nz;
    for (int c = 0; c < 6; ++c) std::memcpy(out + c * plen, k[c].data(), plen * sizeof(double));
}

// Grid transfer operators (AoS vectors): native -> convolution and back.
void oracle_push_m(const double *m, const uint64_t *nc, const double *nd, const uint64_t *cc, const double *cdd,
                   double *out) {
    const size_t n[3] = {nc[0], nc[1], nc[2]}, c[3] = {cc[0], cc[1], cc[2]};
    std::vector<V3> v(n[0] * n[1] * n[2]);
    std::memcpy(v.data(), m, v.size() * sizeof(V3));
    const std::vector<V3> r = push_m(v, n, nd, c, cdd);
    std::memcpy(out, r.data(), r.size() * sizeof(V3));
}
void oracle_pull_h(const double *h, const uint64_t *cc, const double *cdd, const uint64_t *nc, const double *nd,
                   double *out) {
    const size_t n[3] = {nc[0], nc[1], nc[2]}, c[3] = {cc[0], cc[1], cc[2]};
    std::vector<V3> v(c[0] * c[1] * c[2]);
    std::memcpy(v.data(), h, v.size() * sizeof(V3));
    const std::vector<V3> r = pull_h(v, c, cdd, n, nd);
    std::memcpy(out, r.data(), r.size() * sizeof(V3));
}

// A stack of layer problems (created with demag off; the stack owns the
// coupling).  The problems stay owned by the caller and must outlive it.
void *oracle_stack_create(uint32_t count, void *const *layers, const double *origins, const uint64_t *common,
                          const double *conv_cell, int demag) {
    auto *s = new Stack();
    for (uint32_t l = 0; l < count; ++l) {
        s->layer.push_back(static_cast<Problem *>(layers[l]));
        s->oz.push_back(origins[3 * l + 2]);
    }
    for (int d = 0; d < 3; ++d) {
        s->cc[d] = common[d];
        s->csz[d] = conv_cell[d];
    }
    s->demag = demag != 0;
    stack_build(*s);
    return s;
}
void oracle_stack_destroy(void *h) { delete static_cast<Stack *>(h); }
int oracle_stack_step(void *h, double dt, double *report) {
    auto *s = static_cast<Stack *>(h);
    if (!(dt > 0.0)) return -1;
    const Report r = stack_heun(*s, dt);
    if (report) report_to(r, report);
    return 0;
}
// fields: per layer in order, 5 AoS arrays (H_ex, H_demag, H_ext, H_eff, rhs)
// of 3 n_l doubles; scalars: the step report layout.
void oracle_stack_observe(void *h, double *fields, double *scalars) {
    auto *s = static_cast<Stack *>(h);
    std::vector<std::vector<V3>> ms;
    for (Problem *p : s->layer) ms.push_back(p->m);
    const StackEval e = stack_eval(*s, ms, true);
    size_t off = 0;
    for (size_t l = 0; l < s->layer.size(); ++l) {
        const size_t n = s->layer[l]->n;
        const std::vector<V3> *arr[5] = {&e.h_ex[l], &e.h_d[l], &e.h_ext[l], &e.h_eff[l], &e.rhs[l]};
        if (fields)
            for (int a = 0; a < 5; ++a) {
                std::memcpy(fields + off, arr[a]->data(), 3 * n * sizeof(double));
                off += 3 * n;
            }
    }
    if (scalars) report_to(e.r, scalars);
}

}  // extern "C"
