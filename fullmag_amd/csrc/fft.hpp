// fft.hpp -- line FFTs staged in LDS for the demag convolution.
//
// A "line" is one 1-D transform of length n held in LDS.  A workgroup holds
// several lines; each line is worked on by `threads` = n / (smallest radix)
// lanes, one butterfly per lane per pass, so every pass is read-all /
// barrier / write-all in place and a line needs only n complex slots of LDS.
//
// The passes are Stockham autosort DIT (natural-order output, no bit
// reversal): for radix R and span ns (product of the radices already
// applied) butterfly j reads x[j + r n/R], multiplies by w^(r k n/(ns R))
// with k = j mod ns, applies the R-point DFT and writes x[(j/ns) ns R + k + r ns].
// Twiddles come from a per-length table w[t] = exp(-2 pi i t / n) rounded
// from long double on the host, so FP64 transforms carry only butterfly
// rounding.  Forward is unnormalised exp(-i...), inverse unnormalised
// exp(+i...), the convention of crates/fullmag-engine/src/lib.rs:343-386.
#pragma once

#include "common.hpp"

#include <vector>

namespace fm {

constexpr int kMaxPass = 16;

struct LinePlan {
    int n = 0;        // line length
    int npass = 0;    // number of Stockham passes
    int threads = 0;  // lanes per line (each holds up to 16 complex values)
    int swz = 0;      // 1: LDS slots XOR-swizzled within 16-element blocks
    int radix[kMaxPass] = {};
};

// LDS slot of line element i.  With swz set (n a multiple of 16) element
// 16b + c lives at 16b + (c ^ (b & 15)): the stride-16 writes of the first
// Stockham pass then spread over 32 banks instead of 2.  Bijective within
// each 16-element block.
__host__ __device__ __forceinline__ int lslot(const LinePlan &lp, int i) { return lp.swz ? (i ^ ((i >> 4) & 15)) : i; }

// Chooses radices from {16, 8, 4, 2, 3, 5, 7, 11, 13}, plus any larger prime
// factor of n as a direct-sum pass, and the lanes per line: the fewest lanes
// such that no lane holds more than 16 complex values in any pass, or
// `min_lanes` if larger (capped at the largest pass's butterfly count);
// power-of-two radices are capped at max_radix (16, 8 or 4).  Returns false
// (with `err`) for lengths needing more than kMaxPass passes or 1024 lanes.
bool plan_line(int n, LinePlan &lp, std::string &err, int min_lanes = 1, int max_radix = 16);

// w[t] = exp(-2 pi i t / n), t in [0, n), correctly rounded from long double.
std::vector<cplx<double>> twiddle_table(int n);

// ---------------------------------------------------------------------------
// Device butterflies
// ---------------------------------------------------------------------------

// cos/sin of 2 pi m / 16, m = 0..15 (exact decimal expansions).
#define FM_C1 0.92387953251128675613
#define FM_S1 0.38268343236508977173
#define FM_H 0.70710678118654752440

template <typename T>
__device__ __forceinline__ cplx<T> w16(int m, bool inv) {
    // returns exp(-+ 2 pi i m / 16) (forward: minus sign)
    const T c[16] = {T(1), T(FM_C1), T(FM_H), T(FM_S1), T(0), T(-FM_S1), T(-FM_H), T(-FM_C1),
                     T(-1), T(-FM_C1), T(-FM_H), T(-FM_S1), T(0), T(FM_S1), T(FM_H), T(FM_C1)};
    const T s[16] = {T(0), T(FM_S1), T(FM_H), T(FM_C1), T(1), T(FM_C1), T(FM_H), T(FM_S1),
                     T(0), T(-FM_S1), T(-FM_H), T(-FM_C1), T(-1), T(-FM_C1), T(-FM_H), T(-FM_S1)};
    m &= 15;
    return {c[m], inv ? s[m] : -s[m]};
}

template <typename T>
__device__ __forceinline__ void dft2(cplx<T> &a, cplx<T> &b) {
    cplx<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// In-place 4-point DFT of (v[0], v[s], v[2s], v[3s]).
template <typename T>
__device__ __forceinline__ void dft4(cplx<T> &a0, cplx<T> &a1, cplx<T> &a2, cplx<T> &a3, bool inv) {
    cplx<T> s0 = cadd(a0, a2), d0 = csub(a0, a2);
    cplx<T> s1 = cadd(a1, a3), d1 = csub(a1, a3);
    // forward: -i * d1 ; inverse: +i * d1
    cplx<T> jd1 = inv ? cplx<T>{-d1.y, d1.x} : cplx<T>{d1.y, -d1.x};
    a0 = cadd(s0, s1);
    a2 = csub(s0, s1);
    a1 = cadd(d0, jd1);
    a3 = csub(d0, jd1);
}

// 8-point DFT: inner 4-point DFTs over even/odd samples, twiddle w8^(r2 q1),
// outer 2-point DFTs: X[q1 + 4 q2].
template <typename T>
__device__ __forceinline__ void dft8(cplx<T> *v, bool inv) {
    cplx<T> u0[4] = {v[0], v[2], v[4], v[6]};
    cplx<T> u1[4] = {v[1], v[3], v[5], v[7]};
    dft4(u0[0], u0[1], u0[2], u0[3], inv);
    dft4(u1[0], u1[1], u1[2], u1[3], inv);
#pragma unroll
    for (int q1 = 1; q1 < 4; ++q1) u1[q1] = cmul(u1[q1], w16<T>(2 * q1, inv));
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
        cplx<T> a = u0[q1], b = u1[q1];
        dft2(a, b);
        v[q1] = a;
        v[q1 + 4] = b;
    }
}

// 16-point DFT as 4 x 4: inner DFTs over r1 (stride 4), twiddle w16^(r2 q1),
// outer DFTs over r2: X[q1 + 4 q2].
template <typename T>
__device__ __forceinline__ void dft16(cplx<T> *v, bool inv) {
    cplx<T> u[4][4];
#pragma unroll
    for (int r2 = 0; r2 < 4; ++r2) {
        cplx<T> a0 = v[r2], a1 = v[r2 + 4], a2 = v[r2 + 8], a3 = v[r2 + 12];
        dft4(a0, a1, a2, a3, inv);
        u[r2][0] = a0; u[r2][1] = a1; u[r2][2] = a2; u[r2][3] = a3;
    }
#pragma unroll
    for (int r2 = 1; r2 < 4; ++r2)
#pragma unroll
        for (int q1 = 1; q1 < 4; ++q1) u[r2][q1] = cmul(u[r2][q1], w16<T>(r2 * q1, inv));
#pragma unroll
    for (int q1 = 0; q1 < 4; ++q1) {
        cplx<T> a0 = u[0][q1], a1 = u[1][q1], a2 = u[2][q1], a3 = u[3][q1];
        dft4(a0, a1, a2, a3, inv);
        v[q1] = a0; v[q1 + 4] = a1; v[q1 + 8] = a2; v[q1 + 12] = a3;
    }
}

// Odd prime R: direct DFT with roots from the line's twiddle table.
template <typename T, int R>
__device__ __forceinline__ void dft_prime(cplx<T> *v, const cplx<T> *tw, int n, bool inv) {
    cplx<T> out[R];
    const int step = n / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        cplx<T> acc = v[0];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            cplx<T> w = tw[((r * q) % R) * step];
            if (inv) w.y = -w.y;
            acc = cadd(acc, cmul(v[r], w));
        }
        out[q] = acc;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
}

template <typename T, int R>
__device__ __forceinline__ void dft_any(cplx<T> *v, const cplx<T> *tw, int n, bool inv) {
    if constexpr (R == 2) {
        dft2(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], inv);
    } else if constexpr (R == 8) {
        dft8(v, inv);
    } else if constexpr (R == 16) {
        dft16(v, inv);
    } else {
        dft_prime<T, R>(v, tw, n, inv);
    }
}

// v[r] *= w^(r st), r = 1..R-1, where w^t = tw[t] (r st < n).  For
// power-of-two R only the roots w^(2^b st) are read from the table; each
// other power is one product of an earlier power and a table root, so no
// twiddle carries more than three extra roundings.
template <typename T, int R>
__device__ __forceinline__ void twiddle(cplx<T> *v, const cplx<T> *tw, int st, bool inv) {
    if constexpr (R == 2 || R == 4 || R == 8 || R == 16) {
        constexpr int LB = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
        // conjugation for the inverse as a product with -1 (exact), so the
        // table loads stay out of any branch
        const T sg = inv ? T(-1) : T(1);
        cplx<T> pw[LB];
#pragma unroll
        for (int b = 0; b < LB; ++b) {
            pw[b] = tw[st << b];
            pw[b].y *= sg;
        }
        cplx<T> w[R];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const int low = r & -r;
            const int lb = low == 1 ? 0 : low == 2 ? 1 : low == 4 ? 2 : 3;
            w[r] = (r == low) ? pw[lb] : cmul(w[r - low], pw[lb]);
            v[r] = cmul(v[r], w[r]);
        }
    } else {
        const T sg = inv ? T(-1) : T(1);
#pragma unroll
        for (int r = 1; r < R; ++r) {
            cplx<T> w = tw[r * st];
            w.y *= sg;
            v[r] = cmul(v[r], w);
        }
    }
}

// Butterflies one lane may hold in a radix-R pass (16 complex values).
template <int R>
constexpr int lane_butterflies() {
    return R >= 16 ? 1 : 16 / R;
}

// One Stockham pass over a line in LDS.  Every lane of the workgroup must
// call it (it contains barriers); `t` is the lane's index within its line
// and `nt` the number of lanes per line.  Lane t handles butterflies t,
// t + nt, ... (at most lane_butterflies<R>()); the planner guarantees that
// covers all n/R of them.
template <typename T, int R>
__device__ __forceinline__ void stockham_pass(cplx<T> *line, const LinePlan &lp, int ns, int t, const cplx<T> *tw, bool inv) {
    constexpr int MB = lane_butterflies<R>();
    const int n = lp.n, nt = lp.threads;
    const int nr = n / R;
    cplx<T> v[MB][R];
    int base[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = t + b * nt;
        base[b] = -1;
        if (j < nr) {
            const int k = j % ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = line[lslot(lp, j + r * nr)];
            if (k != 0) twiddle<T, R>(v[b], tw, k * (n / (ns * R)), inv);
            dft_any<T, R>(v[b], tw, n, inv);
            base[b] = (j / ns) * ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        if (base[b] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) line[lslot(lp, base[b] + r * ns)] = v[b][r];
        }
    }
    __syncthreads();
}

// One Stockham pass of a prime radix R above 13, known only at run time.
// Lane t owns the output slots o = t + m nt (m < 16; the planner gives such
// a line at least n / 16 lanes).  Slot o = (j / ns) ns R + q ns + k of
// butterfly j (k = j mod ns) is sum_r x[j + r n/R] w^(r e) with
// e = k n/(ns R) + q n/R, the twiddle and the R-point DFT in one direct sum
// over table roots (O(n R) work per pass).
template <typename T>
__device__ __forceinline__ void stockham_pass_prime(cplx<T> *line, const LinePlan &lp, int ns, int t, const cplx<T> *tw,
                                                 bool inv, int R) {
    const int n = lp.n, nt = lp.threads, nr = n / R, st = n / (ns * R);
    cplx<T> out[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int o = t + m * nt;
        if (o < n) {
            const int k = o % ns, q = (o / ns) % R, jb = o / (ns * R);
            const int j = jb * ns + k, e = k * st + q * nr;
            cplx<T> acc = line[lslot(lp, j)];
            int idx = e;
            for (int r = 1; r < R; ++r) {
                cplx<T> w = tw[idx];
                if (inv) w.y = -w.y;
                acc = cadd(acc, cmul(line[lslot(lp, j + r * nr)], w));
                idx += e;
                if (idx >= n) idx -= n;
            }
            out[m] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int o = t + m * nt;
        if (o < n) line[lslot(lp, o)] = out[m];
    }
    __syncthreads();
}

// Full transform of one line.  Uniform across the workgroup.
template <typename T>
__device__ __forceinline__ void line_fft(cplx<T> *line, const LinePlan &lp, int t, const cplx<T> *tw, bool inv) {
    int ns = 1;
    for (int p = 0; p < lp.npass; ++p) {
        const int R = lp.radix[p];
        switch (R) {
            case 2: stockham_pass<T, 2>(line, lp, ns, t, tw, inv); break;
            case 3: stockham_pass<T, 3>(line, lp, ns, t, tw, inv); break;
            case 4: stockham_pass<T, 4>(line, lp, ns, t, tw, inv); break;
            case 5: stockham_pass<T, 5>(line, lp, ns, t, tw, inv); break;
            case 7: stockham_pass<T, 7>(line, lp, ns, t, tw, inv); break;
            case 8: stockham_pass<T, 8>(line, lp, ns, t, tw, inv); break;
            case 11: stockham_pass<T, 11>(line, lp, ns, t, tw, inv); break;
            case 13: stockham_pass<T, 13>(line, lp, ns, t, tw, inv); break;
            case 16: stockham_pass<T, 16>(line, lp, ns, t, tw, inv); break;
            default: stockham_pass_prime<T>(line, lp, ns, t, tw, inv, R); break;
        }
        ns *= R;
    }
}

// ---------------------------------------------------------------------------
// Compile-time schedules for power-of-two lines, 16 <= n <= 4096
// ---------------------------------------------------------------------------
// The same radices plan_line() picks for these lengths with default knobs
// (radix-16 passes, then an 8/4 tail; 2^(4a+1) uses 16^(a-1) 8 4), so a line
// transformed by either path carries the same roundings.  With n, the radix
// and the span known at compile time every index is a shift or mask, the
// swizzle is fixed and each kernel only holds one pass's registers.
constexpr int clog2(int n) { return n <= 1 ? 0 : 1 + clog2(n / 2); }

constexpr int p2_npass(int n) {
    return (clog2(n) % 4 == 0) ? clog2(n) / 4 : (clog2(n) % 4 == 1 && clog2(n) < 4) ? 1 : clog2(n) / 4 + 1;
}

constexpr int p2_radix(int n, int p) {
    const int a = clog2(n) / 4, r = clog2(n) % 4;
    if (r == 0) return 16;
    if (r == 1) return a == 0 ? 2 : p < a - 1 ? 16 : p == a - 1 ? 8 : 4;
    if (r == 2) return p < a ? 16 : 4;
    return p < a ? 16 : 8;
}

// Host check: `lp` is the default plan of a length with a fixed schedule.
inline bool p2_matches(const LinePlan &lp, int n) {
    if (lp.n != n || lp.threads != n / 16 || lp.npass != p2_npass(n) || !lp.swz) return false;
    for (int p = 0; p < lp.npass; ++p)
        if (lp.radix[p] != p2_radix(n, p)) return false;
    return true;
}

__host__ __device__ constexpr int p2slot(int i) { return i ^ ((i >> 4) & 15); }

// Ordering point between the reads and the writes of a pass.  With N / 16
// <= 64 lanes per line, every line lies inside one wavefront, whose LDS
// operations execute in program order, so a wavefront-scope fence (which
// also keeps the compiler from moving LDS accesses across it) replaces the
// workgroup barrier and the other lines' waves run on independently.
// Callers that read other lines after the transform add their own barrier.
template <int N>
__device__ __forceinline__ void pass_sync() {
    if constexpr (N / 16 <= 64) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// `lx` XORs every slot of the line (a per-line bank rotation for workgroups
// that read several lines at one element index; 0 or a multiple of 4 < 32).
template <typename T, int N, int R, int NS>
__device__ __forceinline__ void stockham_pass_p2(cplx<T> *line, int t, const cplx<T> *tw, bool inv, int lx) {
    constexpr int MB = lane_butterflies<R>();
    constexpr int NT = N / 16, NR = N / R, ST = N / (NS * R);
    static_assert(MB * NT == NR, "every butterfly has a lane");
    cplx<T> v[MB][R];
    int base[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int j = t + b * NT;
        const int k = j % NS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = line[p2slot(j + r * NR) ^ lx];
        // k = 0 multiplies by the table's exact w^0 = 1 (no branch: lanes of
        // one wave differ in k, and a branch would split the table loads)
        if (NS > 1) twiddle<T, R>(v[b], tw, k * ST, inv);
        dft_any<T, R>(v[b], tw, N, inv);
        base[b] = (j / NS) * NS * R + k;
    }
    pass_sync<N>();
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) line[p2slot(base[b] + r * NS) ^ lx] = v[b][r];
    pass_sync<N>();
}

template <typename T, int N, int P, int NS>
__device__ __forceinline__ void p2_passes(cplx<T> *line, int t, const cplx<T> *tw, bool inv, int lx) {
    if constexpr (P < p2_npass(N)) {
        constexpr int R = p2_radix(N, P);
        stockham_pass_p2<T, N, R, NS>(line, t, tw, inv, lx);
        p2_passes<T, N, P + 1, NS * R>(line, t, tw, inv, lx);
    }
}

// The first pass of a fixed schedule whose first radix is 16, fed from
// registers: lane t holds elements t + r N/16 (r < 16), exactly the inputs
// of its butterfly, so the line need not be written to LDS and read back
// before it.  Continue with p2_passes<T, N, 1, 16>.
template <int N>
constexpr bool p2_first_from_regs() {
    return N >= 256 && p2_radix(N, 0) == 16;
}

template <typename T, int N>
__device__ __forceinline__ void p2_first_pass_regs(cplx<T> *line, int t, cplx<T> v[16], const cplx<T> *tw,
                                                   bool inv, int lx) {
    static_assert(p2_first_from_regs<N>(), "first radix 16");
    dft_any<T, 16>(v, tw, N, inv);
#pragma unroll
    for (int r = 0; r < 16; ++r) line[p2slot(t * 16 + r) ^ lx] = v[r];
    pass_sync<N>();
}

// Line access policy: N > 0 a fixed power-of-two schedule (N / 16 lanes per
// line), N == 0 the runtime plan.
template <int N>
struct LineOps {
    __device__ __forceinline__ static int n(const LinePlan &lp) { return N ? N : lp.n; }
    __device__ __forceinline__ static int threads(const LinePlan &lp) { return N ? N / 16 : lp.threads; }
    __device__ __forceinline__ static int slot(const LinePlan &lp, int i) { return N ? p2slot(i) : lslot(lp, i); }
    // Bank rotation of line j of a multi-line workgroup: fixed schedules of
    // n >= 32 XOR the slots with (j mod 8) * 4, so eight lines read at one
    // element index hit eight different bank groups without padding.
    __device__ __forceinline__ static int line_xor(int j) { return N >= 32 ? ((j & 7) << 2) : 0; }
    __device__ __forceinline__ static int slot(const LinePlan &lp, int i, int lx) {
        return N ? (p2slot(i) ^ lx) : lslot(lp, i);
    }
    template <typename T>
    __device__ __forceinline__ static void fft(cplx<T> *line, const LinePlan &lp, int t, const cplx<T> *tw, bool inv,
                                               int lx = 0) {
        if constexpr (N > 0) {
            (void)lp;
            p2_passes<T, N, 0, 1>(line, t, tw, inv, lx);
        } else {
            (void)lx;
            line_fft(line, lp, t, tw, inv);
        }
    }
};

// Register-resident DFT of a short line (all n values in one lane), used for
// the z transforms of thin films.  Direct O(n^2) sum with table roots for n
// not a power of two <= 16.
template <typename T, int N>
__device__ __forceinline__ void reg_dft(cplx<T> *v, const cplx<T> *tw, bool inv) {
    if constexpr (N == 1) {
        (void)v; (void)tw; (void)inv;
    } else if constexpr (N == 2 || N == 4 || N == 8 || N == 16) {
        dft_any<T, N>(v, tw, N, inv);
    } else {
        cplx<T> out[N];
#pragma unroll
        for (int q = 0; q < N; ++q) {
            cplx<T> acc = v[0];
#pragma unroll
            for (int r = 1; r < N; ++r) {
                cplx<T> w = tw[(r * q) % N];
                if (inv) w.y = -w.y;
                acc = cadd(acc, cmul(v[r], w));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < N; ++q) v[q] = out[q];
    }
}

}  // namespace fm
