// common.hpp -- shared types and helpers for the HIP FDM backend (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace fm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMu0 = 4.0 * kPi * 1e-7;  // crates/fullmag-engine/src/lib.rs:25

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ inline cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }

// Integer ceil-div for launch geometry.
__host__ __device__ inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Division of 32-bit unsigned integers by a fixed divisor d >= 1 with one
// multiply-high (the round-up method of Granlund & Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994): with l = ceil(log2 d)
// and m = floor(2^32 (2^l - d) / d) + 1, x / d = (mulhi(m, x) + x) >> l for
// every x < 2^32 (the sum taken in 64 bits).
struct FastDiv {
    uint32_t d = 1, m = 1, l = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t div) : d(div) {
        l = 0;
        while ((uint64_t(1) << l) < div) ++l;
        m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - div)) / div + 1);
    }
    __host__ __device__ inline uint32_t div(uint32_t x) const {
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t t = __umulhi(m, x);
#else
        const uint32_t t = (uint32_t)(((uint64_t)m * x) >> 32);
#endif
        return (uint32_t)(((uint64_t)t + x) >> l);
    }
};

// Records a HIP failure into `err` and returns false.
inline bool hip_ok(hipError_t e, const char *what, std::string &err) {
    if (e == hipSuccess) return true;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

}  // namespace fm
