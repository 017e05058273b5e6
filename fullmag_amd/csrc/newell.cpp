// newell.cpp -- Newell-Williams-Dunlop demag tensor on the displacement
// octant (host, FP64), for the in-library kernel build.
//
// Same tensor as crates/fullmag-fdm-demag/src/newell.rs:254-519 (Boris-style
// f/g tables, 27-point second-difference stencil with Kahan-Neumaier
// summation, point-dipole far field beyond 40 cells).  The reference places
// the octant into all eight sign octants of the padded grid on the host; here
// only the octant is produced and the device mirrors it with the parity signs
// while it transforms (demag.hip, NewellSource), so the host never holds the
// padded arrays.  A nonzero z_offset gives the cross-layer kernel of layers
// that far apart in z (newell.rs:528-651: every z coordinate, of the f/g
// tables and of the far field, is offset).
#include "context.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace fm {
namespace {

// f(x, y, z): newell.rs:21-61 (log1p form, valid for signed arguments).
double nw_f(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    const double r = std::sqrt(r2);
    double v = (2.0 * x2 - y2 - z2) * r / 6.0;
    const double sxz = x2 + z2, sxy = x2 + y2;
    if (sxz > 1e-300) {
        const double t = 2.0 * y * (y + r) / sxz;
        if (t > -1.0) v += y * (z2 - x2) / 4.0 * std::log1p(t);
    }
    if (sxy > 1e-300) {
        const double t = 2.0 * z * (z + r) / sxy;
        if (t > -1.0) v += z * (y2 - x2) / 4.0 * std::log1p(t);
    }
    if (std::fabs(x) > 1e-300) v -= x * y * z * std::atan(y * z / (x * r));
    return v;
}

// g(x, y, z): newell.rs:63-123.
double nw_g(double x, double y, double z) {
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    if (r2 < 1e-300) return 0.0;
    const double r = std::sqrt(r2);
    double v = -x * y * r / 3.0;
    const double sxy = x2 + y2, syz = y2 + z2, sxz = x2 + z2;
    if (sxy > 1e-300) {
        const double t = 2.0 * z * (z + r) / sxy;
        if (t > -1.0) v += x * y * z * std::log1p(t) / 2.0;
    }
    if (syz > 1e-300) {
        const double t = 2.0 * x * (x + r) / syz;
        if (t > -1.0) v += y * (3.0 * z2 - y2) * std::log1p(t) / 12.0;
    }
    if (sxz > 1e-300) {
        const double t = 2.0 * y * (y + r) / sxz;
        if (t > -1.0) v += x * (3.0 * z2 - x2) * std::log1p(t) / 12.0;
    }
    if (std::fabs(z) > 1e-300) v -= z2 * z / 6.0 * std::atan(x * y / (z * r));
    if (std::fabs(y) > 1e-300) v -= y2 * z / 2.0 * std::atan(x * z / (y * r));
    if (std::fabs(x) > 1e-300) v -= x2 * z / 2.0 * std::atan(y * z / (x * r));
    return v;
}

constexpr size_t kFar = 40;  // newell.rs:229

struct Table {
    size_t sx, sy, sz;
    std::vector<double> v;  // value at grid offset (i-1, j-1, k-1)
    double at(size_t i, size_t j, size_t k) const { return v[(k * sy + j) * sx + i]; }
};

// Second difference over the 27 neighbours of (i, j, k) with weights
// 8 / -4 / 2 / -1 (centre / faces / edges / corners), summed with
// Kahan-Neumaier compensation: newell.rs:150-203.  Table offsets are +1.
double stencil(const Table &t, size_t i, size_t j, size_t k, double hx, double hy, double hz) {
    double sum = 0.0, comp = 0.0;
    auto add = [&](double val) {
        const double s = sum + val;
        if (std::fabs(sum) >= std::fabs(val)) comp += (sum - s) + val;
        else comp += (val - s) + sum;
        sum = s;
    };
    const size_t a = i + 1, b = j + 1, c = k + 1;
    add(8.0 * t.at(a, b, c));
    add(-4.0 * t.at(a + 1, b, c)); add(-4.0 * t.at(a - 1, b, c));
    add(-4.0 * t.at(a, b + 1, c)); add(-4.0 * t.at(a, b - 1, c));
    add(-4.0 * t.at(a, b, c + 1)); add(-4.0 * t.at(a, b, c - 1));
    add(2.0 * t.at(a - 1, b - 1, c)); add(2.0 * t.at(a - 1, b + 1, c));
    add(2.0 * t.at(a + 1, b - 1, c)); add(2.0 * t.at(a + 1, b + 1, c));
    add(2.0 * t.at(a - 1, b, c - 1)); add(2.0 * t.at(a - 1, b, c + 1));
    add(2.0 * t.at(a + 1, b, c - 1)); add(2.0 * t.at(a + 1, b, c + 1));
    add(2.0 * t.at(a, b - 1, c - 1)); add(2.0 * t.at(a, b - 1, c + 1));
    add(2.0 * t.at(a, b + 1, c - 1)); add(2.0 * t.at(a, b + 1, c + 1));
    add(-t.at(a - 1, b - 1, c - 1)); add(-t.at(a - 1, b - 1, c + 1));
    add(-t.at(a - 1, b + 1, c - 1)); add(-t.at(a + 1, b - 1, c - 1));
    add(-t.at(a - 1, b + 1, c + 1)); add(-t.at(a + 1, b - 1, c + 1));
    add(-t.at(a + 1, b + 1, c - 1)); add(-t.at(a + 1, b + 1, c + 1));
    return (sum + comp) / (4.0 * kPi * hx * hy * hz);
}

// Point-dipole limit: newell.rs:212-226.
double dip_diag(double x, double y, double z, double vol) {
    const double r2 = x * x + y * y + z * z, r = std::sqrt(r2), r3 = r2 * r;
    return (1.0 / r3 - 3.0 * x * x / (r3 * r2)) / (4.0 * kPi) * vol;
}
double dip_off(double x, double y, double z, double vol) {
    const double r2 = x * x + y * y + z * z, r = std::sqrt(r2), r5 = r2 * r2 * r;
    return -3.0 * x * y / (4.0 * kPi * r5) * vol;
}

template <class F>
void parallel_for(size_t n, F f) {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(hw, std::max<size_t>(1, n));
    if (nt <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (size_t i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

void newell_octant(size_t nx, size_t ny, size_t nz, double dx, double dy, double dz, std::vector<double> &out,
                   double z_offset) {
    const size_t sx = std::min(nx, kFar) + 2, sy = std::min(ny, kFar) + 2, sz = std::min(nz, kFar) + 2;
    // six base-function tables on offsets -1 .. min(n, 40) along each axis
    Table t[6];
    for (auto &tb : t) {
        tb.sx = sx; tb.sy = sy; tb.sz = sz;
        tb.v.assign(sx * sy * sz, 0.0);
    }
    parallel_for(sz, [&](size_t k) {
        const double z = ((double)k - 1.0) * dz + z_offset;
        for (size_t j = 0; j < sy; ++j) {
            const double y = ((double)j - 1.0) * dy;
            for (size_t i = 0; i < sx; ++i) {
                const double x = ((double)i - 1.0) * dx;
                const size_t id = (k * sy + j) * sx + i;
                t[0].v[id] = nw_f(x, y, z);
                t[1].v[id] = nw_f(y, x, z);
                t[2].v[id] = nw_f(z, y, x);
                t[3].v[id] = nw_g(x, y, z);
                t[4].v[id] = nw_g(x, z, y);
                t[5].v[id] = nw_g(y, z, x);
            }
        }
    });
    const size_t per = nx * ny * nz;
    out.assign(6 * per, 0.0);
    const double vol = dx * dy * dz;
    parallel_for(nz, [&](size_t k) {
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t o = (k * ny + j) * nx + i;
                const bool far = i >= kFar || j >= kFar || k >= kFar || i * i + j * j + k * k >= kFar * kFar;
                double v[6];
                if (far) {
                    const double x = (double)i * dx, y = (double)j * dy, z = (double)k * dz + z_offset;
                    v[0] = dip_diag(x, y, z, vol);
                    v[1] = dip_diag(y, x, z, vol);
                    v[2] = dip_diag(z, y, x, vol);
                    v[3] = dip_off(x, y, z, vol);
                    v[4] = dip_off(x, z, y, vol);
                    v[5] = dip_off(y, z, x, vol);
                } else {
                    // argument order of the cell sizes follows newell.rs:386-392
                    v[0] = stencil(t[0], i, j, k, dx, dy, dz);
                    v[1] = stencil(t[1], i, j, k, dy, dx, dz);
                    v[2] = stencil(t[2], i, j, k, dz, dy, dx);
                    v[3] = stencil(t[3], i, j, k, dx, dy, dz);
                    v[4] = stencil(t[4], i, j, k, dx, dz, dy);
                    v[5] = stencil(t[5], i, j, k, dy, dz, dx);
                }
                for (int c = 0; c < 6; ++c) out[c * per + o] = v[c];
            }
    });
}

}  // namespace fm
