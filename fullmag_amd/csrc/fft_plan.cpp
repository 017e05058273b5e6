// fft_plan.cpp -- host planning for the LDS line FFTs (see fft.hpp).
#include "fft.hpp"

#include <algorithm>
#include <cmath>

namespace fm {

namespace {

void pow2_radices(int k, bool wide, std::vector<int> &out, int max_radix) {
    // k = log2 of the power-of-two factor.  Default: as many radix-16 passes
    // as possible with a radix-8/4 tail (few LDS round trips, >= 4 lanes of
    // work per pass).  `wide` restricts to radices 8 and 16 so that long
    // lines fit in 1024 lanes.
    if (k == 0) return;
    if (max_radix < 16) {
        // radix <= max_radix (8 or 4), largest first, remainder last
        const int lr = max_radix >= 8 ? 3 : 2;
        while (k >= lr) { out.push_back(1 << lr); k -= lr; }
        if (k > 0) out.push_back(1 << k);
        return;
    }
    if (wide) {
        int y = k / 4;
        while (y >= 0 && (k - 4 * y) % 3 != 0) --y;
        if (y < 0) { out.push_back(-1); return; }
        for (int i = 0; i < y; ++i) out.push_back(16);
        for (int i = 0; i < (k - 4 * y) / 3; ++i) out.push_back(8);
        return;
    }
    int a = k / 4, r = k % 4;
    switch (r) {
        case 0: for (int i = 0; i < a; ++i) out.push_back(16); break;
        case 1:
            if (a >= 1) {
                for (int i = 0; i < a - 1; ++i) out.push_back(16);
                out.push_back(8);
                out.push_back(4);
            } else {
                out.push_back(2);
            }
            break;
        case 2: for (int i = 0; i < a; ++i) out.push_back(16); out.push_back(4); break;
        case 3: for (int i = 0; i < a; ++i) out.push_back(16); out.push_back(8); break;
    }
}

}  // namespace

bool plan_line(int n, LinePlan &lp, std::string &err, int min_lanes, int max_radix) {
    lp = LinePlan{};
    if (n < 1) {
        err = "FFT line length must be >= 1";
        return false;
    }
    lp.n = n;
    if (n == 1) {
        lp.npass = 0;
        lp.threads = 1;
        return true;
    }
    int m = n, k = 0;
    while (m % 2 == 0) { m /= 2; ++k; }
    std::vector<int> odd;
    for (int p : {3, 5, 7, 11, 13})
        while (m % p == 0) { odd.push_back(p); m /= p; }
    for (int p = 17; m > 1; p += 2)
        while (m % p == 0) { odd.push_back(p); m /= p; }
    std::vector<int> rad;
    pow2_radices(k, false, rad, max_radix);
    for (int p : odd) rad.push_back(p);
    if ((int)rad.size() > kMaxPass) {
        err = "padded FFT length " + std::to_string(n) + " needs too many passes";
        return false;
    }
    // lanes: every pass's n/R butterflies must fit in lanes * (16 / R)
    int lanes = 1, most = 1;
    for (int r : rad) {
        const int mb = r >= 16 ? 1 : 16 / r;
        const int nr = n / r;
        lanes = std::max(lanes, (nr + mb - 1) / mb);
        most = std::max(most, nr);
        // a prime pass above 13 owns up to 16 output slots per lane
        if (r > 16) lanes = std::max(lanes, (n + 15) / 16), most = std::max(most, (n + 15) / 16);
    }
    lanes = std::max(lanes, std::min(min_lanes, most));
    if (lanes > 1024) {
        err = "padded FFT length " + std::to_string(n) + " needs more than 1024 lanes per line";
        return false;
    }
    lp.npass = (int)rad.size();
    for (int i = 0; i < lp.npass; ++i) lp.radix[i] = rad[i];
    lp.threads = lanes;
    lp.swz = (n % 16 == 0) ? 1 : 0;
    return true;
}

std::vector<cplx<double>> twiddle_table(int n) {
    std::vector<cplx<double>> w(n > 0 ? n : 1);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
        long double a = -two_pi * (long double)t / (long double)n;
        w[t] = {(double)cosl(a), (double)sinl(a)};
    }
    if (n <= 0) w[0] = {1.0, 0.0};
    return w;
}

}  // namespace fm
