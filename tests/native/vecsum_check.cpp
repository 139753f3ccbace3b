// Host check of the Eigen-order vector sums (nexdynipc_amd/csrc/vecsum.h).
//
// Eigen 3.4 (the version the reference's vcpkg baseline resolves, vcpkg.json:20) sums a
// dynamic-size vector expression with SSE2 packets of two doubles (Redux.h, linear
// vectorized traversal): two packet accumulators over the aligned body, their sum, one
// more packet if two aligned entries remain, the horizontal add of the packet
// (predux: low lane + high lane), then the scalar tail. The check runs that sequence
// with the SSE2 instructions themselves (_mm_add_pd, _mm_mul_pd, and predux as
// _mm_add_sd(p, unpackhi(p, p))) and compares bit for bit with eigen_vec_sum /
// eigen_dot6 / eigen_sqnorm6, for every length up to 64 and random lengths to 5000, on
// cancellation-heavy terms. Known answers: [1e16, 1, -1e16, 1] sums to 2 (left to right
// gives 1), and integer terms below 2^53 sum exactly. Prints "OK <cases>".
#include <emmintrin.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../nexdynipc_amd/csrc/vecsum.h"

static double sse2_redux_dot(const double* a, const double* b, int n) {
    if (n == 0) return 0.0;
    const int ps = 2;
    const int size2 = (n / (2 * ps)) * (2 * ps), size1 = (n / ps) * ps;
    if (size1 == 0) return a[0] * b[0];
    __m128d p0 = _mm_mul_pd(_mm_loadu_pd(a), _mm_loadu_pd(b));
    if (size1 > ps) {
        __m128d p1 = _mm_mul_pd(_mm_loadu_pd(a + 2), _mm_loadu_pd(b + 2));
        for (int i = 2 * ps; i < size2; i += 2 * ps) {
            p0 = _mm_add_pd(p0, _mm_mul_pd(_mm_loadu_pd(a + i), _mm_loadu_pd(b + i)));
            p1 = _mm_add_pd(p1, _mm_mul_pd(_mm_loadu_pd(a + i + ps), _mm_loadu_pd(b + i + ps)));
        }
        p0 = _mm_add_pd(p0, p1);
        if (size1 > size2) p0 = _mm_add_pd(p0, _mm_mul_pd(_mm_loadu_pd(a + size2), _mm_loadu_pd(b + size2)));
    }
    double r = _mm_cvtsd_f64(_mm_add_sd(p0, _mm_unpackhi_pd(p0, p0)));
    for (int i = size1; i < n; ++i) r += a[i] * b[i];
    return r;
}

static bool same(double x, double y) { return std::memcmp(&x, &y, sizeof(double)) == 0; }

int main() {
    std::mt19937_64 rng(77);
    std::normal_distribution<double> N(0.0, 1.0);
    long cases = 0;
    {
        const double t[4] = {1e16, 1.0, -1e16, 1.0};
        const double r = nd::eigen_vec_sum(4, [&](int i) { return t[i]; });
        double ltr = 0.0;
        for (double v : t) ltr += v;
        if (r != 2.0 || ltr != 1.0) {
            std::printf("FAIL known answer: %g (left to right %g)\n", r, ltr);
            return 1;
        }
        ++cases;
    }
    for (int trial = 0; trial < 20000; ++trial) {
        const int n = trial < 6500 ? trial % 65 : 1 + int(rng() % 5000);
        std::vector<double> a(n), b(n);
        for (int i = 0; i < n; ++i) {
            a[i] = N(rng) * std::exp(12.0 * N(rng));
            b[i] = (i % 3 == 0) ? -a[i] * (1.0 + 1e-13 * N(rng)) : N(rng);
            if (trial % 5 == 0) b[i] = 1.0;  // plain sums: exact cancellations across lanes
        }
        const double ref = sse2_redux_dot(a.data(), b.data(), n);
        const double got = nd::eigen_vec_sum(n, [&](int i) { return a[i] * b[i]; });
        if (!same(ref, got)) {
            std::printf("FAIL n %d: sse2 %.17g vecsum %.17g\n", n, ref, got);
            return 1;
        }
        if (n == 6) {
            const double sq = sse2_redux_dot(a.data(), a.data(), 6);
            if (!same(nd::eigen_dot6(a.data(), b.data()), ref) || !same(nd::eigen_sqnorm6(a.data()), sq)) {
                std::printf("FAIL dot6 / sqnorm6\n");
                return 1;
            }
        }
        ++cases;
    }
    for (int trial = 0; trial < 2000; ++trial) {  // integers below 2^53 in total: exact
        const int n = 1 + int(rng() % 3000);
        std::vector<double> t(n);
        long long exact = 0;
        for (int i = 0; i < n; ++i) {
            const long long v = (long long)(rng() % 2000001) - 1000000;
            t[i] = double(v);
            exact += v;
        }
        if (nd::eigen_vec_sum(n, [&](int i) { return t[i]; }) != double(exact)) {
            std::printf("FAIL exact integer sum n %d\n", n);
            return 1;
        }
        ++cases;
    }
    std::printf("OK %ld\n", cases);
    return 0;
}
