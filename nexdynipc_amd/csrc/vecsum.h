// Sums of dynamic-size Eigen vector expressions, in Eigen's order.
//
// The reference forms a few sums with Eigen reductions over dynamic-size vectors:
// VectorXd::dot for the line-search slopes (LineSearch.cpp:17,24) and lambda.dot(C),
// C.squaredNorm() in Joint::value (Joint.cpp:12-14, dim() = 6 for fixed and hinge
// joints). Eigen (Redux.h, linear vectorized traversal without unrolling) sums such an
// expression over packets of two doubles, the packet width of x86-64's baseline SSE2
// (the reference's build selects no wider instruction set). With n terms t_i:
//   n == 0: 0;  n == 1: t_0;  n == 2 or 3: t_0 + t_1, then + t_2;
//   n >= 4: packet accumulators P = (t_0, t_1) and Q = (t_2, t_3) add the packets at
//     4, 8, ... (P) and 6, 10, ... (Q) below n4 = n - n % 4; then P = P + Q; if
//     n2 = n - n % 2 exceeds n4, P = P + (t_n4, t_n4+1); the result is P.x + P.y, plus
//     t_{n-1} when n is odd.
// Lane r of the two packets (r = i mod 4, i < n4) is therefore a left-to-right chain
// that starts at t_r, and the four chains c_r combine as
//   ((c0 + c2) [+ t_n4]) + ((c1 + c3) [+ t_n4+1]) [+ t_{n-1}].
// Each term is a single product (pmul of the packets) or a square, as in scalar code.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ND_VS __host__ __device__ __forceinline__
#else
#define ND_VS inline
#endif

namespace nd {

// Combination of the four chains for n >= 4; t_n4, t_n4p1 are read when n - n % 2 > n4
// and t_last when n is odd.
ND_VS double eigen_sum_finish(double c0, double c1, double c2, double c3, int n, double t_n4, double t_n4p1,
                              double t_last) {
    const int n4 = n - n % 4, n2 = n - n % 2;
    double px = c0 + c2;
    double py = c1 + c3;
    if (n2 > n4) {
        px = px + t_n4;
        py = py + t_n4p1;
    }
    double r = px + py;
    if (n % 2) r = r + t_last;
    return r;
}

// Eigen's sum of t(0), ..., t(n - 1) for a dynamic-size vector expression.
template <typename F>
ND_VS double eigen_vec_sum(int n, const F& t) {
    if (n <= 0) return 0.0;
    if (n == 1) return t(0);
    if (n < 4) {
        double r = t(0) + t(1);
        if (n == 3) r = r + t(2);
        return r;
    }
    const int n4 = n - n % 4, n2 = n - n % 2;
    double c0 = t(0), c1 = t(1), c2 = t(2), c3 = t(3);
    for (int i = 4; i < n4; i += 4) {
        c0 = c0 + t(i);
        c1 = c1 + t(i + 1);
        c2 = c2 + t(i + 2);
        c3 = c3 + t(i + 3);
    }
    const double a = n2 > n4 ? t(n4) : 0.0;
    const double b = n2 > n4 ? t(n4 + 1) : 0.0;
    const double l = (n % 2) ? t(n - 1) : 0.0;
    return eigen_sum_finish(c0, c1, c2, c3, n, a, b, l);
}

// a.dot(b) and v.squaredNorm() for dynamic-size vectors of six entries (Joint::value).
ND_VS double eigen_dot6(const double* a, const double* b) {
    return eigen_vec_sum(6, [&](int i) { return a[i] * b[i]; });
}
ND_VS double eigen_sqnorm6(const double* v) {
    return eigen_vec_sum(6, [&](int i) { return v[i] * v[i]; });
}

}  // namespace nd
