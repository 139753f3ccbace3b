// Elementary functions used on the contact-stage path: log, sin, cos, atan2, acos.
//
// The reference calls the platform's libm (std::log in the barrier potential, std::sin /
// std::cos in Eigen's AngleAxis, std::atan2 in the quaternion -> AngleAxis conversion,
// std::acos / std::sin in Eigen's slerp). Two libms may differ in the last bit, and a
// one-bit difference in the barrier or in a rotation is enough to flip the line-search
// and ALM stopping tests of a Newton step, after which two runs drift apart. These
// implementations use only IEEE-754 +, -, *, /, sqrt, fma, floor and fmod, each of
// which is correctly rounded (or exact) on every conforming machine; compiled without
// FP contraction they therefore give the same bits on the GPU and on any host.
// Accuracy is within about 1 ulp of the true value on the domains used here; the host
// tests compare every function with the C library on random and special arguments.
//
// Methods (all from the mathematics; coefficients are Taylor coefficients, rounded to
// nearest double):
//  log:  x = 2^k m, m in (sqrt(2)/2, sqrt(2)], f = m - 1, s = f / (2 + f),
//        log(1 + f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))) (the reconstruction
//        of Sun's fdlibm e_log.c), R = sum_{i>=1} 2 z^i / (2i + 1) to i = 12,
//        plus k ln2 with ln2 split so that k * ln2_hi is exact.
//  sin, cos: Cody-Waite reduction by pi/2 split in three 33-bit parts (exact products
//        for |x| < 2^19), Taylor kernels on |r| <= pi/4 (sin to r^19, cos to r^20 with
//        the compensated 1 - r^2/2 of fdlibm's __kernel_cos).
//  atan: for t in [0, 1]: Taylor series to t^31 below tan(pi/12), otherwise
//        atan(t) = pi/6 + atan((sqrt(3) t - 1) / (sqrt(3) + t)) with sqrt(3) in two parts.
//  atan2: C99 Annex F special cases around atan of min/max(|x|, |y|).
//  acos: pi/2 - atan(x / sqrt((1 - x)(1 + x))) for |x| <= 1/2,
//        2 atan(sqrt((1 - x) / (1 + x))) above, pi - acos(-x) below.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ND_EF __host__ __device__ __forceinline__
#else
#define ND_EF inline
#endif

namespace nd {
namespace ef {

ND_EF double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
ND_EF uint64_t to_bits(double x) { return __builtin_bit_cast(uint64_t, x); }
ND_EF double qnan() { return from_bits(0x7ff8000000000000ull); }
ND_EF double inf() { return from_bits(0x7ff0000000000000ull); }
ND_EF bool is_nan(double x) { return x != x; }
ND_EF double copy_sign(double a, double s) {
    return from_bits((to_bits(a) & 0x7fffffffffffffffull) | (to_bits(s) & 0x8000000000000000ull));
}

// ln 2 = LN2_HI + LN2_LO, LN2_HI with 21 significant bits
constexpr double LN2_HI = 0x1.62e42p-1;
constexpr double LN2_LO = 0x1.fdf473de6af28p-22;
constexpr double SQRT2 = 0x1.6a09e667f3bcdp+0;
// pi/2 = PIO2_1 + PIO2_2 + PIO2_3, the first two with 33 significant bits
constexpr double PIO2_1 = 0x1.921fb544p+0;
constexpr double PIO2_2 = 0x1.0b4611a6p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double TWO_PI = 0x1.921fb54442d18p+2;
// two-part constants: value = HI + LO
constexpr double PI_HI = 0x1.921fb54442d18p+1, PI_LO = 0x1.1a62633145c07p-53;
constexpr double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
constexpr double PIO4_HI = 0x1.921fb54442d18p-1;
constexpr double PI6_HI = 0x1.0c152382d7366p-1, PI6_LO = -0x1.ee6913347c2a6p-55;
constexpr double S3_HI = 0x1.bb67ae8584caap+0, S3_LO = 0x1.cec95d0b5c1e3p-54;
constexpr double TAN_PI12 = 0x1.126145e9ecd56p-2;  // 2 - sqrt(3)

// ------------------------------------------------------------------ log
ND_EF double log(double x) {
    if (is_nan(x)) return x;
    if (x < 0.0) return qnan();
    if (x == 0.0) return -inf();
    if (x == inf()) return x;
    uint64_t b = to_bits(x);
    int k = 0;
    if (b < 0x0010000000000000ull) {  // subnormal: scale by 2^54
        x = x * 0x1p54;
        b = to_bits(x);
        k = -54;
    }
    k += int(b >> 52) - 1023;
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
    if (m > SQRT2) {
        m = m * 0.5;
        k += 1;
    }
    const double f = m - 1.0;  // exact
    const double s = f / (2.0 + f);
    const double z = s * s;
    const double R =
        z * (0x1.5555555555555p-1 +
        z * (0x1.999999999999ap-2 +
        z * (0x1.2492492492492p-2 +
        z * (0x1.c71c71c71c71cp-3 +
        z * (0x1.745d1745d1746p-3 +
        z * (0x1.3b13b13b13b14p-3 +
        z * (0x1.1111111111111p-3 +
        z * (0x1.e1e1e1e1e1e1ep-4 +
        z * (0x1.af286bca1af28p-4 +
        z * (0x1.8618618618618p-4 +
        z * (0x1.642c8590b2164p-4 +
        z * 0x1.47ae147ae147bp-4)))))))))));
    const double hfsq = 0.5 * f * f;
    const double dk = double(k);
    return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

// ------------------------------------------------------------------ sin / cos
ND_EF double sin_kernel(double r) {  // |r| <= ~pi/4
    const double z = r * r;
    const double p =
        -0x1.5555555555555p-3 +
        z * (0x1.1111111111111p-7 +
        z * (-0x1.a01a01a01a01ap-13 +
        z * (0x1.71de3a556c734p-19 +
        z * (-0x1.ae64567f544e4p-26 +
        z * (0x1.6124613a86d09p-33 +
        z * (-0x1.ae7f3e733b81fp-41 +
        z * (0x1.952c77030ad4ap-49 +
        z * -0x1.2f49b46814157p-57)))))));
    return r + (r * z) * p;
}
ND_EF double cos_kernel(double r) {
    const double z = r * r;
    const double c =
        0x1.5555555555555p-5 +
        z * (-0x1.6c16c16c16c17p-10 +
        z * (0x1.a01a01a01a01ap-16 +
        z * (-0x1.27e4fb7789f5cp-22 +
        z * (0x1.1eed8eff8d898p-29 +
        z * (-0x1.93974a8c07c9dp-37 +
        z * (0x1.ae7f3e733b81fp-45 +
        z * (-0x1.6827863b97d97p-53 +
        z * 0x1.e542ba4020225p-62)))))));
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + (z * z) * c);
}
// x = n pi/2 + r; returns r and n mod 4.
ND_EF double reduce_pio2(double x, int* quadrant) {
    if (fabs(x) > 0x1p19) x = fmod(x, TWO_PI);  // exact; beyond any rotation angle here
    const double fn = floor(x * TWO_OVER_PI + 0.5);
    const double r = ((x - fn * PIO2_1) - fn * PIO2_2) - fn * PIO2_3;
    *quadrant = int(fn) & 3;
    return r;
}
ND_EF void sincos(double x, double* s, double* c) {
    if (is_nan(x) || fabs(x) == inf()) {
        *s = qnan();
        *c = qnan();
        return;
    }
    if (x == 0.0) {
        *s = x;
        *c = 1.0;
        return;
    }
    int q;
    const double r = reduce_pio2(x, &q);
    const double sr = sin_kernel(r), cr = cos_kernel(r);
    switch (q) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}
ND_EF double sin(double x) {
    double s, c;
    sincos(x, &s, &c);
    return s;
}
ND_EF double cos(double x) {
    double s, c;
    sincos(x, &s, &c);
    return c;
}

// ------------------------------------------------------------------ atan / atan2 / acos
ND_EF double atan_small(double u) {  // |u| <= tan(pi/12)
    const double z = u * u;
    const double p =
        -0x1.5555555555555p-2 +
        z * (0x1.999999999999ap-3 +
        z * (-0x1.2492492492492p-3 +
        z * (0x1.c71c71c71c71cp-4 +
        z * (-0x1.745d1745d1746p-4 +
        z * (0x1.3b13b13b13b14p-4 +
        z * (-0x1.1111111111111p-4 +
        z * (0x1.e1e1e1e1e1e1ep-5 +
        z * (-0x1.af286bca1af28p-5 +
        z * (0x1.8618618618618p-5 +
        z * (-0x1.642c8590b2164p-5 +
        z * (0x1.47ae147ae147bp-5 +
        z * (-0x1.2f684bda12f68p-5 +
        z * (0x1.1a7b9611a7b96p-5 +
        z * -0x1.0842108421084p-5)))))))))))));
    return u + (u * z) * p;
}
// atan(t) for t in [0, 1]
ND_EF double atan01(double t) {
    if (t <= TAN_PI12) return atan_small(t);
    const double num = __builtin_fma(S3_HI, t, -1.0) + S3_LO * t;
    const double u = num / (S3_HI + t);
    return PI6_HI + (atan_small(u) + PI6_LO);
}
ND_EF double atan2(double y, double x) {
    if (is_nan(x) || is_nan(y)) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    const bool xneg = (to_bits(x) >> 63) != 0;
    if (ay == 0.0) return xneg ? copy_sign(PI_HI, y) : y;
    if (ax == 0.0) return copy_sign(PIO2_HI, y);
    if (ax == inf()) {
        if (ay == inf()) return copy_sign(xneg ? 3.0 * PIO4_HI : PIO4_HI, y);
        return xneg ? copy_sign(PI_HI, y) : copy_sign(0.0, y);
    }
    if (ay == inf()) return copy_sign(PIO2_HI, y);
    double a;
    if (ay > ax) {
        a = atan01(ax / ay);
        a = PIO2_HI - (a - PIO2_LO);
    } else {
        a = atan01(ay / ax);
    }
    if (xneg) a = PI_HI - (a - PI_LO);
    return copy_sign(a, y);
}
ND_EF double acos(double x) {
    if (is_nan(x)) return x;
    const double ax = fabs(x);
    if (ax > 1.0) return qnan();
    if (ax <= 0.5) {
        const double a = atan01(ax / sqrt((1.0 - ax) * (1.0 + ax)));
        return x >= 0.0 ? PIO2_HI - (a - PIO2_LO) : PIO2_HI + (a + PIO2_LO);
    }
    const double a = 2.0 * atan01(sqrt((1.0 - ax) / (1.0 + ax)));
    return x > 0.0 ? a : PI_HI - (a - PI_LO);
}

}  // namespace ef
}  // namespace nd
