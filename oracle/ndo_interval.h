// ORACLE — test infrastructure only.
//
// Interval type restating NexDynIPC's Math::Interval
// (nexdynipc/include/Math/Interval.h:25-47): boost::numeric::interval<double,
// save_state<rounded_transc_std<double>>, checking_catch_nan<double>>.
// Boost.Interval is not vendored by the reference (vcpkg baseline
// a2a478a93d582a4b395a4dc4b7052bfcb42c1f8e); its published semantics are restated:
//   - every endpoint is rounded outward with the hardware rounding mode switched per
//     operation (here: fesetround around a volatile result, compiled -frounding-math);
//   - products use the sign-case analysis of boost/numeric/interval/arith.hpp;
//   - median() rounds to nearest, width() rounds up;
//   - constructing an interval from a NaN endpoint throws (checking_catch_nan); the
//     oracle records that as an error flag instead of throwing.
#pragma once
#include <cfenv>
#include <cmath>
#include <limits>

namespace ndo {

struct RoundGuard {
    int old;
    explicit RoundGuard(int mode) : old(std::fegetround()) { std::fesetround(mode); }
    ~RoundGuard() { std::fesetround(old); }
};

inline double add_dn(double a, double b) { RoundGuard g(FE_DOWNWARD); volatile double r = a + b; return r; }
inline double add_up(double a, double b) { RoundGuard g(FE_UPWARD); volatile double r = a + b; return r; }
inline double sub_dn(double a, double b) { RoundGuard g(FE_DOWNWARD); volatile double r = a - b; return r; }
inline double sub_up(double a, double b) { RoundGuard g(FE_UPWARD); volatile double r = a - b; return r; }
inline double mul_dn(double a, double b) { RoundGuard g(FE_DOWNWARD); volatile double r = a * b; return r; }
inline double mul_up(double a, double b) { RoundGuard g(FE_UPWARD); volatile double r = a * b; return r; }

extern thread_local bool g_interval_nan;

struct I {
    double lo = 0, hi = 0;
    I() = default;
    explicit I(double v) : lo(v), hi(v) { if (std::isnan(v)) g_interval_nan = true; }
    I(double l, double h) : lo(l), hi(h) {
        if (std::isnan(l) || std::isnan(h)) g_interval_nan = true;
    }
};

inline I operator+(const I& a, const I& b) { return I(add_dn(a.lo, b.lo), add_up(a.hi, b.hi)); }
inline I operator-(const I& a, const I& b) { return I(sub_dn(a.lo, b.hi), sub_up(a.hi, b.lo)); }
inline I operator-(double a, const I& b) { return I(sub_dn(a, b.hi), sub_up(a, b.lo)); }
inline I operator*(const I& x, const I& y) {
    const double xl = x.lo, xu = x.hi, yl = y.lo, yu = y.hi;
    if (xu < 0) {
        if (yu < 0) return I(mul_dn(xu, yu), mul_up(xl, yl));
        if (yl < 0) return I(mul_dn(xl, yu), mul_up(xl, yl));
        return I(mul_dn(xl, yu), mul_up(xu, yl));
    }
    if (xl < 0) {
        if (yu < 0) return I(mul_dn(xu, yl), mul_up(xl, yl));
        if (yl < 0) {
            return I(std::min(mul_dn(xl, yu), mul_dn(xu, yl)), std::max(mul_up(xl, yl), mul_up(xu, yu)));
        }
        return I(mul_dn(xl, yu), mul_up(xu, yu));
    }
    if (yu < 0) return I(mul_dn(xu, yl), mul_up(xl, yu));
    if (yl < 0) return I(mul_dn(xu, yl), mul_up(xu, yu));
    return I(mul_dn(xl, yl), mul_up(xu, yu));
}
inline double width(const I& a) { return sub_up(a.hi, a.lo); }
inline double median(const I& a) { return (a.lo + a.hi) / 2; }
inline bool zero_in(const I& a) { return a.lo <= 0.0 && 0.0 <= a.hi; }

}  // namespace ndo
