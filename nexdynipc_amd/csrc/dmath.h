// Fixed-size FP64 geometry for the contact-stage kernels.
//
// Every formula keeps the operation order of the reference's Eigen expressions so that
// device results are bit-identical to an unvectorised host evaluation when both are
// compiled without FP contraction (-ffp-contract=off; sqrt and division are IEEE
// correctly rounded on gfx950):
//   dot = (a0*b0 + a1*b1) + a2*b2; cross = (ay bz - az by, az bx - ax bz, ax by - ay bx);
//   q*v = v + w*uv + qv x uv with uv = 2 (qv x v); Hamilton product;
//   AngleAxis -> quaternion (cos, sin of the half angle); normalized = each / sqrt(|.|^2).
// Elementary functions come from elemfn.h, which rounds identically on every machine.
#pragma once
#include <math.h>
#include <stdint.h>

#include "elemfn.h"
#include "vecsum.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ND_HD __host__ __device__ __forceinline__
#else
#define ND_HD inline
#endif

namespace nd {

struct V3 {
    double x, y, z;
};
ND_HD V3 v3(double a, double b, double c) { V3 r; r.x = a; r.y = b; r.z = c; return r; }
ND_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
ND_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
ND_HD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
ND_HD V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
ND_HD V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
ND_HD V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
ND_HD double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
ND_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ND_HD double sqnorm(V3 a) { return dot(a, a); }
ND_HD double norm(V3 a) { return sqrt(sqnorm(a)); }
ND_HD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
ND_HD V3 normalized(V3 a) {
    const double n2 = sqnorm(a);
    return n2 > 0.0 ? a / sqrt(n2) : a;
}
ND_HD double dmin(double a, double b) { return (b < a) ? b : a; }
ND_HD double dmax(double a, double b) { return (a < b) ? b : a; }
ND_HD V3 vmin(V3 a, V3 b) { return v3(dmin(a.x, b.x), dmin(a.y, b.y), dmin(a.z, b.z)); }
ND_HD V3 vmax(V3 a, V3 b) { return v3(dmax(a.x, b.x), dmax(a.y, b.y), dmax(a.z, b.z)); }

struct Q {
    double w, x, y, z;
};
ND_HD Q quat(double w, double x, double y, double z) { Q q; q.w = w; q.x = x; q.y = y; q.z = z; return q; }
ND_HD V3 qvec(Q q) { return v3(q.x, q.y, q.z); }
ND_HD V3 rotate(Q q, V3 v) {
    V3 uv = cross(qvec(q), v);
    uv = uv + uv;
    return (v + q.w * uv) + cross(qvec(q), uv);
}
ND_HD Q qmul(Q a, Q b) {
    return quat(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x);
}
ND_HD Q qconj(Q q) { return quat(q.w, -q.x, -q.y, -q.z); }
ND_HD Q qnormalized(Q q) {
    const double n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n2 > 0.0) {
        const double n = sqrt(n2);
        return quat(q.w / n, q.x / n, q.y / n, q.z / n);
    }
    return q;
}
ND_HD Q from_angle_axis(double angle, V3 axis) {
    const double ha = 0.5 * angle;
    double s, c;
    ef::sincos(ha, &s, &c);
    return quat(c, s * axis.x, s * axis.y, s * axis.z);
}
// BarrierPotential (Barrier.cpp:11-46): b = -k (d - dh)^2 ln(d / dh) on (0, dh), 0 at
// and above dh, +inf at and below 0; its first and second derivatives in d.
ND_HD double bar_v(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    if (d <= 0.0) return 1.0 / 0.0;
    const double r = d / dh;
    return -k * (d - dh) * (d - dh) * ef::log(r);
}
ND_HD double bar_g(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    if (d <= 0.0) return -1.0 / 0.0;
    const double diff = d - dh;
    const double lr = ef::log(d / dh);
    return -k * (2 * diff * lr + diff * diff / d);
}
ND_HD double bar_h(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    if (d <= 0.0) return 1.0 / 0.0;
    const double diff = d - dh;
    const double lr = ef::log(d / dh);
    return -k * (2 * lr + 4 * diff / d - (diff * diff) / (d * d));
}
// Pose of a body at Newton variable theta: normalize(AngleAxis(|t|, t/|t|) * q) for
// |t| > 1e-10 (IPCSolver.cpp:616-626, FrictionForm.cpp:206-223).
ND_HD Q apply_theta(Q q, V3 th) {
    const double ang = norm(th);
    if (ang > 1e-10) return qnormalized(qmul(from_angle_axis(ang, th / ang), q));
    return q;
}
// JointPrimitives::updateRotation (JointPrimitives.h:11-19): "< 1e-10 -> q".
ND_HD Q update_rotation(Q q, V3 th) {
    const double ang = norm(th);
    if (ang < 1e-10) return q;
    return qnormalized(qmul(from_angle_axis(ang, th / ang), q));
}
// Rotation vector of a unit quaternion, angle = 2 atan2(|v|, |w|) (Eigen AngleAxis).
ND_HD V3 rotation_vector(Q q) {
    const double n = norm(qvec(q));
    if (n != 0.0) {
        const double ang = 2.0 * ef::atan2(n, fabs(q.w));
        const V3 axis = (q.w < 0.0) ? (-qvec(q)) / n : qvec(q) / n;
        return ang * axis;
    }
    return v3(0.0, 0.0, 0.0);
}
ND_HD Q slerp(Q a, double t, Q b) {
    const double one = 1.0 - 2.220446049250313e-16;
    const double d = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    const double ad = fabs(d);
    double s0, s1;
    if (ad >= one) {
        s0 = 1.0 - t;
        s1 = t;
    } else {
        const double th = ef::acos(ad);
        const double st = ef::sin(th);
        s0 = ef::sin((1.0 - t) * th) / st;
        s1 = ef::sin(t * th) / st;
    }
    if (d < 0.0) s1 = -s1;
    return quat(s0 * a.w + s1 * b.w, s0 * a.x + s1 * b.x, s0 * a.y + s1 * b.y, s0 * a.z + s1 * b.z);
}

// 3x3 row-major
struct M3 {
    double m[9];
};
ND_HD M3 m3zero() { M3 r; for (int i = 0; i < 9; ++i) r.m[i] = 0.0; return r; }
ND_HD M3 skew(V3 v) {
    M3 s = m3zero();
    s.m[1] = -v.z; s.m[2] = v.y;
    s.m[3] = v.z;  s.m[5] = -v.x;
    s.m[6] = -v.y; s.m[7] = v.x;
    return s;
}
ND_HD M3 m3mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i * 3 + j] = a.m[i * 3 + 0] * b.m[0 * 3 + j] + a.m[i * 3 + 1] * b.m[1 * 3 + j] + a.m[i * 3 + 2] * b.m[2 * 3 + j];
    return r;
}
ND_HD V3 m3v(const M3& a, V3 v) {
    return v3(a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
              a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z);
}
// Inverse by cofactors; det from the first-column cofactors.
ND_HD M3 m3inverse(const M3& a) {
#define ND_A(i, j) a.m[(i) * 3 + (j)]
#define ND_COF(i, j) (ND_A(((i) + 1) % 3, ((j) + 1) % 3) * ND_A(((i) + 2) % 3, ((j) + 2) % 3) - \
                      ND_A(((i) + 1) % 3, ((j) + 2) % 3) * ND_A(((i) + 2) % 3, ((j) + 1) % 3))
    const double c00 = ND_COF(0, 0), c10 = ND_COF(1, 0), c20 = ND_COF(2, 0);
    const double det = c00 * ND_A(0, 0) + c10 * ND_A(1, 0) + c20 * ND_A(2, 0);
    const double inv = 1.0 / det;
    M3 r;
    r.m[0] = c00 * inv; r.m[1] = c10 * inv; r.m[2] = c20 * inv;
    r.m[3] = ND_COF(0, 1) * inv; r.m[4] = ND_COF(1, 1) * inv; r.m[5] = ND_COF(2, 1) * inv;
    r.m[6] = ND_COF(0, 2) * inv; r.m[7] = ND_COF(1, 2) * inv; r.m[8] = ND_COF(2, 2) * inv;
    return r;
#undef ND_COF
#undef ND_A
}

struct Box3 {
    V3 mn, mx;
};
ND_HD bool box_overlap(const Box3& a, const Box3& b) {
    return a.mn.x <= b.mx.x && a.mn.y <= b.mx.y && a.mn.z <= b.mx.z && a.mx.x >= b.mn.x && a.mx.y >= b.mn.y &&
           a.mx.z >= b.mn.z;
}
ND_HD Box3 box_inflated(const Box3& a, double r) {
    Box3 o;
    o.mn = v3(a.mn.x - r, a.mn.y - r, a.mn.z - r);
    o.mx = v3(a.mx.x + r, a.mx.y + r, a.mx.z + r);
    return o;
}
// aabb_distance (ContactAssembler.cpp:45-55): std::max({0, a, b}) keeps the first
// largest value.
ND_HD double max3(double a, double b, double c) {
    double m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    return m;
}
ND_HD double box_distance(const Box3& a, const Box3& b) {
    const double dx = max3(0.0, a.mn.x - b.mx.x, b.mn.x - a.mx.x);
    const double dy = max3(0.0, a.mn.y - b.mx.y, b.mn.y - a.mx.y);
    const double dz = max3(0.0, a.mn.z - b.mx.z, b.mn.z - a.mx.z);
    return sqrt(dx * dx + dy * dy + dz * dz);
}
// point_aabb_distance (ContactAssembler.cpp:57-65)
ND_HD double point_box_distance(V3 p, const Box3& b) {
    const double dx = max3(b.mn.x - p.x, 0.0, p.x - b.mx.x);
    const double dy = max3(b.mn.y - p.y, 0.0, p.y - b.mx.y);
    const double dz = max3(b.mn.z - p.z, 0.0, p.z - b.mx.z);
    return sqrt(dx * dx + dy * dy + dz * dz);
}
// transform_local_aabb (ContactAssembler.cpp:22-43): bounds of the 8 rotated corners.
ND_HD Box3 world_box(V3 lmin, V3 lmax, V3 p, Q q) {
    const double big = 1.7976931348623157e308;
    Box3 b;
    b.mn = v3(big, big, big);
    b.mx = v3(-big, -big, -big);
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int iz = 0; iz < 2; ++iz) {
                const V3 c = v3(ix ? lmax.x : lmin.x, iy ? lmax.y : lmin.y, iz ? lmax.z : lmin.z);
                const V3 w = p + rotate(q, c);
                b.mn = vmin(b.mn, w);
                b.mx = vmax(b.mx, w);
            }
    return b;
}

// ------------------------------------------------------------------ directed rounding
// Boost.Interval rounds every endpoint outward. gfx950 has no per-instruction
// directed-rounding FP64 adds, so the rounding direction is recovered from the exact
// error of the round-to-nearest result (TwoSum, Knuth; TwoProduct via FMA,
// Ogita-Rump-Oishi 2005) and the result is stepped one ulp when the error points the
// other way. Exact for finite, non-underflowing operands.
ND_HD double next_down(double s) {
    if (s != s) return s;
    if (s == 0.0) return -4.9406564584124654e-324;
    int64_t b = __builtin_bit_cast(int64_t, s);
    b += (s > 0.0) ? -1 : 1;
    return __builtin_bit_cast(double, b);
}
ND_HD double next_up(double s) { return -next_down(-s); }
ND_HD double two_sum_err(double a, double b, double s) {
    const double bb = s - a;
    return (a - (s - bb)) + (b - bb);
}
ND_HD double add_dn(double a, double b) {
    const double s = a + b;
    if (s == 0.0) {
        const bool pos = (a == 0.0 && b == 0.0 && !__builtin_signbit(a) && !__builtin_signbit(b));
        return pos ? 0.0 : -0.0;
    }
    return two_sum_err(a, b, s) < 0.0 ? next_down(s) : s;
}
ND_HD double add_up(double a, double b) {
    const double s = a + b;
    if (s == 0.0) return s;
    return two_sum_err(a, b, s) > 0.0 ? next_up(s) : s;
}
ND_HD double sub_dn(double a, double b) { return add_dn(a, -b); }
ND_HD double sub_up(double a, double b) { return add_up(a, -b); }
ND_HD double mul_dn(double a, double b) {
    const double p = a * b;
    if (p == 0.0) return p;
    return fma(a, b, -p) < 0.0 ? next_down(p) : p;
}
ND_HD double mul_up(double a, double b) {
    const double p = a * b;
    if (p == 0.0) return p;
    return fma(a, b, -p) > 0.0 ? next_up(p) : p;
}

// Interval with Boost.Interval semantics (interval<double, save_state<rounded_transc_std>,
// checking_catch_nan>): outward rounding, sign-case products, median rounds to nearest,
// width rounds up. NaN endpoints set a flag (the reference throws).
struct Iv {
    double lo, hi;
};
ND_HD Iv iv(double l, double h) { Iv r; r.lo = l; r.hi = h; return r; }
ND_HD bool iv_nan(Iv a) { return a.lo != a.lo || a.hi != a.hi; }
ND_HD Iv operator+(Iv a, Iv b) { return iv(add_dn(a.lo, b.lo), add_up(a.hi, b.hi)); }
ND_HD Iv operator-(Iv a, Iv b) { return iv(sub_dn(a.lo, b.hi), sub_up(a.hi, b.lo)); }
ND_HD Iv operator-(double a, Iv b) { return iv(sub_dn(a, b.hi), sub_up(a, b.lo)); }
ND_HD Iv operator*(Iv x, Iv y) {
    const double xl = x.lo, xu = x.hi, yl = y.lo, yu = y.hi;
    if (xu < 0) {
        if (yu < 0) return iv(mul_dn(xu, yu), mul_up(xl, yl));
        if (yl < 0) return iv(mul_dn(xl, yu), mul_up(xl, yl));
        return iv(mul_dn(xl, yu), mul_up(xu, yl));
    }
    if (xl < 0) {
        if (yu < 0) return iv(mul_dn(xu, yl), mul_up(xl, yl));
        if (yl < 0) return iv(dmin(mul_dn(xl, yu), mul_dn(xu, yl)), dmax(mul_up(xl, yl), mul_up(xu, yu)));
        return iv(mul_dn(xl, yu), mul_up(xu, yu));
    }
    if (yu < 0) return iv(mul_dn(xu, yl), mul_up(xl, yu));
    if (yl < 0) return iv(mul_dn(xu, yl), mul_up(xu, yu));
    return iv(mul_dn(xl, yl), mul_up(xu, yu));
}
ND_HD double iv_width(Iv a) { return sub_up(a.hi, a.lo); }
ND_HD double iv_median(Iv a) { return (a.lo + a.hi) / 2; }
ND_HD bool iv_zero_in(Iv a) { return a.lo <= 0.0 && 0.0 <= a.hi; }

}  // namespace nd
