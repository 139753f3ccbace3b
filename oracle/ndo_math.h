// ORACLE — test infrastructure only. Nothing in the product links or calls this code.
//
// Small fixed-size linear algebra for the CPU restatement of NexDynIPC's contact
// stage. The reference uses Eigen; the formulas below follow the operation order of
// Eigen's public algorithms (re-derived, not copied) so that the restatement rounds
// the way an unvectorised Eigen build does:
//   - dot / squaredNorm: left-to-right sum (a0*b0 + a1*b1) + a2*b2
//   - cross: (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   - quaternion * vector: uv = 2 (q.vec x v);  v + w*uv + q.vec x uv
//   - Hamilton product, AngleAxis -> quaternion (cos/sin of the half angle)
//   - quaternion -> AngleAxis: angle = 2 atan2(|vec|, |w|), axis = sign(w) vec/|vec|
//   - normalized(): divide every coefficient by sqrt(squaredNorm)
//   - slerp with the 1 - machine-epsilon linear fallback
//   - 3x3 inverse by cofactors
// Eigen's SSE paths (vectorised quaternion product and 4-wide reductions) round
// differently in the last bit; that difference is "parity unpinned" against the
// real reference and is recorded in DESIGN.md.
#pragma once
#include <cmath>

// log, sin, cos, atan2 and acos come from the library's elemfn.h: built from correctly
// rounded IEEE operations only, they give the same bits here as on the GPU, whereas two
// libms may differ in the last bit. They are pinned against the C library (<= 2 ulp).
#include "elemfn.h"
#include "vecsum.h"
#include <limits>
#include <algorithm>

namespace ndo {

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(double s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, const V3& b) { a = a + b; return a; }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double sqnorm(const V3& a) { return dot(a, a); }
inline double norm(const V3& a) { return std::sqrt(sqnorm(a)); }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 normalized(const V3& a) {
    const double n2 = sqnorm(a);
    if (n2 > 0.0) return a / std::sqrt(n2);
    return a;
}
inline V3 cmin(const V3& a, const V3& b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 cmax(const V3& a, const V3& b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline double absmax(const V3& a) { return std::max(std::max(std::abs(a.x), std::abs(a.y)), std::abs(a.z)); }

struct Q {  // w + (x, y, z)
    double w = 1, x = 0, y = 0, z = 0;
    Q() = default;
    Q(double w_, double x_, double y_, double z_) : w(w_), x(x_), y(y_), z(z_) {}
    V3 vec() const { return {x, y, z}; }
};
inline V3 rotate(const Q& q, const V3& v) {
    V3 uv = cross(q.vec(), v);
    uv = uv + uv;
    return v + q.w * uv + cross(q.vec(), uv);
}
inline Q mul(const Q& a, const Q& b) {
    return Q(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
             a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
             a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
             a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x);
}
inline Q conj(const Q& q) { return Q(q.w, -q.x, -q.y, -q.z); }
inline double qsqnorm(const Q& q) { return q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w; }
inline Q qnormalized(const Q& q) {
    const double n2 = qsqnorm(q);
    if (n2 > 0.0) {
        const double n = std::sqrt(n2);
        return Q(q.w / n, q.x / n, q.y / n, q.z / n);
    }
    return q;
}
inline Q from_angle_axis(double angle, const V3& axis) {
    const double ha = 0.5 * angle;
    double s, c;
    nd::ef::sincos(ha, &s, &c);
    return Q(c, s * axis.x, s * axis.y, s * axis.z);
}
// Incremental rotation used throughout IPCSolver / FrictionForm / JointPrimitives:
// q(theta) = normalize(AngleAxis(|theta|, theta/|theta|) * q_ref) if |theta| > 1e-10.
inline Q apply_theta(const Q& q_ref, const V3& theta) {
    const double angle = norm(theta);
    if (angle > 1e-10) {
        const V3 axis = theta / angle;
        return qnormalized(mul(from_angle_axis(angle, axis), q_ref));
    }
    return q_ref;
}
// JointPrimitives::updateRotation uses "< 1e-10 -> q_ref" (same as above for all
// inputs except |theta| == 1e-10 exactly).
inline Q update_rotation(const Q& q_ref, const V3& theta) {
    const double angle = norm(theta);
    if (angle < 1e-10) return q_ref;
    const V3 axis = theta / angle;
    return qnormalized(mul(from_angle_axis(angle, axis), q_ref));
}
// AngleAxis(q): angle*axis as a rotation vector.
inline V3 to_rotation_vector(const Q& q) {
    double n = norm(q.vec());
    if (n != 0.0) {
        const double angle = 2.0 * nd::ef::atan2(n, std::abs(q.w));
        const V3 axis = (q.w < 0.0) ? (-q.vec()) / n : q.vec() / n;
        return angle * axis;
    }
    return V3(0.0, 0.0, 0.0);
}
inline Q slerp(const Q& a, double t, const Q& b) {
    const double one = 1.0 - std::numeric_limits<double>::epsilon();
    const double d = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    const double absD = std::abs(d);
    double scale0, scale1;
    if (absD >= one) {
        scale0 = 1.0 - t;
        scale1 = t;
    } else {
        const double theta = nd::ef::acos(absD);
        const double sinTheta = nd::ef::sin(theta);
        scale0 = nd::ef::sin((1.0 - t) * theta) / sinTheta;
        scale1 = nd::ef::sin(t * theta) / sinTheta;
    }
    if (d < 0.0) scale1 = -scale1;
    return Q(scale0 * a.w + scale1 * b.w, scale0 * a.x + scale1 * b.x,
             scale0 * a.y + scale1 * b.y, scale0 * a.z + scale1 * b.z);
}

struct M3 {
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double operator()(int i, int j) const { return m[i][j]; }
    double& operator()(int i, int j) { return m[i][j]; }
    static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0; return r; }
};
inline V3 mul(const M3& a, const V3& v) {
    return {a(0, 0) * v.x + a(0, 1) * v.y + a(0, 2) * v.z,
            a(1, 0) * v.x + a(1, 1) * v.y + a(1, 2) * v.z,
            a(2, 0) * v.x + a(2, 1) * v.y + a(2, 2) * v.z};
}
inline M3 mul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r;
}
inline M3 outer(const V3& a, const V3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = a[i] * b[j];
    return r;
}
inline M3 skew(const V3& v) {
    M3 s;
    s(0, 1) = -v.z; s(0, 2) = v.y;
    s(1, 0) = v.z;  s(1, 2) = -v.x;
    s(2, 0) = -v.y; s(2, 1) = v.x;
    return s;
}
// Inverse by cofactors: det from the first column's cofactors, inv(i,j) = cof(j,i)/det.
inline M3 inverse(const M3& a) {
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return a(i1, j1) * a(i2, j2) - a(i1, j2) * a(i2, j1);
    };
    const double c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const double det = c00 * a(0, 0) + c10 * a(1, 0) + c20 * a(2, 0);
    const double invdet = 1.0 / det;
    M3 r;
    r(0, 0) = c00 * invdet; r(0, 1) = c10 * invdet; r(0, 2) = c20 * invdet;
    r(1, 0) = cof(0, 1) * invdet; r(1, 1) = cof(1, 1) * invdet; r(1, 2) = cof(2, 1) * invdet;
    r(2, 0) = cof(0, 2) * invdet; r(2, 1) = cof(1, 2) * invdet; r(2, 2) = cof(2, 2) * invdet;
    return r;
}

// Physics::Geometry::AABB (nexdynipc/include/Physics/Geometry/AABB.h:15-236).
struct AABB {
    V3 mn{std::numeric_limits<double>::max(), std::numeric_limits<double>::max(), std::numeric_limits<double>::max()};
    V3 mx{std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest()};
    AABB() = default;
    AABB(const V3& a, const V3& b) : mn(a), mx(b) {}
    V3 center() const { return (mn + mx) * 0.5; }
    V3 extent() const { return mx - mn; }
    double surfaceArea() const {
        const V3 e = extent();
        return 2.0 * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
    void expand(const V3& p) { mn = cmin(mn, p); mx = cmax(mx, p); }
    void expand(const AABB& o) { mn = cmin(mn, o.mn); mx = cmax(mx, o.mx); }
    AABB merged(const AABB& o) const { return AABB(cmin(mn, o.mn), cmax(mx, o.mx)); }
    // AABB.h:149-153 — touching boxes count as overlapping.
    bool intersects(const AABB& o) const {
        return mn.x <= o.mx.x && mn.y <= o.mx.y && mn.z <= o.mx.z &&
               mx.x >= o.mn.x && mx.y >= o.mn.y && mx.z >= o.mn.z;
    }
    AABB inflated(double r) const {
        return AABB(V3(mn.x - r, mn.y - r, mn.z - r), V3(mx.x + r, mx.y + r, mx.z + r));
    }
    int maxExtentAxis() const {
        const V3 e = extent();
        if (e.x >= e.y && e.x >= e.z) return 0;
        if (e.y >= e.z) return 1;
        return 2;
    }
};

}  // namespace ndo
