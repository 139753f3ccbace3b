// Forward-mode dual numbers and the element functions the adjoint differentiates.
//
// The adjoint of a step needs, besides the Newton Hessian, the derivatives of the
// objective's gradient with respect to the step-start pose (friction's lagged frame and
// the joints' reference rotations), the friction coefficient, and the derivatives of the
// quaternion update. Those are small dense Jacobians of short straight-line formulas,
// evaluated here with Dn<N> (value + N directional derivatives, each operation applying
// the chain rule), which gives them to rounding without hand-derived expressions. The
// formulas follow the operation order of dmath.h.
//
// Host and device: the oracle's CPU adjoint includes this header too.
#pragma once
#include <math.h>

#include "elemfn.h"

#if defined(__HIPCC__)
#define ND_DU __host__ __device__ __forceinline__
#else
#define ND_DU inline
#endif

namespace nd {
namespace du {

template <int N>
struct Dn {
    double v;
    double d[N];
};

template <int N>
ND_DU Dn<N> cst(double v) {
    Dn<N> r;
    r.v = v;
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}
template <int N>
ND_DU Dn<N> var(double v, int i) {
    Dn<N> r = cst<N>(v);
    r.d[i] = 1.0;
    return r;
}
template <int N>
ND_DU Dn<N> operator+(const Dn<N>& a, const Dn<N>& b) {
    Dn<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
ND_DU Dn<N> operator-(const Dn<N>& a, const Dn<N>& b) {
    Dn<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
ND_DU Dn<N> operator-(const Dn<N>& a) {
    Dn<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
ND_DU Dn<N> operator*(const Dn<N>& a, const Dn<N>& b) {
    Dn<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
ND_DU Dn<N> operator/(const Dn<N>& a, const Dn<N>& b) {
    Dn<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N>
ND_DU Dn<N> operator+(const Dn<N>& a, double b) { Dn<N> r = a; r.v = a.v + b; return r; }
template <int N>
ND_DU Dn<N> operator+(double a, const Dn<N>& b) { return b + a; }
template <int N>
ND_DU Dn<N> operator-(const Dn<N>& a, double b) { Dn<N> r = a; r.v = a.v - b; return r; }
template <int N>
ND_DU Dn<N> operator-(double a, const Dn<N>& b) {
    Dn<N> r;
    r.v = a - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
    return r;
}
template <int N>
ND_DU Dn<N> operator*(const Dn<N>& a, double b) {
    Dn<N> r;
    r.v = a.v * b;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N>
ND_DU Dn<N> operator*(double a, const Dn<N>& b) { return b * a; }
template <int N>
ND_DU Dn<N> operator/(const Dn<N>& a, double b) {
    Dn<N> r;
    r.v = a.v / b;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
    return r;
}
template <int N>
ND_DU Dn<N> operator/(double a, const Dn<N>& b) { return cst<N>(a) / b; }

ND_DU double val(double x) { return x; }
ND_DU double with_val(double, double v) { return v; }
template <int N>
ND_DU double val(const Dn<N>& x) { return x.v; }
template <int N>
ND_DU Dn<N> with_val(Dn<N> x, double v) {
    x.v = v;
    return x;
}

ND_DU double dsqrt(double x) { return sqrt(x); }
template <int N>
ND_DU Dn<N> dsqrt(const Dn<N>& a) {
    Dn<N> r;
    r.v = sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}
ND_DU void dsincos(double x, double* s, double* c) { ef::sincos(x, s, c); }
template <int N>
ND_DU void dsincos(const Dn<N>& a, Dn<N>* s, Dn<N>* c) {
    double sv, cv;
    ef::sincos(a.v, &sv, &cv);
    s->v = sv;
    c->v = cv;
    for (int k = 0; k < N; ++k) {
        s->d[k] = cv * a.d[k];
        c->d[k] = -sv * a.d[k];
    }
}
ND_DU double datan2(double y, double x) { return ef::atan2(y, x); }
template <int N>
ND_DU Dn<N> datan2(const Dn<N>& y, const Dn<N>& x) {
    Dn<N> r;
    r.v = ef::atan2(y.v, x.v);
    const double den = x.v * x.v + y.v * y.v;
    for (int k = 0; k < N; ++k) r.d[k] = den > 0.0 ? (x.v * y.d[k] - y.v * x.d[k]) / den : 0.0;
    return r;
}
ND_DU double dfabs(double x) { return fabs(x); }
template <int N>
ND_DU Dn<N> dfabs(const Dn<N>& a) { return a.v < 0.0 ? -a : a; }

template <typename T>
struct Lift;
template <>
struct Lift<double> {
    ND_DU static double f(double v) { return v; }
};
template <int N>
struct Lift<Dn<N>> {
    ND_DU static Dn<N> f(double v) { return cst<N>(v); }
};
template <typename T>
ND_DU T lift(double v) { return Lift<T>::f(v); }

// ------------------------------------------------------------------ geometry over T
template <typename T>
struct V3T {
    T x, y, z;
};
template <typename T>
struct QT {
    T w, x, y, z;
};
template <typename T>
ND_DU V3T<T> mk3(T a, T b, T c) { V3T<T> r; r.x = a; r.y = b; r.z = c; return r; }
template <typename T>
ND_DU QT<T> mkq(T w, T x, T y, T z) { QT<T> r; r.w = w; r.x = x; r.y = y; r.z = z; return r; }
template <typename T>
ND_DU V3T<T> lift3(double a, double b, double c) { return mk3<T>(lift<T>(a), lift<T>(b), lift<T>(c)); }
template <typename T>
ND_DU QT<T> liftq(const double* q) { return mkq<T>(lift<T>(q[0]), lift<T>(q[1]), lift<T>(q[2]), lift<T>(q[3])); }
template <typename T>
ND_DU V3T<T> add(V3T<T> a, V3T<T> b) { return mk3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T>
ND_DU V3T<T> sub(V3T<T> a, V3T<T> b) { return mk3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T, typename S>
ND_DU V3T<T> scl(S s, V3T<T> a) { return mk3<T>(s * a.x, s * a.y, s * a.z); }
template <typename T, typename S>
ND_DU V3T<T> dvs(V3T<T> a, S s) { return mk3<T>(a.x / s, a.y / s, a.z / s); }
template <typename T>
ND_DU T dot(V3T<T> a, V3T<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T>
ND_DU V3T<T> cross(V3T<T> a, V3T<T> b) {
    return mk3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename T>
ND_DU QT<T> qmul(QT<T> a, QT<T> b) {
    return mkq<T>(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                  a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z, a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x);
}
template <typename T>
ND_DU QT<T> qconj(QT<T> q) { return mkq<T>(q.w, -q.x, -q.y, -q.z); }
template <typename T>
ND_DU QT<T> qnormalized(QT<T> q) {
    const T n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (val(n2) > 0.0) {
        const T n = dsqrt(n2);
        return mkq<T>(q.w / n, q.x / n, q.y / n, q.z / n);
    }
    return q;
}
// q * v as v + w uv + qv x uv, uv = 2 (qv x v)
template <typename T>
ND_DU V3T<T> rotate(QT<T> q, V3T<T> v) {
    const V3T<T> qv = mk3<T>(q.x, q.y, q.z);
    V3T<T> uv = cross(qv, v);
    uv = add(uv, uv);
    return add(add(v, scl(q.w, uv)), cross(qv, uv));
}
// normalize(AngleAxis(|th|, th/|th|) * q) for |th| > thr (apply) or >= thr (update)
template <typename T>
ND_DU QT<T> rot_update(QT<T> q, V3T<T> th, bool strict) {
    const T ang = dsqrt(dot(th, th));
    const bool skip = strict ? !(val(ang) > 1e-10) : (val(ang) < 1e-10);
    if (skip) {
        // below the threshold the update leaves q as it is; the derivative taken is that
        // of the rotation exp(th/2) q which it stands for (at th = 0: (0, th/2) q), the
        // limit that finite differences with |th| above the threshold see
        const QT<T> lin = qmul(mkq<T>(lift<T>(1.0), th.x * 0.5, th.y * 0.5, th.z * 0.5), q);
        return mkq<T>(with_val(lin.w, val(q.w)), with_val(lin.x, val(q.x)), with_val(lin.y, val(q.y)),
                      with_val(lin.z, val(q.z)));
    }
    const V3T<T> axis = dvs(th, ang);
    T s, c;
    dsincos(ang * 0.5, &s, &c);
    const QT<T> aa = mkq<T>(c, s * axis.x, s * axis.y, s * axis.z);
    return qnormalized(qmul(aa, q));
}
// rotation vector of a unit quaternion (Eigen AngleAxis): 2 atan2(|v|, |w|) axis
template <typename T>
ND_DU V3T<T> rotation_vector(QT<T> q) {
    const V3T<T> v = mk3<T>(q.x, q.y, q.z);
    const T n = dsqrt(dot(v, v));
    if (val(n) != 0.0) {
        const T ang = datan2(n, dfabs(q.w)) * 2.0;
        const V3T<T> axis = val(q.w) < 0.0 ? dvs(scl(-1.0, v), n) : dvs(v, n);
        return scl(ang, axis);
    }
    // |v| = 0: the value is 0, and the derivative is that of the limit 2 sign(w) v / |w|
    const T sc = (val(q.w) < 0.0 ? -2.0 : 2.0) / dfabs(q.w);
    return scl(sc, v);
}
// 3x3 helpers, row-major
template <typename T>
struct M3T {
    T m[9];
};
template <typename T>
ND_DU M3T<T> skew(V3T<T> a) {
    M3T<T> r;
    const T z = lift<T>(0.0);
    r.m[0] = z;    r.m[1] = -a.z; r.m[2] = a.y;
    r.m[3] = a.z;  r.m[4] = z;    r.m[5] = -a.x;
    r.m[6] = -a.y; r.m[7] = a.x;  r.m[8] = z;
    return r;
}
template <typename T>
ND_DU M3T<T> m3mul(const M3T<T>& a, const M3T<T>& b) {
    M3T<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = a.m[3 * i] * b.m[j] + a.m[3 * i + 1] * b.m[3 + j] + a.m[3 * i + 2] * b.m[6 + j];
    return r;
}

// ------------------------------------------------------------------ element functions
// Friction gradient of one contact (FrictionForm::gradient, FrictionForm.cpp:44-90 with
// FrictionPotential::gradient, Friction.cpp:20-31): g[0..5] on body A, g[6..11] on B,
// as a function of the Newton variables of A (xa[6]) and B (xb[6]), of the step-start
// poses (pA0, qA0, pB0, qB0), of mu, and of the contact point P, normal nr and lagged
// normal force N (which follow the state the narrow phase ran at).
template <typename T>
ND_DU void friction_grad_g(const T* xa, const T* xb, V3T<T> pA0, QT<T> qA0, V3T<T> pB0, QT<T> qB0, V3T<T> P,
                           V3T<T> nr, T N, T mu, double eps, double dt, T* g) {
    const V3T<T> pA = mk3<T>(xa[0], xa[1], xa[2]), pB = mk3<T>(xb[0], xb[1], xb[2]);
    const V3T<T> thA = mk3<T>(xa[3], xa[4], xa[5]), thB = mk3<T>(xb[3], xb[4], xb[5]);
    const QT<T> qA = rot_update(qA0, thA, true), qB = rot_update(qB0, thB, true);
    const V3T<T> dA = sub(pA, pA0), dB = sub(pB, pB0);
    const V3T<T> lA = rotate(qconj(qA0), sub(P, pA0)), lB = rotate(qconj(qB0), sub(P, pB0));
    const V3T<T> rA = rotate(qA, lA), rB = rotate(qB, lB);
    V3T<T> wA = lift3<T>(0.0, 0.0, 0.0), wB = wA, vA = wA, vB = wA;
    if (dt > 1e-12) {
        wA = dvs(thA, dt);
        wB = dvs(thB, dt);
        vA = dvs(dA, dt);
        vB = dvs(dB, dt);
    }
    vA = add(vA, cross(wA, rA));
    vB = add(vB, cross(wB, rB));
    const V3T<T> rel = sub(vB, vA);
    T n0 = nr.x, n1 = nr.y, n2 = nr.z;
    const T nn2 = n0 * n0 + n1 * n1 + n2 * n2;
    if (val(nn2) > 1e-16) {
        const T l = dsqrt(nn2);
        n0 = n0 / l; n1 = n1 / l; n2 = n2 / l;
    } else {
        n0 = lift<T>(0.0); n1 = lift<T>(1.0); n2 = lift<T>(0.0);
    }
    const T nv[3] = {n0, n1, n2};
    const T rn = rel.x * n0 + rel.y * n1 + rel.z * n2;
    const V3T<T> tv = mk3<T>(rel.x - rn * n0, rel.y - rn * n1, rel.z - rn * n2);
    const V3T<T> u = scl(dt, tv);
    // FrictionPotential::gradient
    V3T<T> fg = lift3<T>(0.0, 0.0, 0.0);
    if (val(N) > 0) {
        const T un = dsqrt(dot(u, u));
        if (!(val(un) < 1e-12)) {
            const T sn = dsqrt(un * un + eps * eps);
            const T s = mu * N / sn;
            fg = scl(s, u);
        }
    }
    // J_A = [-P, P [rA]x], J_B = [P, -P [rB]x], g = J^T fg
    M3T<T> Pm;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Pm.m[3 * i + j] = (i == j ? 1.0 : 0.0) - nv[i] * nv[j];
    const M3T<T> PA = m3mul(Pm, skew(rA)), PB = m3mul(Pm, skew(rB));
    const T f[3] = {fg.x, fg.y, fg.z};
    for (int j = 0; j < 3; ++j) {
        g[j] = (-Pm.m[j]) * f[0] + (-Pm.m[3 + j]) * f[1] + (-Pm.m[6 + j]) * f[2];
        g[3 + j] = PA.m[j] * f[0] + PA.m[3 + j] * f[1] + PA.m[6 + j] * f[2];
        g[6 + j] = Pm.m[j] * f[0] + Pm.m[3 + j] * f[1] + Pm.m[6 + j] * f[2];
        g[9 + j] = (-PB.m[j]) * f[0] + (-PB.m[3 + j]) * f[1] + (-PB.m[6 + j]) * f[2];
    }
}
// The same with a fixed contact point, normal and normal force.
template <typename T>
ND_DU void friction_grad_t(const T* xa, const T* xb, V3T<T> pA0, QT<T> qA0, V3T<T> pB0, QT<T> qB0,
                           const double* pt, const double* nr, double N, T mu, double eps, double dt, T* g) {
    friction_grad_g<T>(xa, xb, pA0, qA0, pB0, qB0, lift3<T>(pt[0], pt[1], pt[2]), lift3<T>(nr[0], nr[1], nr[2]),
                       lift<T>(N), mu, eps, dt, g);
}

// ------------------------------------------------------------------ contact geometry
// The narrow phase's point and normal of one contact (IPCSolver.cpp:32-134, 394-806) as
// functions of the world vertices of its primitives and the two body positions; the
// branch decisions follow the values, so these are the derivatives of the formulas on
// the branch the narrow phase took.
// safeNormalized (IPCSolver.cpp:92-105)
template <typename T>
ND_DU V3T<T> safe_normalized_t(V3T<T> v, V3T<T> fb) {
    const T n = dsqrt(dot(v, v));
    if (val(n) > 1e-12) return dvs(v, n);
    const T fn = dsqrt(dot(fb, fb));
    if (val(fn) > 1e-12) return dvs(fb, fn);
    return lift3<T>(0.0, 1.0, 0.0);
}
template <typename T>
ND_DU T clamp01_t(T x) { return val(x) < 0.0 ? lift<T>(0.0) : (1.0 < val(x) ? lift<T>(1.0) : x); }
// closestPointsOnSegments (IPCSolver.cpp:32-90)
template <typename T>
ND_DU void closest_points_t(V3T<T> p1, V3T<T> q1, V3T<T> p2, V3T<T> q2, V3T<T>& c1, V3T<T>& c2) {
    const double eps = 1e-12;
    const V3T<T> d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    T s, t;
    if (val(a) <= eps && val(e) <= eps) {
        c1 = p1;
        c2 = p2;
        return;
    }
    if (val(a) <= eps) {
        s = lift<T>(0.0);
        t = clamp01_t(f / e);
    } else {
        const T c = dot(d1, r);
        if (val(e) <= eps) {
            t = lift<T>(0.0);
            s = clamp01_t(-c / a);
        } else {
            const T b = dot(d1, d2);
            const T den = a * e - b * b;
            s = (fabs(val(den)) > eps) ? clamp01_t((b * f - c * e) / den) : lift<T>(0.0);
            t = (b * s + f) / e;
            if (val(t) < 0.0) {
                t = lift<T>(0.0);
                s = clamp01_t(-c / a);
            } else if (val(t) > 1.0) {
                t = lift<T>(1.0);
                s = clamp01_t((b - c) / a);
            }
        }
    }
    c1 = add(p1, scl(s, d1));
    c2 = add(p2, scl(t, d2));
}
// kind 0 face-vertex: v[0..2] the face of body A, v[3] the vertex of body B;
// kind 1 edge-edge: v[0], v[1] the edge of A, v[2], v[3] the edge of B;
// kind 2 vertex-vertex: v[0] of A, v[1] of B. pA, pB: the body positions; the normal
// is oriented towards pB - pA (append_contact).
template <typename T>
ND_DU void contact_geom_t(int kind, const V3T<T>* v, V3T<T> pA, V3T<T> pB, V3T<T>& point, V3T<T>& normal) {
    V3T<T> nrm;
    if (kind == 0) {
        const V3T<T> nn = safe_normalized_t(cross(sub(v[1], v[0]), sub(v[2], v[0])), lift3<T>(0.0, 1.0, 0.0));
        const T sd = dot(sub(v[3], v[0]), nn);
        const V3T<T> proj = sub(v[3], scl(sd, nn));
        point = scl(0.5, add(proj, v[3]));
        nrm = val(sd) >= 0.0 ? nn : scl(-1.0, nn);
    } else if (kind == 1) {
        V3T<T> c0, c1;
        closest_points_t(v[0], v[1], v[2], v[3], c0, c1);
        point = scl(0.5, add(c0, c1));
        nrm = sub(c1, c0);
    } else {
        point = scl(0.5, add(v[0], v[1]));
        nrm = sub(v[1], v[0]);
    }
    const V3T<T> ab = sub(pB, pA);
    normal = safe_normalized_t(nrm, ab);
    if (val(dot(normal, ab)) < 0.0) normal = scl(-1.0, normal);
}

// Joint gradient J^T (mu C + lambda) (Joint.cpp:21-45) and the constraint C, as
// functions of the Newton variables of both bodies and of the reference rotations
// q_refA (and q_refB for a hinge). type 0 FixedJoint (jd: target p(3), q(4)),
// 1 HingeJoint (jd: anchorA, anchorB, axisA, axisB). g[0..5] on A, g[6..11] on B.
template <typename T>
ND_DU void joint_grad_t(int type, const double* jd, const T* xa, const T* xb, QT<T> qrA, QT<T> qrB, double mu,
                        const double* lambda, T* g, T* C) {
    const T z = lift<T>(0.0);
    for (int k = 0; k < 12; ++k) g[k] = z;
    if (type == 0) {
        const QT<T> tq = mkq<T>(lift<T>(jd[3]), lift<T>(jd[4]), lift<T>(jd[5]), lift<T>(jd[6]));
        const V3T<T> tt = rotation_vector(qmul(tq, qconj(qrA)));
        C[0] = xa[0] - jd[0];
        C[1] = xa[1] - jd[1];
        C[2] = xa[2] - jd[2];
        C[3] = xa[3] - tt.x;
        C[4] = xa[4] - tt.y;
        C[5] = xa[5] - tt.z;
        for (int r = 0; r < 6; ++r) g[r] = C[r] * mu + lambda[r];
        return;
    }
    const V3T<T> anA = lift3<T>(jd[0], jd[1], jd[2]), anB = lift3<T>(jd[3], jd[4], jd[5]);
    const V3T<T> axA = lift3<T>(jd[6], jd[7], jd[8]), axB = lift3<T>(jd[9], jd[10], jd[11]);
    const V3T<T> pA = mk3<T>(xa[0], xa[1], xa[2]), pB = mk3<T>(xb[0], xb[1], xb[2]);
    const QT<T> qA = rot_update(qrA, mk3<T>(xa[3], xa[4], xa[5]), false);
    const QT<T> qB = rot_update(qrB, mk3<T>(xb[3], xb[4], xb[5]), false);
    const V3T<T> rA = rotate(qA, anA), rB = rotate(qB, anB);
    const V3T<T> pv = sub(add(pA, rA), add(pB, rB));
    const V3T<T> nA = rotate(qA, axA), nB = rotate(qB, axB);
    const V3T<T> ac = cross(nA, nB);
    C[0] = pv.x; C[1] = pv.y; C[2] = pv.z; C[3] = ac.x; C[4] = ac.y; C[5] = ac.z;
    (void)z;
    T f[6];
    for (int r = 0; r < 6; ++r) f[r] = C[r] * mu + lambda[r];
    // Jacobian rows (HingeJoint via JointPrimitives.h): translation rows
    // [I, -[rA]x, -I, [rB]x], rotation rows [0, [nB]x[nA]x, 0, -[nA]x[nB]x]
    const M3T<T> sA = skew(rA), sB = skew(rB);
    const M3T<T> dA = m3mul(skew(nB), skew(nA));
    M3T<T> mnA = skew(nA);
    for (int k = 0; k < 9; ++k) mnA.m[k] = -mnA.m[k];
    const M3T<T> dB = m3mul(mnA, skew(nB));
    for (int c = 0; c < 3; ++c) {
        g[c] = f[c];
        g[6 + c] = -f[c];
        T a3 = z, b3 = z;
        for (int i = 0; i < 3; ++i) {
            a3 = a3 + (-sA.m[3 * i + c]) * f[i];
            b3 = b3 + sB.m[3 * i + c] * f[i];
        }
        for (int i = 0; i < 3; ++i) {
            a3 = a3 + dA.m[3 * i + c] * f[3 + i];
            b3 = b3 + dB.m[3 * i + c] * f[3 + i];
        }
        g[3 + c] = a3;
        g[9 + c] = b3;
    }
}

}  // namespace du
}  // namespace nd
