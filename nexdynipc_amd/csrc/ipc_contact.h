// Contact energy of the `ipc` mode: the primitive-distance barrier of rigid IPC that the
// reference's authors plan to move to (documents/2026.02.28/
// NexDynIPC_接触链路rigid-ipc完整替换实施规划.md:37-81) in place of the centre-distance
// barrier of DistanceBarrierForm (ref mode). Per
// point-triangle (PT) or edge-edge (EE) pair of world points x[12] (PT: p, t0, t1, t2;
// EE: a0, a1, b0, b1):
//   - the distance type: which Voronoi region of the triangle / which clamped closest-
//     point case of the two segments holds the closest points;
//   - the squared distance of that type (point-point, point-edge, point-plane, line-line)
//     with its gradient and Hessian in x, by hyper-dual numbers (value, 12 first and 78
//     second derivatives carried through every operation);
//   - E = kappa b(d), b the reference's BarrierPotential (Barrier.cpp:11-46) of the
//     distance d = sqrt(d2), chained through d2: E' = b'(d) / (2 d),
//     E'' = (b''(d) - b'(d) / d) / (4 d^2);
//   - the 12 x 12 Hessian projected to the positive semidefinite cone (cyclic Jacobi
//     eigen-decomposition, negative eigenvalues set to 0);
//   - the fold through the rigid-body map x_v = p_b + R(theta_b) r_v into the two
//     bodies' 6 x 6 blocks (Gauss-Newton in the map: J^T H J, gradient J^T g).
// Host and device: the oracle includes this header too.
#pragma once
#include "dmath.h"
#include "dual.h"

namespace nd {
namespace ipc {

constexpr int kN = 12;
constexpr int kH = kN * (kN + 1) / 2;

// index of (i, j) in the packed upper triangle, row-major
ND_DU int hix(int i, int j) {
    if (i > j) {
        const int t = i;
        i = j;
        j = t;
    }
    return i * kN - (i * (i - 1)) / 2 + (j - i);
}

struct H12 {
    double v;
    double d[kN];
    double h[kH];
};

ND_DU H12 hcst(double v) {
    H12 r;
    r.v = v;
    for (int k = 0; k < kN; ++k) r.d[k] = 0.0;
    for (int k = 0; k < kH; ++k) r.h[k] = 0.0;
    return r;
}
ND_DU H12 hvar(double v, int i) {
    H12 r = hcst(v);
    r.d[i] = 1.0;
    return r;
}
ND_DU H12 operator+(const H12& a, const H12& b) {
    H12 r;
    r.v = a.v + b.v;
    for (int k = 0; k < kN; ++k) r.d[k] = a.d[k] + b.d[k];
    for (int k = 0; k < kH; ++k) r.h[k] = a.h[k] + b.h[k];
    return r;
}
ND_DU H12 operator-(const H12& a, const H12& b) {
    H12 r;
    r.v = a.v - b.v;
    for (int k = 0; k < kN; ++k) r.d[k] = a.d[k] - b.d[k];
    for (int k = 0; k < kH; ++k) r.h[k] = a.h[k] - b.h[k];
    return r;
}
ND_DU H12 operator*(const H12& a, const H12& b) {
    H12 r;
    r.v = a.v * b.v;
    for (int k = 0; k < kN; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    int t = 0;
    for (int i = 0; i < kN; ++i)
        for (int j = i; j < kN; ++j, ++t) r.h[t] = (a.h[t] * b.v + a.v * b.h[t]) + (a.d[i] * b.d[j] + a.d[j] * b.d[i]);
    return r;
}
// a / b through 1 / b: f = 1/b, f' = -1/b^2, f'' = 2/b^3
ND_DU H12 operator/(const H12& a, const H12& b) {
    H12 ib;
    const double f = 1.0 / b.v, f1 = -f * f, f2 = -2.0 * f1 * f;
    ib.v = f;
    for (int k = 0; k < kN; ++k) ib.d[k] = f1 * b.d[k];
    int t = 0;
    for (int i = 0; i < kN; ++i)
        for (int j = i; j < kN; ++j, ++t) ib.h[t] = f1 * b.h[t] + f2 * b.d[i] * b.d[j];
    return a * ib;
}
ND_DU double hval(double x) { return x; }
ND_DU double hval(const H12& x) { return x.v; }

template <typename T>
struct P3 {
    T x, y, z;
};
template <typename T>
ND_DU P3<T> psub(const P3<T>& a, const P3<T>& b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename T>
ND_DU T pdot(const P3<T>& a, const P3<T>& b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <typename T>
ND_DU P3<T> pcross(const P3<T>& a, const P3<T>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// squared distances of the types (T = double or H12)
template <typename T>
ND_DU T d2_point_point(const P3<T>& a, const P3<T>& b) {
    const P3<T> d = psub(a, b);
    return pdot(d, d);
}
template <typename T>
ND_DU T d2_point_edge(const P3<T>& p, const P3<T>& e0, const P3<T>& e1) {
    const P3<T> e = psub(e1, e0);
    const P3<T> c = pcross(e, psub(p, e0));
    return pdot(c, c) / pdot(e, e);
}
template <typename T>
ND_DU T d2_point_plane(const P3<T>& p, const P3<T>& t0, const P3<T>& t1, const P3<T>& t2) {
    const P3<T> n = pcross(psub(t1, t0), psub(t2, t0));
    const T s = pdot(psub(p, t0), n);
    return (s * s) / pdot(n, n);
}
template <typename T>
ND_DU T d2_line_line(const P3<T>& a0, const P3<T>& a1, const P3<T>& b0, const P3<T>& b1) {
    const P3<T> n = pcross(psub(a1, a0), psub(b1, b0));
    const T s = pdot(psub(b0, a0), n);
    return (s * s) / pdot(n, n);
}

// distance types: point-triangle 0..2 point-vertex t0..t2, 3..5 point-edge t0t1, t1t2,
// t2t0, 6 point-plane; edge-edge 0..3 point-point (a0b0, a0b1, a1b0, a1b1), 4..7
// point-edge (a0-b, a1-b, b0-a, b1-a), 8 line-line
enum { PT_T0 = 0, PT_T1, PT_T2, PT_E01, PT_E12, PT_E20, PT_FACE };
enum { EE_A0B0 = 0, EE_A0B1, EE_A1B0, EE_A1B1, EE_A0_B, EE_A1_B, EE_B0_A, EE_B1_A, EE_LINES };

ND_DU P3<double> p3(const double* x) { return {x[0], x[1], x[2]}; }

// Voronoi region of the closest point of triangle (a, b, c) to p (Ericson, Real-Time
// Collision Detection 5.1.5)
ND_DU int pt_type(const double* x) {
    const P3<double> p = p3(x), a = p3(x + 3), b = p3(x + 6), c = p3(x + 9);
    const P3<double> ab = psub(b, a), ac = psub(c, a), ap = psub(p, a);
    const double d1 = pdot(ab, ap), d2 = pdot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return PT_T0;
    const P3<double> bp = psub(p, b);
    const double d3 = pdot(ab, bp), d4 = pdot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return PT_T1;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return PT_E01;
    const P3<double> cp = psub(p, c);
    const double d5 = pdot(ab, cp), d6 = pdot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return PT_T2;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return PT_E20;
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) return PT_E12;
    return PT_FACE;
}

// case of the clamped closest points s, t of segments a0a1 and b0b1 (Ericson 5.1.9; the
// parallel case takes s = 0 as the reference's closestPointsOnSegments does)
ND_DU int ee_type(const double* x) {
    const P3<double> a0 = p3(x), a1 = p3(x + 3), b0 = p3(x + 6), b1 = p3(x + 9);
    const P3<double> d1 = psub(a1, a0), d2 = psub(b1, b0), r = psub(a0, b0);
    const double a = pdot(d1, d1), e = pdot(d2, d2), f = pdot(d2, r);
    const double eps = 1e-12;
    double s, t;
    if (a <= eps && e <= eps) {
        s = 0.0;
        t = 0.0;
    } else if (a <= eps) {
        s = 0.0;
        t = dmin(dmax(f / e, 0.0), 1.0);
    } else {
        const double c = pdot(d1, r);
        if (e <= eps) {
            t = 0.0;
            s = dmin(dmax(-c / a, 0.0), 1.0);
        } else {
            const double b = pdot(d1, d2);
            const double denom = a * e - b * b;
            s = fabs(denom) > eps * a * e ? dmin(dmax((b * f - c * e) / denom, 0.0), 1.0) : 0.0;
            t = (b * s + f) / e;
            if (t < 0.0) {
                t = 0.0;
                s = dmin(dmax(-c / a, 0.0), 1.0);
            } else if (t > 1.0) {
                t = 1.0;
                s = dmin(dmax((b - c) / a, 0.0), 1.0);
            }
        }
    }
    const bool si = s > 0.0 && s < 1.0, ti = t > 0.0 && t < 1.0;
    if (si && ti) return EE_LINES;
    if (si) return t == 0.0 ? EE_B0_A : EE_B1_A;
    if (ti) return s == 0.0 ? EE_A0_B : EE_A1_B;
    return s == 0.0 ? (t == 0.0 ? EE_A0B0 : EE_A0B1) : (t == 0.0 ? EE_A1B0 : EE_A1B1);
}

// the squared distance of the type, T = double or H12
template <typename T>
ND_DU T pair_d2(int kind, int type, const P3<T>* v) {
    if (kind == 0) {
        switch (type) {
            case PT_T0: return d2_point_point(v[0], v[1]);
            case PT_T1: return d2_point_point(v[0], v[2]);
            case PT_T2: return d2_point_point(v[0], v[3]);
            case PT_E01: return d2_point_edge(v[0], v[1], v[2]);
            case PT_E12: return d2_point_edge(v[0], v[2], v[3]);
            case PT_E20: return d2_point_edge(v[0], v[3], v[1]);
            default: return d2_point_plane(v[0], v[1], v[2], v[3]);
        }
    }
    switch (type) {
        case EE_A0B0: return d2_point_point(v[0], v[2]);
        case EE_A0B1: return d2_point_point(v[0], v[3]);
        case EE_A1B0: return d2_point_point(v[1], v[2]);
        case EE_A1B1: return d2_point_point(v[1], v[3]);
        case EE_A0_B: return d2_point_edge(v[0], v[2], v[3]);
        case EE_A1_B: return d2_point_edge(v[1], v[2], v[3]);
        case EE_B0_A: return d2_point_edge(v[2], v[0], v[1]);
        case EE_B1_A: return d2_point_edge(v[3], v[0], v[1]);
        default: return d2_line_line(v[0], v[1], v[2], v[3]);
    }
}

ND_DU int pair_type(int kind, const double* x) { return kind == 0 ? pt_type(x) : ee_type(x); }

// distance of the pair (its type's closed form)
ND_DU double pair_distance(int kind, const double* x, int* type_out) {
    const int type = pair_type(kind, x);
    P3<double> v[4];
    for (int k = 0; k < 4; ++k) v[k] = p3(x + 3 * k);
    const double d2 = pair_d2<double>(kind, type, v);
    if (type_out) *type_out = type;
    return sqrt(d2 > 0.0 ? d2 : 0.0);
}

// Conservative advancement of one pair (Mirtich 1996; the additive form of Li, Kaufman
// and Jiang 2021) along t in [0, 1]: world(t, xw) writes the pair's four world points at
// t; mu bounds how fast the distance can shrink (|d'(t)| <= mu); the pair keeps
// sep * d(0) + margin. From d = d(0), t advances by 0.9 (d - keep) / mu, so d stays
// above keep up to the returned t. Returns that t once the gap d - keep is down to 1e-6
// of its start (or after max_iter advances), +inf when t passes 1 first or nothing
// moves. done(t) may end the advance early (returning t, a lower bound of the full
// result) when the caller knows the pair cannot give the least result any more.
template <class F, class G>
ND_DU double pair_advance(int kind, F&& world, double mu, double sep, double margin, int max_iter, G&& done) {
    double xw[kN];
    world(0.0, xw);
    const double d0 = pair_distance(kind, xw, nullptr);
    const double keep = sep * d0 + margin;
    if (!(mu > 0.0)) return __builtin_inf();
    const double stop = 1e-6 * (d0 - keep);
    double t = 0.0, d = d0;
    for (int k = 0; k < max_iter; ++k) {
        const double gap = d - keep;
        if (!(gap > stop) || !(gap > 0.0) || done(t)) return t;
        t = t + 0.9 * gap / mu;
        if (!(t < 1.0)) return __builtin_inf();
        world(t, xw);
        d = pair_distance(kind, xw, nullptr);
    }
    return t;
}

// E = kappa b(d) of the pair with gradient g[12] and Hessian H[144] (row-major) in x.
// Returns false (E = 0, g = H = 0) when d >= dhat; E = +inf when d <= 0.
ND_DU bool pair_energy(int kind, const double* x, double dhat, double kappa, double* E, double* g, double* H,
                       int* type_out, double* d_out) {
    const int type = pair_type(kind, x);
    if (type_out) *type_out = type;
    P3<H12> v[4];
    for (int k = 0; k < 4; ++k)
        v[k] = {hvar(x[3 * k], 3 * k), hvar(x[3 * k + 1], 3 * k + 1), hvar(x[3 * k + 2], 3 * k + 2)};
    const H12 d2 = pair_d2<H12>(kind, type, v);
    const double d = sqrt(d2.v > 0.0 ? d2.v : 0.0);
    if (d_out) *d_out = d;
    for (int k = 0; k < kN; ++k) g[k] = 0.0;
    for (int k = 0; k < kN * kN; ++k) H[k] = 0.0;
    if (!(d < dhat)) {
        *E = 0.0;
        return false;
    }
    *E = bar_v(d, dhat, kappa);
    if (!(d > 0.0)) return true;
    const double b1 = bar_g(d, dhat, kappa), b2 = bar_h(d, dhat, kappa);
    const double e1 = b1 / (2.0 * d), e2 = (b2 - b1 / d) / (4.0 * d * d);
    for (int i = 0; i < kN; ++i) g[i] = e1 * d2.d[i];
    for (int i = 0; i < kN; ++i)
        for (int j = i; j < kN; ++j) {
            const double hv = e2 * d2.d[i] * d2.d[j] + e1 * d2.h[hix(i, j)];
            H[i * kN + j] = hv;
            H[j * kN + i] = hv;
        }
    return true;
}

// Projects the symmetric H (n x n, row-major) onto the positive semidefinite cone: cyclic
// Jacobi rotations until the off-diagonal mass is below 1e-30 of the total (at most 50
// sweeps), eigenvalues below 0 set to 0, H = V diag(max(l, 0)) V^T. Returns the smallest
// eigenvalue before the projection.
template <int NN>
ND_DU double psd_project(double* H) {
    double A[NN * NN], V[NN * NN];
    double tot = 0.0;
    for (int i = 0; i < NN * NN; ++i) {
        A[i] = H[i];
        V[i] = (i / NN == i % NN) ? 1.0 : 0.0;
        tot += H[i] * H[i];
    }
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < NN; ++i)
            for (int j = i + 1; j < NN; ++j) off += A[i * NN + j] * A[i * NN + j];
        if (!(off > 1e-30 * tot)) break;
        for (int p = 0; p < NN; ++p)
            for (int q = p + 1; q < NN; ++q) {
                const double apq = A[p * NN + q];
                if (apq == 0.0) continue;
                const double app = A[p * NN + p], aqq = A[q * NN + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < NN; ++k) {  // A = J^T A J on rows/columns p, q
                    const double akp = A[k * NN + p], akq = A[k * NN + q];
                    A[k * NN + p] = c * akp - s * akq;
                    A[k * NN + q] = s * akp + c * akq;
                }
                for (int k = 0; k < NN; ++k) {
                    const double apk = A[p * NN + k], aqk = A[q * NN + k];
                    A[p * NN + k] = c * apk - s * aqk;
                    A[q * NN + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < NN; ++k) {
                    const double vkp = V[k * NN + p], vkq = V[k * NN + q];
                    V[k * NN + p] = c * vkp - s * vkq;
                    V[k * NN + q] = s * vkp + c * vkq;
                }
            }
    }
    double lmin = A[0];
    for (int i = 1; i < NN; ++i) lmin = dmin(lmin, A[i * NN + i]);
    for (int i = 0; i < NN; ++i)
        for (int j = 0; j < NN; ++j) {
            double s = 0.0;
            for (int k = 0; k < NN; ++k) {
                const double l = A[k * NN + k];
                if (l > 0.0) s += V[i * NN + k] * l * V[j * NN + k];
            }
            H[i * NN + j] = s;
        }
    return lmin;
}

// gravity (0, -9.81, 0) (IPCSolver.cpp:142)
ND_DU double gravity(int k) { return k == 1 ? -9.81 : 0.0; }

// A moving body's inertia terms of the ipc-mode step at its Newton variable xb (p,
// theta), x^ = xh (InertiaForm.cpp:14-53, body-frame inertia I, row-major): energy
// e = 1/2 m |dp|^2 + (I^T dtheta / 2) . dtheta, gradient g = (m dp, I dtheta), Hessian
// block H = diag(m, m, m, I); and its gravity energy -m g.x_p added to *grav.
ND_DU void body_terms(double m, const double* I, const double* xb, const double* xh, double& e, double* g,
                      double* H, double* grav) {
    const double dp0 = xb[0] - xh[0], dp1 = xb[1] - xh[1], dp2 = xb[2] - xh[2];
    const double dt0 = xb[3] - xh[3], dt1 = xb[4] - xh[4], dt2 = xb[5] - xh[5];
    e = 0.5 * m * ((dp0 * dp0 + dp1 * dp1) + dp2 * dp2);
    const double r0 = 0.5 * dt0, r1 = 0.5 * dt1, r2 = 0.5 * dt2;
    const double s0 = r0 * I[0] + r1 * I[3] + r2 * I[6];
    const double s1 = r0 * I[1] + r1 * I[4] + r2 * I[7];
    const double s2 = r0 * I[2] + r1 * I[5] + r2 * I[8];
    e = e + (s0 * dt0 + s1 * dt1 + s2 * dt2);
    g[0] = m * dp0;
    g[1] = m * dp1;
    g[2] = m * dp2;
    g[3] = (I[0] * dt0 + I[1] * dt1) + I[2] * dt2;
    g[4] = (I[3] * dt0 + I[4] * dt1) + I[5] * dt2;
    g[5] = (I[6] * dt0 + I[7] * dt1) + I[8] * dt2;
    for (int k = 0; k < 36; ++k) H[k] = 0.0;
    H[0] = m;
    H[7] = m;
    H[14] = m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) H[6 * (3 + r) + 3 + c] = I[3 * r + c];
    *grav = *grav - m * ((gravity(0) * xb[0] + gravity(1) * xb[1]) + gravity(2) * xb[2]);
}

// World position of body-frame point r of a body at Newton variable (p, theta) with
// step-start orientation q (the narrow phase's pose, IPCSolver.cpp:616-626), and its
// derivative in theta (3 x 3, column k = d x / d theta_k).
ND_DU void vertex_and_jacobian(const double* xb, const double* q, const double* r, double* xw, double* Jth) {
    using D3 = du::Dn<3>;
    const du::V3T<D3> th = du::mk3<D3>(du::var<3>(xb[3], 0), du::var<3>(xb[4], 1), du::var<3>(xb[5], 2));
    const du::QT<D3> qq = du::rot_update(du::liftq<D3>(q), th, true);
    const du::V3T<D3> w = du::rotate(qq, du::lift3<D3>(r[0], r[1], r[2]));
    xw[0] = xb[0] + w.x.v;
    xw[1] = xb[1] + w.y.v;
    xw[2] = xb[2] + w.z.v;
    for (int k = 0; k < 3; ++k) {
        Jth[0 * 3 + k] = w.x.d[k];
        Jth[1 * 3 + k] = w.y.d[k];
        Jth[2 * 3 + k] = w.z.d[k];
    }
}

// Fold of the vertex-space gradient g[12] and Hessian H[144] of a pair into body DOFs:
// owner[v] in {0, 1} names the body (slot 0 or 1) of vertex v, Jth[v] its 3 x 3 rotation
// Jacobian. gb[12] = (slot 0 DOFs, slot 1 DOFs); Hb[144] the 12 x 12 block matrix
// (slot-major), J^T H J.
ND_DU void fold_to_bodies(const double* g, const double* H, const int* owner, const double (*Jth)[9], double* gb,
                          double* Hb) {
    // J: 12 vertex coordinates x 12 body DOFs; row 3v + a: d x_v,a / d (slot DOFs)
    double J[kN * kN];
    for (int i = 0; i < kN * kN; ++i) J[i] = 0.0;
    for (int v = 0; v < 4; ++v) {
        const int o = 6 * owner[v];
        for (int a = 0; a < 3; ++a) {
            J[(3 * v + a) * kN + o + a] = 1.0;
            for (int k = 0; k < 3; ++k) J[(3 * v + a) * kN + o + 3 + k] = Jth[v][a * 3 + k];
        }
    }
    for (int c = 0; c < kN; ++c) {
        double s = 0.0;
        for (int r = 0; r < kN; ++r) s += J[r * kN + c] * g[r];
        gb[c] = s;
    }
    double HJ[kN * kN];
    for (int r = 0; r < kN; ++r)
        for (int c = 0; c < kN; ++c) {
            double s = 0.0;
            for (int k = 0; k < kN; ++k) s += H[r * kN + k] * J[k * kN + c];
            HJ[r * kN + c] = s;
        }
    for (int r = 0; r < kN; ++r)
        for (int c = 0; c < kN; ++c) {
            double s = 0.0;
            for (int k = 0; k < kN; ++k) s += J[k * kN + r] * HJ[k * kN + c];
            Hb[r * kN + c] = s;
        }
}

}  // namespace ipc
}  // namespace nd
