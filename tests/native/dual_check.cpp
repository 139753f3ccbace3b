// Host check of the adjoint's element derivatives (nexdynipc_amd/csrc/dual.h): every
// dual-number Jacobian the device adjoint uses is compared with central finite
// differences of the same function evaluated in plain doubles.
//   - rot_update (quaternion update by a rotation vector), at random angles and at 0;
//   - rotation_vector at a quaternion with zero vector part;
//   - friction_grad_t: Jacobians in the Newton variables, the step-start poses and mu;
//   - friction_grad_g: Jacobians in the contact point, normal and lagged normal force;
//   - contact_geom_t: the narrow phase's contact point and normal (face-vertex, edge-edge
//     with interior and clamped closest points, vertex-vertex) in its vertices and the
//     body positions;
//   - joint_grad_t for a hinge and a fixed joint: Jacobians in the Newton variables and
//     the reference rotations, and the constraint's Jacobian.
// Prints "OK" and exits 0 when every relative error is below its bound.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../nexdynipc_amd/csrc/dual.h"

using namespace nd::du;

static int g_fail = 0;
static double g_worst = 0.0;

static void expect(const char* what, double ad, double fd, double scale, double tol) {
    const double err = std::fabs(ad - fd) / std::max(scale, 1e-300);
    if (err > g_worst) g_worst = err;
    if (!(err <= tol)) {
        ++g_fail;
        if (g_fail < 20) std::printf("FAIL %s: dual %.12e fd %.12e rel %.2e\n", what, ad, fd, err);
    }
}

static double maxabs(const double* a, int n) {
    double m = 0.0;
    for (int i = 0; i < n; ++i) m = std::max(m, std::fabs(a[i]));
    return m;
}

static void unit_quat(std::mt19937_64& rng, double* q) {
    std::normal_distribution<double> N(0.0, 1.0);
    double n = 0.0;
    for (int k = 0; k < 4; ++k) {
        q[k] = N(rng);
        n += q[k] * q[k];
    }
    n = std::sqrt(n);
    for (int k = 0; k < 4; ++k) q[k] /= n;
}

// rot_update: 3 rotation-vector and 4 quaternion inputs, 4 outputs
static void check_rot_update(std::mt19937_64& rng, double scale_th, bool strict) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double q[4], th[3];
    unit_quat(rng, q);
    for (int k = 0; k < 3; ++k) th[k] = scale_th * U(rng);
    typedef Dn<7> D;
    const QT<D> r = rot_update(mkq(var<7>(q[0], 3), var<7>(q[1], 4), var<7>(q[2], 5), var<7>(q[3], 6)),
                               mk3(var<7>(th[0], 0), var<7>(th[1], 1), var<7>(th[2], 2)), strict);
    const D out[4] = {r.w, r.x, r.y, r.z};
    auto f = [&](const double* in, double* o) {
        const QT<double> rr = rot_update(mkq(in[3], in[4], in[5], in[6]), mk3(in[0], in[1], in[2]), strict);
        o[0] = rr.w; o[1] = rr.x; o[2] = rr.y; o[3] = rr.z;
    };
    double in[7] = {th[0], th[1], th[2], q[0], q[1], q[2], q[3]};
    const double h = 1e-6;
    for (int k = 0; k < 7; ++k) {
        double ip[7], im[7], op[4], om[4];
        for (int i = 0; i < 7; ++i) ip[i] = im[i] = in[i];
        ip[k] += h;
        im[k] -= h;
        f(ip, op);
        f(im, om);
        for (int r2 = 0; r2 < 4; ++r2) expect("rot_update", out[r2].d[k], (op[r2] - om[r2]) / (2 * h), 1.0, 1e-8);
    }
}

static void check_rotation_vector_at_identity_axis() {
    // q = (w, 0, 0, 0) with |w| = 1: the derivative of the limit 2 sign(w) v / |w|
    for (double w : {1.0, -1.0}) {
        typedef Dn<4> D;
        const V3T<D> r = rotation_vector(mkq(var<4>(w, 0), var<4>(0.0, 1), var<4>(0.0, 2), var<4>(0.0, 3)));
        const D o[3] = {r.x, r.y, r.z};
        const double h = 1e-7;
        for (int k = 1; k < 4; ++k) {
            double qp[4] = {w, 0, 0, 0}, qm[4] = {w, 0, 0, 0};
            qp[k] += h;
            qm[k] -= h;
            const V3T<double> a = rotation_vector(mkq(qp[0], qp[1], qp[2], qp[3]));
            const V3T<double> b = rotation_vector(mkq(qm[0], qm[1], qm[2], qm[3]));
            const double da[3] = {a.x, a.y, a.z}, db[3] = {b.x, b.y, b.z};
            for (int r2 = 0; r2 < 3; ++r2) expect("rotation_vector", o[r2].d[k], (da[r2] - db[r2]) / (2 * h), 1.0, 1e-8);
        }
    }
}

static void check_friction(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double qA[4], qB[4], pA0[3], pB0[3], xa[6], xb[6], pt[3], nr[3];
    unit_quat(rng, qA);
    unit_quat(rng, qB);
    for (int k = 0; k < 3; ++k) {
        pA0[k] = U(rng);
        pB0[k] = U(rng);
        pt[k] = 0.5 * (pA0[k] + pB0[k]) + 0.1 * U(rng);
        nr[k] = U(rng);
        xa[k] = pA0[k] + 0.01 * U(rng);
        xb[k] = pB0[k] + 0.01 * U(rng);
        xa[3 + k] = 0.02 * U(rng);
        xb[3 + k] = 0.02 * U(rng);
    }
    const double N = 3.0, mu = 0.4, eps = 1e-8, dt = 1e-2;
    // all 27 inputs: xa(6) xb(6) pA0(3) qA0(4) pB0(3) qB0(4) mu(1)
    double in[27];
    for (int k = 0; k < 6; ++k) {
        in[k] = xa[k];
        in[6 + k] = xb[k];
    }
    for (int k = 0; k < 3; ++k) {
        in[12 + k] = pA0[k];
        in[19 + k] = pB0[k];
    }
    for (int k = 0; k < 4; ++k) {
        in[15 + k] = qA[k];
        in[22 + k] = qB[k];
    }
    in[26] = mu;
    auto f = [&](const double* v, double* g) {
        friction_grad_t<double>(v, v + 6, mk3(v[12], v[13], v[14]), mkq(v[15], v[16], v[17], v[18]),
                                mk3(v[19], v[20], v[21]), mkq(v[22], v[23], v[24], v[25]), pt, nr, N, v[26], eps,
                                dt, g);
    };
    typedef Dn<27> D;
    D dv[27];
    for (int k = 0; k < 27; ++k) dv[k] = var<27>(in[k], k);
    D g[12];
    friction_grad_t<D>(dv, dv + 6, mk3(dv[12], dv[13], dv[14]), mkq(dv[15], dv[16], dv[17], dv[18]),
                       mk3(dv[19], dv[20], dv[21]), mkq(dv[22], dv[23], dv[24], dv[25]), pt, nr, N, dv[26], eps, dt,
                       g);
    double g0[12];
    f(in, g0);
    const double sc = maxabs(g0, 12);
    for (int k = 0; k < 27; ++k) {
        const double h = 1e-7;
        double ip[27], im[27], gp[12], gm[12];
        for (int i = 0; i < 27; ++i) ip[i] = im[i] = in[i];
        ip[k] += h;
        im[k] -= h;
        f(ip, gp);
        f(im, gm);
        // scale of a derivative column: |g| over the input's natural length scale
        for (int r2 = 0; r2 < 12; ++r2)
            expect("friction_grad", g[r2].d[k], (gp[r2] - gm[r2]) / (2 * h), sc / 1e-2, 1e-6);
    }
}

// friction_grad_g: Jacobian in the contact point, normal and normal force
static void check_friction_geom(std::mt19937_64& rng) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double qA[4], qB[4], pA0[3], pB0[3], xa[6], xb[6], in[7];
    unit_quat(rng, qA);
    unit_quat(rng, qB);
    for (int k = 0; k < 3; ++k) {
        pA0[k] = U(rng);
        pB0[k] = U(rng);
        in[k] = 0.5 * (pA0[k] + pB0[k]) + 0.1 * U(rng);  // point
        in[3 + k] = U(rng);                              // normal (not unit)
        xa[k] = pA0[k] + 0.01 * U(rng);
        xb[k] = pB0[k] + 0.01 * U(rng);
        xa[3 + k] = 0.02 * U(rng);
        xb[3 + k] = 0.02 * U(rng);
    }
    in[6] = 3.0;  // N
    const double mu = 0.4, eps = 1e-8, dt = 1e-2;
    auto f = [&](const double* v, double* g) {
        friction_grad_g<double>(xa, xb, mk3(pA0[0], pA0[1], pA0[2]), mkq(qA[0], qA[1], qA[2], qA[3]),
                                mk3(pB0[0], pB0[1], pB0[2]), mkq(qB[0], qB[1], qB[2], qB[3]), mk3(v[0], v[1], v[2]),
                                mk3(v[3], v[4], v[5]), v[6], mu, eps, dt, g);
    };
    typedef Dn<7> D;
    D dv[7], xad[6], xbd[6];
    for (int k = 0; k < 7; ++k) dv[k] = var<7>(in[k], k);
    for (int k = 0; k < 6; ++k) {
        xad[k] = cst<7>(xa[k]);
        xbd[k] = cst<7>(xb[k]);
    }
    D g[12];
    friction_grad_g<D>(xad, xbd, lift3<D>(pA0[0], pA0[1], pA0[2]), liftq<D>(qA), lift3<D>(pB0[0], pB0[1], pB0[2]),
                       liftq<D>(qB), mk3(dv[0], dv[1], dv[2]), mk3(dv[3], dv[4], dv[5]), dv[6], cst<7>(mu), eps, dt,
                       g);
    double g0[12];
    f(in, g0);
    const double sc = maxabs(g0, 12);
    for (int k = 0; k < 7; ++k) {
        const double h = 1e-7;
        double ip[7], im[7], gp[12], gm[12];
        for (int i = 0; i < 7; ++i) ip[i] = im[i] = in[i];
        ip[k] += h;
        im[k] -= h;
        f(ip, gp);
        f(im, gm);
        for (int r2 = 0; r2 < 12; ++r2)
            expect("friction_grad_geom", g[r2].d[k], (gp[r2] - gm[r2]) / (2 * h), sc, 1e-6);
    }
}

// contact_geom_t: point and normal of a face-vertex, edge-edge (interior and clamped
// closest points) and vertex-vertex contact, in its four vertices and two body positions
static void check_contact_geom(std::mt19937_64& rng, int kind, bool clamp) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double in[18];
    for (int k = 0; k < 18; ++k) in[k] = U(rng);
    if (kind == 0) {  // a vertex above the triangle's interior
        for (int k = 0; k < 3; ++k) in[9 + k] = (in[k] + in[3 + k] + in[6 + k]) / 3.0 + 0.2 * U(rng);
    }
    if (kind == 1 && clamp) {  // the second segment far beyond the first's end
        for (int k = 0; k < 3; ++k) {
            in[6 + k] += 3.0 * (in[3 + k] - in[k]);
            in[9 + k] += 3.0 * (in[3 + k] - in[k]);
        }
    }
    auto f = [&](const double* v, double* o) {
        V3T<double> vv[4];
        for (int i = 0; i < 4; ++i) vv[i] = mk3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        V3T<double> pt, nr;
        contact_geom_t<double>(kind, vv, mk3(v[12], v[13], v[14]), mk3(v[15], v[16], v[17]), pt, nr);
        o[0] = pt.x; o[1] = pt.y; o[2] = pt.z; o[3] = nr.x; o[4] = nr.y; o[5] = nr.z;
    };
    typedef Dn<18> D;
    D dv[18];
    for (int k = 0; k < 18; ++k) dv[k] = var<18>(in[k], k);
    V3T<D> vv[4];
    for (int i = 0; i < 4; ++i) vv[i] = mk3(dv[3 * i], dv[3 * i + 1], dv[3 * i + 2]);
    V3T<D> pt, nr;
    contact_geom_t<D>(kind, vv, mk3(dv[12], dv[13], dv[14]), mk3(dv[15], dv[16], dv[17]), pt, nr);
    const D o[6] = {pt.x, pt.y, pt.z, nr.x, nr.y, nr.z};
    double o0[6];
    f(in, o0);
    for (int r2 = 0; r2 < 6; ++r2) expect("contact_geom value", o[r2].v, o0[r2], 1.0, 0.0);
    for (int k = 0; k < 18; ++k) {
        const double h = 1e-7;
        double ip[18], im[18], op[6], om[6];
        for (int i = 0; i < 18; ++i) ip[i] = im[i] = in[i];
        ip[k] += h;
        im[k] -= h;
        f(ip, op);
        f(im, om);
        for (int r2 = 0; r2 < 6; ++r2) expect("contact_geom", o[r2].d[k], (op[r2] - om[r2]) / (2 * h), 1.0, 1e-6);
    }
}

static void check_joint(std::mt19937_64& rng, int type) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double qA[4], qB[4], jd[16] = {0}, xa[6], xb[6], lam[6];
    unit_quat(rng, qA);
    unit_quat(rng, qB);
    for (int k = 0; k < 16; ++k) jd[k] = U(rng);
    if (type == 0) {
        double tq[4];
        unit_quat(rng, tq);
        for (int k = 0; k < 4; ++k) jd[3 + k] = tq[k];
    } else {
        for (int a = 6; a <= 9; a += 3) {  // unit axes
            const double n = std::sqrt(jd[a] * jd[a] + jd[a + 1] * jd[a + 1] + jd[a + 2] * jd[a + 2]);
            for (int k = 0; k < 3; ++k) jd[a + k] /= n;
        }
    }
    for (int k = 0; k < 6; ++k) {
        xa[k] = 0.3 * U(rng);
        xb[k] = 0.3 * U(rng);
        lam[k] = U(rng);
    }
    const double mu = 50.0;
    // inputs: xa(6) xb(6) qA(4) qB(4)
    double in[20];
    for (int k = 0; k < 6; ++k) {
        in[k] = xa[k];
        in[6 + k] = xb[k];
    }
    for (int k = 0; k < 4; ++k) {
        in[12 + k] = qA[k];
        in[16 + k] = qB[k];
    }
    auto f = [&](const double* v, double* g, double* C) {
        joint_grad_t<double>(type, jd, v, v + 6, mkq(v[12], v[13], v[14], v[15]), mkq(v[16], v[17], v[18], v[19]),
                             mu, lam, g, C);
    };
    typedef Dn<20> D;
    D dv[20];
    for (int k = 0; k < 20; ++k) dv[k] = var<20>(in[k], k);
    D g[12], C[6];
    joint_grad_t<D>(type, jd, dv, dv + 6, mkq(dv[12], dv[13], dv[14], dv[15]), mkq(dv[16], dv[17], dv[18], dv[19]),
                    mu, lam, g, C);
    double g0[12], C0[6];
    f(in, g0, C0);
    const double sg = maxabs(g0, 12);
    for (int k = 0; k < 20; ++k) {
        const double h = 1e-7;
        double ip[20], im[20], gp[12], gm[12], Cp[6], Cm[6];
        for (int i = 0; i < 20; ++i) ip[i] = im[i] = in[i];
        ip[k] += h;
        im[k] -= h;
        f(ip, gp, Cp);
        f(im, gm, Cm);
        for (int r2 = 0; r2 < 12; ++r2) expect(type ? "hinge_grad" : "fixed_grad", g[r2].d[k], (gp[r2] - gm[r2]) / (2 * h), sg, 1e-6);
        for (int r2 = 0; r2 < 6; ++r2) expect(type ? "hinge_C" : "fixed_C", C[r2].d[k], (Cp[r2] - Cm[r2]) / (2 * h), 1.0, 1e-7);
    }
}

int main() {
    std::mt19937_64 rng(20261203);
    for (int t = 0; t < 50; ++t) {
        check_rot_update(rng, 0.5, true);
        check_rot_update(rng, 0.5, false);
    }
    for (int t = 0; t < 10; ++t) {
        check_rot_update(rng, 0.0, true);  // th = 0: the smooth limit
        check_rot_update(rng, 0.0, false);
    }
    check_rotation_vector_at_identity_axis();
    for (int t = 0; t < 50; ++t) check_friction(rng);
    for (int t = 0; t < 50; ++t) check_friction_geom(rng);
    for (int t = 0; t < 50; ++t) {
        check_contact_geom(rng, 0, false);
        check_contact_geom(rng, 1, false);
        check_contact_geom(rng, 1, true);
        check_contact_geom(rng, 2, false);
    }
    for (int t = 0; t < 50; ++t) {
        check_joint(rng, 0);
        check_joint(rng, 1);
    }
    if (g_fail) {
        std::printf("FAILED %d checks (worst relative error %.2e)\n", g_fail, g_worst);
        return 1;
    }
    std::printf("OK worst relative error %.2e\n", g_worst);
    return 0;
}
