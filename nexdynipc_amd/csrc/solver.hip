// Newton / ALM solver kernels: one work-group per scene runs
// NewtonSolver::minimize (NewtonSolver.cpp:10-95) with the backtracking line search
// (LineSearch.cpp:6-58) on the SimulationProblem objective (IPCSolver.cpp:137-262):
// inertia (body-frame, unrotated), h^2 gravity, ALM joints, h^2 barrier and h^2 lagged
// friction. The linear solve replaces SimplicialLDLT with an LDL^T per connected
// component of the coupling graph (block-Jacobi PCG for components above 64 bodies).
// Scenes are independent, so a batch of scenes is one launch of each kernel.
#include "ndipc_internal.h"
#include "dual.h"

namespace nd {

constexpr int kBlk = 256;
constexpr int kTermLds = 8192;  // doubles of LDS for the ordered sums of the Newton kernel
constexpr double kG = -9.81;  // gravity (0, -9.81, 0), IPCSolver.cpp:142

// ------------------------------------------------------------------ potentials
// BarrierPotential (Barrier.cpp:11-46): bar_v / bar_g / bar_h in dmath.h

struct ScnCtx {
    int s, b0, n, nd, j0, j1, nc;
    long long cb;  // contact base (raw contact index)
    double h2, dhat, kappa, fmu, feps, fdt;
    int has_barrier, has_friction, nfc;
};

__device__ __forceinline__ V3 xpos(const double* x, int b) { return v3(x[6 * b], x[6 * b + 1], x[6 * b + 2]); }
__device__ __forceinline__ V3 xrot(const double* x, int b) { return v3(x[6 * b + 3], x[6 * b + 4], x[6 * b + 5]); }

// DistanceBarrierForm::contactDistance (DistanceBarrierForm.cpp:89-98); xa, xb: the
// Newton variables of the contact's two bodies.
__device__ __forceinline__ double contact_dist_x(const Work& w, long long c, const double* xa, const double* xb) {
    const double* cd = w.cdbl + 7 * c;
    const V3 n = v3(cd[3], cd[4], cd[5]);
    const double pr = dot(v3(xb[0], xb[1], xb[2]) - v3(xa[0], xa[1], xa[2]), n);
    return pr > 1e-9 ? pr : 1e-9;
}
__device__ __forceinline__ double contact_dist(const Work& w, long long c, const double* X) {
    const int A = w.cint[5 * c], B = w.cint[5 * c + 1];
    return contact_dist_x(w, c, X + 6 * A, X + 6 * B);
}

// FrictionForm::computeTangentDisplacement + Jacobians (FrictionForm.cpp:44-204)
struct FKin {
    V3 u;
    double JA[3][6], JB[3][6];
};
__device__ void fric_kin(const Dev& d, const Work& w, long long c, const double* X, double dt, FKin& k) {
    const int A = w.cint[5 * c], B = w.cint[5 * c + 1];
    const double* cd = w.cdbl + 7 * c;
    const V3 pt = v3(cd[0], cd[1], cd[2]);
    const V3 nr = v3(cd[3], cd[4], cd[5]);
    const V3 pA0 = ld3(d.p, A), pB0 = ld3(d.p, B);
    const Q qA0 = ld4(d.q, A), qB0 = ld4(d.q, B);
    const V3 pA = xpos(X, A), pB = xpos(X, B);
    const Q qA = apply_theta(qA0, xrot(X, A)), qB = apply_theta(qB0, xrot(X, B));
    const V3 dA = pA - pA0, dB = pB - pB0;
    const V3 lA = rotate(qconj(qA0), pt - pA0), lB = rotate(qconj(qB0), pt - pB0);
    const V3 rA = rotate(qA, lA), rB = rotate(qB, lB);
    V3 wA = v3(0, 0, 0), wB = v3(0, 0, 0), vA = v3(0, 0, 0), vB = v3(0, 0, 0);
    if (dt > 1e-12) {
        wA = xrot(X, A) / dt;
        wB = xrot(X, B) / dt;
        vA = dA / dt;
        vB = dB / dt;
    }
    vA = vA + cross(wA, rA);
    vB = vB + cross(wB, rB);
    const V3 rel = vB - vA;
    const V3 nn = (sqnorm(nr) > 1e-16) ? normalized(nr) : v3(0.0, 1.0, 0.0);
    const V3 tv = rel - dot(rel, nn) * nn;
    k.u = tv * dt;
    M3 P;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P.m[3 * i + j] = (i == j ? 1.0 : 0.0) - comp(nn, i) * comp(nn, j);
    const M3 PA = m3mul(P, skew(rA)), PB = m3mul(P, skew(rB));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            k.JA[i][j] = -P.m[3 * i + j];
            k.JB[i][j] = P.m[3 * i + j];
            k.JA[i][3 + j] = PA.m[3 * i + j];
            k.JB[i][3 + j] = -PB.m[3 * i + j];
        }
}
// The parts of the friction displacement that do not depend on the Newton variables, computed once per
// contact: the step-start poses, the contact point in both body frames and the unit
// normal (the same expressions as in fric_kin).
struct FricPre {
    V3 pA0, pB0, lA, lB, nn;
    Q qA0, qB0;
};
__device__ __forceinline__ FricPre fric_pre(const Dev& d, const Work& w, long long c) {
    const int A = w.cint[5 * c], B = w.cint[5 * c + 1];
    const double* cd = w.cdbl + 7 * c;
    const V3 pt = v3(cd[0], cd[1], cd[2]);
    const V3 nr = v3(cd[3], cd[4], cd[5]);
    FricPre f;
    f.pA0 = ld3(d.p, A);
    f.pB0 = ld3(d.p, B);
    f.qA0 = ld4(d.q, A);
    f.qB0 = ld4(d.q, B);
    f.lA = rotate(qconj(f.qA0), pt - f.pA0);
    f.lB = rotate(qconj(f.qB0), pt - f.pB0);
    f.nn = (sqnorm(nr) > 1e-16) ? normalized(nr) : v3(0.0, 1.0, 0.0);
    return f;
}
// fric_kin's tangential displacement u alone (the friction value needs no Jacobian),
// from the Newton variables xa, xb of the two bodies: the same expressions.
__device__ __forceinline__ V3 fric_u_pre(const FricPre& f, const double* xa, const double* xb, double dt) {
    const V3 pA = v3(xa[0], xa[1], xa[2]), pB = v3(xb[0], xb[1], xb[2]);
    const V3 thA = v3(xa[3], xa[4], xa[5]), thB = v3(xb[3], xb[4], xb[5]);
    const Q qA = apply_theta(f.qA0, thA), qB = apply_theta(f.qB0, thB);
    const V3 dA = pA - f.pA0, dB = pB - f.pB0;
    const V3 rA = rotate(qA, f.lA), rB = rotate(qB, f.lB);
    V3 wA = v3(0, 0, 0), wB = v3(0, 0, 0), vA = v3(0, 0, 0), vB = v3(0, 0, 0);
    if (dt > 1e-12) {
        wA = thA / dt;
        wB = thB / dt;
        vA = dA / dt;
        vB = dB / dt;
    }
    vA = vA + cross(wA, rA);
    vB = vB + cross(wB, rB);
    const V3 rel = vB - vA;
    const V3 tv = rel - dot(rel, f.nn) * f.nn;
    return tv * dt;
}
// FrictionPotential (Friction.cpp:5-62)
__device__ __forceinline__ double fric_v(V3 u, double N, double mu, double eps) {
    if (N <= 0) return 0.0;
    const double un = norm(u);
    const double sn = sqrt(un * un + eps * eps);
    return mu * N * (sn - eps);
}
__device__ __forceinline__ V3 fric_g(V3 u, double N, double mu, double eps) {
    if (N <= 0) return v3(0, 0, 0);
    const double un = norm(u);
    if (un < 1e-12) return v3(0, 0, 0);
    const double sn = sqrt(un * un + eps * eps);
    const double scale = mu * N / sn;
    return scale * u;
}
__device__ __forceinline__ void fric_h(V3 u, double N, double mu, double eps, double H[9]) {
    for (int i = 0; i < 9; ++i) H[i] = 0.0;
    if (N <= 0) return;
    const double un = norm(u);
    if (un < 1e-12) {
        const double s = mu * N / eps;
        H[0] = s * 1.0; H[4] = s * 1.0; H[8] = s * 1.0;
        H[1] = H[2] = H[3] = H[5] = H[6] = H[7] = s * 0.0;
        return;
    }
    const double sn = sqrt(un * un + eps * eps);
    const double scale = mu * N / sn;
    const double s2 = sn * sn;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) H[3 * i + j] = scale * ((i == j ? 1.0 : 0.0) - (comp(u, i) * comp(u, j)) / s2);
}

// ------------------------------------------------------------------ joints
// FixedJoint / HingeJoint constraint and Jacobian (FixedJoint.cpp:141-173,
// HingeJoint.cpp:36-69, JointPrimitives.h:63-108,208-245,281-324). J is 6 x 12
// (columns: body A dofs, body B dofs).
struct JEval {
    double C[6];
    double J[6][12];
};
__device__ void joint_eval(const Dev& d, int j, const double* X, JEval& e) {
    const int type = d.jtype[j];
    const int A = d.ja[j], B = d.jb[j];
    const double* jd = d.jdata + 16 * j;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 12; ++c) e.J[r][c] = 0.0;
    if (type == 0) {
        const V3 tp = v3(jd[0], jd[1], jd[2]);
        const Q tq = quat(jd[3], jd[4], jd[5], jd[6]);
        const V3 dp = xpos(X, A) - tp;
        const V3 tt = rotation_vector(qmul(tq, qconj(ld4(d.q, A))));
        const V3 e3 = xrot(X, A) - tt;
        e.C[0] = dp.x; e.C[1] = dp.y; e.C[2] = dp.z; e.C[3] = e3.x; e.C[4] = e3.y; e.C[5] = e3.z;
        for (int r = 0; r < 6; ++r) e.J[r][r] = 1.0;
        return;
    }
    const V3 anA = v3(jd[0], jd[1], jd[2]), anB = v3(jd[3], jd[4], jd[5]);
    const V3 axA = v3(jd[6], jd[7], jd[8]), axB = v3(jd[9], jd[10], jd[11]);
    const V3 pA = xpos(X, A), pB = xpos(X, B);
    const Q qA = update_rotation(ld4(d.q, A), xrot(X, A));
    const Q qB = update_rotation(ld4(d.q, B), xrot(X, B));
    const V3 rA = rotate(qA, anA), rB = rotate(qB, anB);
    const V3 pv = (pA + rA) - (pB + rB);
    const V3 nA = rotate(qA, axA), nB = rotate(qB, axB);
    const V3 ac = cross(nA, nB);
    e.C[0] = pv.x; e.C[1] = pv.y; e.C[2] = pv.z; e.C[3] = ac.x; e.C[4] = ac.y; e.C[5] = ac.z;
    const M3 sA = skew(rA), sB = skew(rB);
    for (int i = 0; i < 3; ++i) {
        e.J[i][i] = 1.0;
        e.J[i][6 + i] = -1.0;
        for (int k = 0; k < 3; ++k) {
            e.J[i][3 + k] = -sA.m[3 * i + k];
            e.J[i][9 + k] = sB.m[3 * i + k];
        }
    }
    const M3 dA = m3mul(skew(nB), skew(nA));
    M3 mnA = skew(nA);
    for (int k = 0; k < 9; ++k) mnA.m[k] = -mnA.m[k];
    const M3 dB = m3mul(mnA, skew(nB));
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            e.J[3 + i][3 + k] = dA.m[3 * i + k];
            e.J[3 + i][9 + k] = dB.m[3 * i + k];
        }
}

// The constraint alone (the value and the ALM updates need no Jacobian): the same
// expressions as joint_eval's C. xa, xb: the Newton variables of bodies A and B (xb
// unused by a fixed joint).
__device__ void joint_C_x(const Dev& d, int j, const double* xa, const double* xb, double* C) {
    const int type = d.jtype[j];
    const int A = d.ja[j], B = d.jb[j];
    const double* jd = d.jdata + 16 * j;
    if (type == 0) {
        const V3 tp = v3(jd[0], jd[1], jd[2]);
        const Q tq = quat(jd[3], jd[4], jd[5], jd[6]);
        const V3 dp = v3(xa[0], xa[1], xa[2]) - tp;
        const V3 tt = rotation_vector(qmul(tq, qconj(ld4(d.q, A))));
        const V3 e3 = v3(xa[3], xa[4], xa[5]) - tt;
        C[0] = dp.x; C[1] = dp.y; C[2] = dp.z; C[3] = e3.x; C[4] = e3.y; C[5] = e3.z;
        return;
    }
    const V3 anA = v3(jd[0], jd[1], jd[2]), anB = v3(jd[3], jd[4], jd[5]);
    const V3 axA = v3(jd[6], jd[7], jd[8]), axB = v3(jd[9], jd[10], jd[11]);
    const V3 pA = v3(xa[0], xa[1], xa[2]), pB = v3(xb[0], xb[1], xb[2]);
    const Q qA = update_rotation(ld4(d.q, A), v3(xa[3], xa[4], xa[5]));
    const Q qB = update_rotation(ld4(d.q, B), v3(xb[3], xb[4], xb[5]));
    const V3 rA = rotate(qA, anA), rB = rotate(qB, anB);
    const V3 pv = (pA + rA) - (pB + rB);
    const V3 nA = rotate(qA, axA), nB = rotate(qB, axB);
    const V3 ac = cross(nA, nB);
    C[0] = pv.x; C[1] = pv.y; C[2] = pv.z; C[3] = ac.x; C[4] = ac.y; C[5] = ac.z;
}
__device__ void joint_C(const Dev& d, int j, const double* X, double* C) {
    const int A = d.ja[j], B = d.jb[j];
    joint_C_x(d, j, X + 6 * A, B >= 0 ? X + 6 * B : nullptr, C);
}

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ void load_scn(const Dev& d, int s, ScnCtx& c) {
    const SceneParams& sp = d.sp[s];
    c.s = s;
    c.b0 = d.body_off[s];
    c.n = d.body_off[s + 1] - c.b0;
    c.nd = 6 * c.n;
    c.j0 = d.joint_off[s];
    c.j1 = d.joint_off[s + 1];
    c.nc = sp.n_contacts;
    c.cb = (long long)s * kMaxRawContacts;
    c.h2 = sp.adt * sp.adt;
    c.dhat = sp.dhat > 1e-9 ? sp.dhat : 1e-9;
    c.kappa = sp.kappa > 0.0 ? sp.kappa : 0.0;
    c.fmu = sp.f_mu;
    c.feps = 1e-8;
    c.fdt = sp.adt > 0.0 ? sp.adt : 0.0;
    c.has_barrier = sp.has_barrier;
    c.has_friction = sp.has_friction;
    c.nfc = sp.nf_count;
}

// Per-contact element data at X: barrier b'' (unscaled), gradients, friction blocks.
// cgrad[c][18] = gbA(3) gbB(3) gfA(6) gfB(6);  cblk[c][4*36] = h2 * friction AA BB AB BA;
// cactive[c] bit0 barrier active, bit1 friction active; the unscaled barrier b'' of
// contact c is written to b2arr[c]. Without kBlocks the friction blocks are left
// unwritten (a caller that needs only the gradients).
template <bool kBlocks = true>
__device__ void eval_elements(const Dev& d, const Work& w, const ScnCtx& sc, const double* X, double* b2arr) {
    for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
        const long long c = sc.cb + i;
        double* g = w.cgrad + 18 * ((long long)sc.s * kMaxContacts + i);
        for (int k = 0; k < 18; ++k) g[k] = 0.0;
        int act = 0;
        const double* cd = w.cdbl + 7 * c;
        const V3 n = v3(cd[3], cd[4], cd[5]);
        b2arr[i] = 0.0;
        if (sc.has_barrier) {
            const double dd = contact_dist(w, c, X);
            if (dd < sc.dhat) {
                const double b1 = bar_g(dd, sc.dhat, sc.kappa);
                g[0] = -b1 * n.x; g[1] = -b1 * n.y; g[2] = -b1 * n.z;
                g[3] = b1 * n.x; g[4] = b1 * n.y; g[5] = b1 * n.z;
                b2arr[i] = bar_h(dd, sc.dhat, sc.kappa);
                act |= 1;
            }
        }
        if (sc.has_friction && i < sc.nfc && w.nf[(long long)sc.s * kMaxContacts + i] > 0.0) {
            const double N = w.nf[(long long)sc.s * kMaxContacts + i];
            FKin k;
            fric_kin(d, w, c, X, sc.fdt, k);
            const V3 fg = fric_g(k.u, N, sc.fmu, sc.feps);
            for (int j = 0; j < 6; ++j) {
                g[6 + j] = k.JA[0][j] * fg.x + k.JA[1][j] * fg.y + k.JA[2][j] * fg.z;
                g[12 + j] = k.JB[0][j] * fg.x + k.JB[1][j] * fg.y + k.JB[2][j] * fg.z;
            }
            double H[9];
            if (kBlocks) fric_h(k.u, N, sc.fmu, sc.feps, H);
            double* blk = w.cblk + 144 * ((long long)sc.s * kMaxContacts + i);
            // (X^T H) Y for (A,A), (B,B), (A,B), (B,A)   (FrictionForm.cpp:137-140)
            for (int which = 0; which < (kBlocks ? 4 : 0); ++which) {
                double (*Xm)[6] = (which == 1 || which == 3) ? k.JB : k.JA;
                double (*Ym)[6] = (which == 0 || which == 3) ? k.JA : k.JB;
                double T[6][3];
                for (int r = 0; r < 6; ++r)
                    for (int cc = 0; cc < 3; ++cc)
                        T[r][cc] = Xm[0][r] * H[0 * 3 + cc] + Xm[1][r] * H[1 * 3 + cc] + Xm[2][r] * H[2 * 3 + cc];
                for (int r = 0; r < 6; ++r)
                    for (int cc = 0; cc < 6; ++cc)
                        blk[which * 36 + r * 6 + cc] = (T[r][0] * Ym[0][cc] + T[r][1] * Ym[1][cc] + T[r][2] * Ym[2][cc]) * sc.h2;
            }
            act |= 2;
        }
        w.cactive[(long long)sc.s * kMaxContacts + i] = act;
    }
    // hinge joint element blocks (mu J^T J) and gradients J^T (mu C + lambda); joints
    // are taken from the last thread down, so that with fewer contacts than threads a
    // joint is evaluated beside the contacts rather than after one
    for (int j = sc.j0 + (blockDim.x - 1 - threadIdx.x); j < sc.j1; j += blockDim.x) {
        JEval e;
        joint_eval(d, j, X, e);
        const double mu = d.jmu[j];
        double f[6];
        for (int r = 0; r < 6; ++r) f[r] = mu * e.C[r] + d.jlambda[6 * j + r];
        double* jg = w.jgrad + 12 * j;
        for (int c = 0; c < 12; ++c) {
            double acc = 0.0;
            for (int r = 0; r < 6; ++r) acc = acc + e.J[r][c] * f[r];
            jg[c] = acc;
        }
        double* jb = w.jblk + 108 * j;
        for (int which = 0; which < 3; ++which) {
            const int oa = (which == 1) ? 6 : 0, ob = (which == 0) ? 0 : 6;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    double acc = 0.0;
                    for (int k = 0; k < 6; ++k) acc = acc + e.J[k][oa + r] * e.J[k][ob + c];
                    jb[which * 36 + r * 6 + c] = acc * mu;
                }
        }
        for (int r = 0; r < 6; ++r) d.jC[6 * j + r] = e.C[r];
    }
    __syncthreads();
}

// Gradient of the Newton objective (IPCSolver.cpp:167-190) and the diagonal 6x6 block
// of the Hessian of body li, gathered through the body CSR.
//   g: inertia, h^2 gravity, joints in joint order, then h^2 barrier and h^2 friction
//      summed over the body's contact entries in contact order; cgv: the contact part
//      alone.
//   D: as body_diag (inertia, joints, barrier contacts, friction contacts).
// A contact entry's data (its gradient rows, active flags, normal and b'') are loaded
// together, with the next entry's index fetched beside them, so each entry waits out
// one memory latency rather than a chain of dependent loads; the friction blocks
// follow in a second pass over the entries found friction-active (a mask for the
// first 64 entries, a plain scan beyond). Every accumulator adds its terms in the same
// order as body_diag and the gradient sums.
__device__ __forceinline__ void body_grad_diag(const Dev& d, const Work& w, const ScnCtx& sc, const double* X,
                                               const double* b2arr, int li, const int* off, const int* ent,
                                               double g[6], double cgv[6], double D[36]) {
    const double* XH = w.xhat;
    const int b = sc.b0 + li;
    const double m = d.mass[b];
    const double* I = d.inertia + 9 * b;
    const V3 dp = xpos(X, b) - xpos(XH, b);
    const V3 gp = m * dp;
    const V3 dt = xrot(X, b) - xrot(XH, b);
    g[0] = 0.0 + gp.x; g[1] = 0.0 + gp.y; g[2] = 0.0 + gp.z;
    g[3] = 0.0 + (I[0] * dt.x + I[1] * dt.y + I[2] * dt.z);
    g[4] = 0.0 + (I[3] * dt.x + I[4] * dt.y + I[5] * dt.z);
    g[5] = 0.0 + (I[6] * dt.x + I[7] * dt.y + I[8] * dt.z);
    const double fy = (-m) * kG;
    g[0] = g[0] + sc.h2 * ((-m) * 0.0);
    g[1] = g[1] + sc.h2 * fy;
    g[2] = g[2] + sc.h2 * ((-m) * 0.0);
    g[3] = g[3] + sc.h2 * 0.0; g[4] = g[4] + sc.h2 * 0.0; g[5] = g[5] + sc.h2 * 0.0;
#pragma unroll
    for (int k = 0; k < 36; ++k) D[k] = 0.0;
    D[0] = m; D[7] = m; D[14] = m;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) D[(3 + r) * 6 + 3 + c] = I[3 * r + c];
    // joints (ConstraintForm accumulates each joint's J^T f and its blocks)
    double gj[6] = {0, 0, 0, 0, 0, 0};
    for (int e = d.bj_off[b]; e < d.bj_off[b + 1]; ++e) {  // the body's joints, j ascending
        const int je = d.bj_ent[e];
        const double* jg = w.jgrad + 12 * (je >> 1) + 6 * (je & 1);
        for (int k = 0; k < 6; ++k) gj[k] += jg[k];
        const double* jb = w.jblk + 108 * (je >> 1) + 36 * (je & 1);
#pragma unroll
        for (int k = 0; k < 36; ++k) D[k] = D[k] + jb[k];
    }
    for (int k = 0; k < 6; ++k) g[k] = g[k] + gj[k];
    // contacts: gradient rows of every entry, barrier blocks of the barrier-active ones
    const long long sbase = (long long)sc.s * kMaxContacts;
    const int e0 = off[li], e1 = off[li + 1];
    double gb[6] = {0, 0, 0, 0, 0, 0}, gf[6] = {0, 0, 0, 0, 0, 0};
    unsigned long long fmask = 0;  // entry e0 + i (i < 64) has friction active
    int en = e0 < e1 ? ent[e0] : 0;
    for (int e = e0; e < e1; ++e) {
        const int c = en >> 1, side = en & 1;
        en = e + 1 < e1 ? ent[e + 1] : 0;
        const double* cg = w.cgrad + 18 * (sbase + c);
        const int act = w.cactive[sbase + c];
        const double* cd = w.cdbl + 7 * (sc.cb + c);
        const double n[3] = {cd[3], cd[4], cd[5]};
        const double b2 = b2arr[c];
        for (int k = 0; k < 3; ++k) gb[k] += cg[side * 3 + k];
        for (int k = 0; k < 6; ++k) gf[k] += cg[6 + side * 6 + k];
        if (act & 1) {
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int q = 0; q < 3; ++q) D[r * 6 + q] = D[r * 6 + q] + (b2 * (n[r] * n[q])) * sc.h2;
        }
        if ((act & 2) && e - e0 < 64) fmask |= 1ull << (e - e0);
    }
    for (int k = 0; k < 6; ++k) {
        g[k] = g[k] + sc.h2 * gb[k];
        g[k] = g[k] + sc.h2 * gf[k];
        cgv[k] = (0.0 + sc.h2 * gb[k]) + sc.h2 * gf[k];
    }
    // friction blocks, entries in order
    int ef = fmask ? ent[e0 + __ffsll(fmask) - 1] : 0;
    while (fmask) {
        fmask &= fmask - 1;
        const int c = ef >> 1, side = ef & 1;
        ef = fmask ? ent[e0 + __ffsll(fmask) - 1] : 0;
        const double* blk = w.cblk + 144 * (sbase + c) + side * 36;
#pragma unroll
        for (int k = 0; k < 36; ++k) D[k] = D[k] + blk[k];
    }
    for (int e = e0 + 64; e < e1; ++e) {
        const int c = ent[e] >> 1, side = ent[e] & 1;
        if (!(w.cactive[sbase + c] & 2)) continue;
        const double* blk = w.cblk + 144 * (sbase + c) + side * 36;
#pragma unroll
        for (int k = 0; k < 36; ++k) D[k] = D[k] + blk[k];
    }
}

// Objective value (IPCSolver.cpp:154-165) and contact value (IPCSolver.cpp:223-235).
// ------------------------------------------------------------------ ordered sums
// The reference accumulates the objective in plain loops over bodies, joints and
// contacts (its line-search slopes are Eigen dot products: serial_dot2 below). A
// parallel tree sum rounds differently, and the Armijo test and the ALM stopping test
// compare these values in the last bits, so they are summed left to right here: all
// lanes compute a window of terms into LDS, then one lane per accumulator adds them in
// order (LDS reads pipelined ahead of the dependent adds); independent accumulators sit
// on neighbouring lanes of one wavefront and advance in the same instructions.
static_assert(kBlk == 256, "ordered sums assume four wavefronts per work-group");

// acc + t[0] + t[1] + ... + t[n-1], t 16-byte aligned; the 16-byte reads of the next 16
// terms are in flight while the current 16 are added.
__device__ __forceinline__ double lds_serial_sum(double acc, const double* t, int n) {
    const double2* t2 = reinterpret_cast<const double2*>(t);
    const int n2 = n >> 1;
    int k = 0;
    if (n2 >= 8) {
        double2 a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = t2[q];
        for (; k + 16 <= n2; k += 8) {
            double2 b[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) b[q] = t2[k + 8 + q];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                acc += a[q].x;
                acc += a[q].y;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) a[q] = b[q];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            acc += a[q].x;
            acc += a[q].y;
        }
        k += 8;
    }
    for (; k < n2; ++k) {
        const double2 v = t2[k];
        acc += v.x;
        acc += v.y;
    }
    if (n & 1) acc += t[n - 1];
    return acc;
}

// Objective value (IPCSolver.cpp:144-160): inertia, h^2 gravity, joints, h^2 barrier,
// h^2 friction, each accumulator in the reference's loop order. Skipped elements add
// +0.0, which leaves a sum started at +0.0 unchanged. T: kTermLds doubles of LDS.
__device__ void objective_value(const Dev& d, const Work& w, const ScnCtx& sc, const double* X, double* red,
                                double* out_val, double* out_cval, double* T) {
    constexpr int WB = 1024;  // bodies / joints / contacts per window
    constexpr int OI = 0, OG = 2 * WB, OJ = 3 * WB, OB = 4 * WB, OF = 5 * WB;
    static_assert(6 * WB <= kTermLds, "term windows exceed the LDS buffer");
    const double* XH = w.xhat;
    const int nj = sc.j1 - sc.j0;
    int rounds = (sc.n + WB - 1) / WB;
    rounds = max(rounds, (nj + WB - 1) / WB);
    rounds = max(rounds, (sc.nc + WB - 1) / WB);
    double acc0 = 0.0;  // lane L < 5: accumulator L (inertia, gravity, joints, barrier, friction)
    for (int r = 0; r < rounds; ++r) {
        const int b_lo = r * WB;
        for (int t = threadIdx.x; t < WB; t += blockDim.x) {
            const int li = b_lo + t;
            if (li < sc.n) {
                const int b = sc.b0 + li;
                const double m = d.mass[b];
                const double* I = d.inertia + 9 * b;
                const V3 dp = xpos(X, b) - xpos(XH, b);
                const V3 dt = xrot(X, b) - xrot(XH, b);
                const V3 rr = 0.5 * dt;
                const double s0 = rr.x * I[0] + rr.y * I[3] + rr.z * I[6];
                const double s1 = rr.x * I[1] + rr.y * I[4] + rr.z * I[7];
                const double s2 = rr.x * I[2] + rr.y * I[5] + rr.z * I[8];
                T[OI + 2 * t] = 0.5 * m * sqnorm(dp);               // InertiaForm::value
                T[OI + 2 * t + 1] = s0 * dt.x + s1 * dt.y + s2 * dt.z;
                // GravityForm::value: gv -= m (g . x), and x - y == x + (-y)
                T[OG + t] = -(m * (0.0 * X[6 * b] + kG * X[6 * b + 1] + 0.0 * X[6 * b + 2]));
            }
            if (li < nj) {  // ConstraintForm::value: lambda.C + mu/2 |C|^2 per joint
                const int j = sc.j0 + li;  // Joint.cpp:5-17, Eigen's order (vecsum.h)
                double C[6];
                joint_C(d, j, X, C);
                const double t1 = eigen_dot6(d.jlambda + 6 * j, C);
                const double sq = eigen_sqnorm6(C);
                T[OJ + t] = t1 + 0.5 * d.jmu[j] * sq;
            }
            if (li < sc.nc) {  // DistanceBarrierForm::value, FrictionForm::value
                const long long c = sc.cb + li;
                T[OB + t] = sc.has_barrier ? bar_v(contact_dist(w, c, X), sc.dhat, sc.kappa) : 0.0;
                double fv = 0.0;
                if (sc.has_friction && li < sc.nfc) {
                    const double N = w.nf[(long long)sc.s * kMaxContacts + li];
                    if (N > 0.0) {
                        FKin k;
                        fric_kin(d, w, c, X, sc.fdt, k);
                        fv = fric_v(k.u, N, sc.fmu, sc.feps);
                    }
                }
                T[OF + t] = fv;
            }
        }
        __syncthreads();
        if (threadIdx.x < 5) {  // lanes 0..4 of the first wavefront, side by side
            const int nb = min(WB, max(0, sc.n - b_lo));
            const int ncw = min(WB, max(0, sc.nc - b_lo));
            const int lc = threadIdx.x;
            const double* base = lc == 0 ? T + OI : lc == 1 ? T + OG : lc == 2 ? T + OJ : lc == 3 ? T + OB : T + OF;
            const int cnt = lc == 0 ? 2 * nb : lc == 1 ? nb : lc == 2 ? min(WB, max(0, nj - b_lo)) : ncw;
            acc0 = lds_serial_sum(acc0, base, cnt);
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) red[threadIdx.x] = acc0;
    __syncthreads();
    const double vin = red[0], vg = red[1], vj = red[2], vb = red[3], vf = red[4];
    __syncthreads();
    double val = vin;
    val += sc.h2 * vg;
    val += vj;
    double cval = 0.0;
    if (sc.has_barrier) { val += sc.h2 * vb; cval += sc.h2 * vb; }
    if (sc.has_friction) { val += sc.h2 * vf; cval += sc.h2 * vf; }
    *out_val = val;
    *out_cval = cval;
}

// Objective and contact values at K trial points X0 + alpha_k DX0 (k < K <= kLsBatch,
// alpha_k = alpha / 2^k; X0, DX0 global body indexing): the line search's next K step
// lengths evaluated together. Each element's lane forms its bodies' trial points in
// registers with the same expression as the line search's x + alpha dx. Every value is
// summed exactly as in objective_value (same terms, same order, from +0.0); the 5 K
// accumulators run on lanes 0..5K-1 (the first two wavefronts), side by side, so K
// trials cost about one trial's serial chain. vals, cvals: K entries each, in LDS.
__device__ void objective_value_batch(const Dev& d, const Work& w, const ScnCtx& sc, int K, const double* X0,
                                      const double* DX0, double alpha, double* red, double* vals, double* cvals,
                                      double* T) {
    // Each trial owns TS doubles of LDS per round: 3 WBb body terms (two inertia terms
    // and one gravity term per body) and 3 WBe element terms (joint, barrier, friction),
    // the two widths chosen per scene so that the fewest rounds cover the bodies and
    // the elements (widths even: every window starts 16-byte aligned; with K > 1 at
    // most one element per lane, a single trial takes its window in slots of 256).
    const int TS = (kTermLds / K) & ~1;
    const double* XH = w.xhat;
    const int nj = sc.j1 - sc.j0;
    const int nel = max(nj, sc.nc);
    int rounds = 1, WBb = 0, WBe = 0;
    for (;; ++rounds) {
        WBb = (sc.n + rounds - 1) / rounds;
        WBe = (nel + rounds - 1) / rounds;
        WBb += WBb & 1;
        WBe += WBe & 1;
        // an element per lane at most when trials share an element's lanes (K > 1)
        if (3 * (WBb + WBe) <= TS && (WBe <= kBlk || K == 1)) break;
    }
    const int OI = 0, OG = 2 * WBb, OJ = 3 * WBb, OB = 3 * WBb + WBe, OF = 3 * WBb + 2 * WBe;
    double acc = 0.0;  // lane L < 5K: accumulator L % 5 of trial L / 5
    const int lk = threadIdx.x / 5, lc = threadIdx.x % 5;
    for (int r = 0; r < rounds; ++r) {
        const int b_lo = r * WBb, e_lo = r * WBe;
        // inertia and gravity: one lane per body, its K trial points formed in registers
        // (x + alpha_k dx, the expression that filled XTB)
        for (int tt = threadIdx.x; tt < WBb; tt += blockDim.x) {
            const int li = b_lo + tt;
            if (li >= sc.n) break;
            const int b = sc.b0 + li;
            const double m = d.mass[b];
            double I[9], xh[6], x0[6], dx[6];
            for (int q = 0; q < 9; ++q) I[q] = d.inertia[9 * b + q];
            for (int q = 0; q < 6; ++q) {
                xh[q] = XH[6 * b + q];
                x0[q] = X0[6 * b + q];
                dx[q] = DX0[6 * b + q];
            }
            double ak = alpha;
#pragma unroll
            for (int k = 0; k < kLsBatch; ++k) {
                if (k < K) {
                    double X[6];
                    for (int q = 0; q < 6; ++q) X[q] = x0[q] + ak * dx[q];
                    double* Tk = T + k * TS;
                    const V3 dp = v3(X[0], X[1], X[2]) - v3(xh[0], xh[1], xh[2]);
                    const V3 dt = v3(X[3], X[4], X[5]) - v3(xh[3], xh[4], xh[5]);
                    const V3 rr = 0.5 * dt;
                    const double s0 = rr.x * I[0] + rr.y * I[3] + rr.z * I[6];
                    const double s1 = rr.x * I[1] + rr.y * I[4] + rr.z * I[7];
                    const double s2 = rr.x * I[2] + rr.y * I[5] + rr.z * I[8];
                    Tk[OI + 2 * tt] = 0.5 * m * sqnorm(dp);
                    Tk[OI + 2 * tt + 1] = s0 * dt.x + s1 * dt.y + s2 * dt.z;
                    Tk[OG + tt] = -(m * (0.0 * X[0] + kG * X[1] + 0.0 * X[2]));
                }
                ak *= 0.5;
            }
        }
        // joints and contacts: the window's ne elements spread over the work-group, each
        // element on G lanes (G = min(K, work-group size / ne)); lane g of an element
        // evaluates trials g, g + G, ... with the element's data loaded once and the
        // trial points of its bodies formed in registers (alpha_k by k halvings, as the
        // line search forms it)
        const int ne = min(WBe, max(0, nel - e_lo));
        const int G = ne > 0 ? min(K, max(1, (int)blockDim.x / ne)) : 0;
        for (int u = threadIdx.x; u < ne * G; u += blockDim.x) {  // slot u: element u % ne, lane g
            const int tt = u % ne, g = u / ne;
            const int li = e_lo + tt;
            if (g < G && li < nj) {
                const int j = sc.j0 + li;
                const int A = d.ja[j], B = d.jb[j];
                double x0a[6], dxa[6], x0b[6], dxb[6];
                for (int q = 0; q < 6; ++q) {
                    x0a[q] = X0[6 * A + q];
                    dxa[q] = DX0[6 * A + q];
                    x0b[q] = B >= 0 ? X0[6 * B + q] : 0.0;
                    dxb[q] = B >= 0 ? DX0[6 * B + q] : 0.0;
                }
                for (int k = g; k < K; k += G) {
                    double ak = alpha;
                    for (int q = 0; q < k; ++q) ak *= 0.5;
                    double xa[6], xb[6], C[6];
                    for (int q = 0; q < 6; ++q) {
                        xa[q] = x0a[q] + ak * dxa[q];
                        xb[q] = B >= 0 ? x0b[q] + ak * dxb[q] : 0.0;
                    }
                    joint_C_x(d, j, xa, xb, C);
                    const double t1 = eigen_dot6(d.jlambda + 6 * j, C);
                    const double sq = eigen_sqnorm6(C);
                    T[k * TS + OJ + tt] = t1 + 0.5 * d.jmu[j] * sq;
                }
            }
            if (g < G && li < sc.nc) {
                const long long c = sc.cb + li;
                const int A = w.cint[5 * c], B = w.cint[5 * c + 1];
                double x0a[6], dxa[6], x0b[6], dxb[6];
                for (int q = 0; q < 6; ++q) {
                    x0a[q] = X0[6 * A + q];
                    dxa[q] = DX0[6 * A + q];
                    x0b[q] = X0[6 * B + q];
                    dxb[q] = DX0[6 * B + q];
                }
                const bool fr = sc.has_friction && li < sc.nfc;
                const double N = fr ? w.nf[(long long)sc.s * kMaxContacts + li] : 0.0;
                const FricPre fp = fric_pre(d, w, c);  // the contact's data, loaded once
                const double* cd = w.cdbl + 7 * c;
                const V3 cn = v3(cd[3], cd[4], cd[5]);
                for (int k = g; k < K; k += G) {
                    double ak = alpha;
                    for (int q = 0; q < k; ++q) ak *= 0.5;
                    double xa[6], xb[6];
                    for (int q = 0; q < 6; ++q) {
                        xa[q] = x0a[q] + ak * dxa[q];
                        xb[q] = x0b[q] + ak * dxb[q];
                    }
                    double* Tk = T + k * TS;
                    if (sc.has_barrier) {  // contact_dist_x with the normal in registers
                        const double pr = dot(v3(xb[0], xb[1], xb[2]) - v3(xa[0], xa[1], xa[2]), cn);
                        Tk[OB + tt] = bar_v(pr > 1e-9 ? pr : 1e-9, sc.dhat, sc.kappa);
                    } else {
                        Tk[OB + tt] = 0.0;
                    }
                    double fv = 0.0;
                    if (fr && N > 0.0) fv = fric_v(fric_u_pre(fp, xa, xb, sc.fdt), N, sc.fmu, sc.feps);
                    Tk[OF + tt] = fv;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < 5 * K) {
            const double* Tk = T + lk * TS;
            const int nb = min(WBb, max(0, sc.n - b_lo));
            const int ncw = min(WBe, max(0, sc.nc - e_lo));
            const int njw = min(WBe, max(0, nj - e_lo));
            const double* base = lc == 0 ? Tk + OI : lc == 1 ? Tk + OG : lc == 2 ? Tk + OJ : lc == 3 ? Tk + OB : Tk + OF;
            const int cnt = lc == 0 ? 2 * nb : lc == 1 ? nb : lc == 2 ? njw : ncw;
            acc = lds_serial_sum(acc, base, cnt);
        }
        __syncthreads();
    }
    if (threadIdx.x < 5 * K) red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < K) {  // lane k: trial k's values (vals, cvals in LDS)
        const int k = threadIdx.x;
        const double vin = red[5 * k], vg = red[5 * k + 1], vj = red[5 * k + 2], vb = red[5 * k + 3],
                     vf = red[5 * k + 4];
        double val = vin;
        val += sc.h2 * vg;
        val += vj;
        double cval = 0.0;
        if (sc.has_barrier) { val += sc.h2 * vb; cval += sc.h2 * vb; }
        if (sc.has_friction) { val += sc.h2 * vf; cval += sc.h2 * vf; }
        vals[k] = val;
        cvals[k] = cval;
    }
    __syncthreads();
}

// (a0.dot(b), a1.dot(b)) for dynamic-size vectors, in Eigen's order (vecsum.h): the
// four packet-lane chains of each product sum run on lanes 0..7 of the first wavefront
// (products stored per chain, contiguous, so each lane streams its own run), then the
// chains combine with the leftover terms. A chain starts from -0.0, the exact identity
// of IEEE addition, so its first add reproduces Eigen's start at t_r.
__device__ void serial_dot2(const double* a0, const double* a1, const double* b, int n, double* red, double* o0,
                            double* o1, double* T) {
    constexpr int Q = kTermLds / 8;  // terms per chain and window (a multiple of 2)
    static_assert(Q % 2 == 0, "chain windows start 16-byte aligned");
    const int n4 = n >= 4 ? n - n % 4 : 0;
    double acc = -0.0;  // lane l < 8: chain l % 4 of product l / 4
    for (int lo = 0; lo < n4; lo += 4 * Q) {
        const int cnt = min(4 * Q, n4 - lo);  // a multiple of 4
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
            const int i = lo + t, r = t & 3, m = t >> 2;
            T[r * Q + m] = a0[i] * b[i];
            T[(4 + r) * Q + m] = a1[i] * b[i];
        }
        __syncthreads();
        if (threadIdx.x < 8) acc = lds_serial_sum(acc, T + threadIdx.x * Q, cnt >> 2);
        __syncthreads();
    }
    if (threadIdx.x < 8) red[threadIdx.x] = acc;
    __syncthreads();
    double v[2];
    for (int k = 0; k < 2; ++k) {
        const double* a = k ? a1 : a0;
        auto t = [&](int i) { return a[i] * b[i]; };
        if (n < 4) {
            v[k] = eigen_vec_sum(n, t);
        } else {
            const int n2 = n - n % 2;
            v[k] = eigen_sum_finish(red[4 * k], red[4 * k + 1], red[4 * k + 2], red[4 * k + 3], n,
                                    n2 > n4 ? t(n4) : 0.0, n2 > n4 ? t(n4 + 1) : 0.0, (n % 2) ? t(n - 1) : 0.0);
        }
    }
    *o0 = v[0];
    *o1 = v[1];
    __syncthreads();
}

// Diagonal 6x6 block of body li, accumulated in registers (constant indices) in the
// reference's triplet order: inertia, joints, barrier contacts, friction contacts.
__device__ __forceinline__ void body_diag(const Dev& d, const Work& w, const ScnCtx& sc, const double* b2arr, int li,
                                          const int* off, const int* ent, double D[36]) {
    const int b = sc.b0 + li;
#pragma unroll
    for (int k = 0; k < 36; ++k) D[k] = 0.0;
    const double m = d.mass[b];
    D[0] = m; D[7] = m; D[14] = m;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) D[(3 + r) * 6 + 3 + c] = d.inertia[9 * b + 3 * r + c];
    for (int e = d.bj_off[b]; e < d.bj_off[b + 1]; ++e) {  // the body's joints, j ascending
        const double* jb = w.jblk + 108 * (d.bj_ent[e] >> 1) + 36 * (d.bj_ent[e] & 1);
#pragma unroll
        for (int k = 0; k < 36; ++k) D[k] = D[k] + jb[k];
    }
    for (int e = off[li]; e < off[li + 1]; ++e) {
        const int c = ent[e] >> 1;
        const long long cc = sc.cb + c;
        const long long ce = (long long)sc.s * kMaxContacts + c;
        if (!(w.cactive[ce] & 1)) continue;
        const double* cd = w.cdbl + 7 * cc;
        const double n[3] = {cd[3], cd[4], cd[5]};
        const double b2 = b2arr[c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) D[r * 6 + q] = D[r * 6 + q] + (b2 * (n[r] * n[q])) * sc.h2;
    }
    for (int e = off[li]; e < off[li + 1]; ++e) {
        const int c = ent[e] >> 1, side = ent[e] & 1;
        const long long ce = (long long)sc.s * kMaxContacts + c;
        if (!(w.cactive[ce] & 2)) continue;
        const double* blk = w.cblk + 144 * ce + side * 36;
#pragma unroll
        for (int k = 0; k < 36; ++k) D[k] = D[k] + blk[k];
    }
}

// Diagonal 6x6 blocks: inertia + joints + h2 (barrier, then friction) contributions.
__device__ void assemble_diag(const Dev& d, const Work& w, const ScnCtx& sc, const double* b2arr) {
    const int* off = w.csr_off + sc.b0 + sc.s;
    const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        double D[36];
        body_diag(d, w, sc, b2arr, li, off, ent, D);
        double* Dout = w.D + 36 * (sc.b0 + li);
#pragma unroll
        for (int k = 0; k < 36; ++k) Dout[k] = D[k];
    }
}

// y = H x over the scene (6 per body).
__device__ void spmv(const Dev& d, const Work& w, const ScnCtx& sc, const double* b2arr, const double* x, double* y) {
    const int* off = w.csr_off + sc.b0 + sc.s;
    const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const int b = sc.b0 + li;
        const double* D = w.D + 36 * b;
        double acc[6];
        for (int r = 0; r < 6; ++r) {
            double s = 0.0;
            for (int c = 0; c < 6; ++c) s += D[r * 6 + c] * x[6 * li + c];
            acc[r] = s;
        }
        for (int e = off[li]; e < off[li + 1]; ++e) {
            const int c = ent[e] >> 1, side = ent[e] & 1;
            const long long cc = sc.cb + c;
            const long long ce = (long long)sc.s * kMaxContacts + c;
            const int act = w.cactive[ce];
            if (!act) continue;
            const int other = (side == 0 ? w.cint[5 * cc + 1] : w.cint[5 * cc]) - sc.b0;
            if (act & 1) {
                const double* cd = w.cdbl + 7 * cc;
                const double n[3] = {cd[3], cd[4], cd[5]};
                for (int r = 0; r < 3; ++r) {
                    double s = 0.0;
                    for (int q = 0; q < 3; ++q) s += (-(b2arr[c] * (n[r] * n[q])) * sc.h2) * x[6 * other + q];
                    acc[r] += s;
                }
            }
            if (act & 2) {
                const double* blk = w.cblk + 144 * ce + (side == 0 ? 72 : 108);  // AB or BA
                for (int r = 0; r < 6; ++r) {
                    double s = 0.0;
                    for (int q = 0; q < 6; ++q) s += blk[r * 6 + q] * x[6 * other + q];
                    acc[r] += s;
                }
            }
        }
        for (int j = sc.j0; j < sc.j1; ++j) {
            if (d.jtype[j] != 1) continue;
            const double* jb = w.jblk + 108 * j + 72;
            int other = -1, side = 0;
            if (d.ja[j] == b) { other = d.jb[j]; side = 0; }
            else if (d.jb[j] == b) { other = d.ja[j]; side = 1; }
            if (other < 0) continue;
            const int lo = other - sc.b0;
            for (int r = 0; r < 6; ++r) {
                double s = 0.0;
                for (int q = 0; q < 6; ++q) s += (side == 0 ? jb[r * 6 + q] : jb[q * 6 + r]) * x[6 * lo + q];
                acc[r] += s;
            }
        }
        for (int r = 0; r < 6; ++r) y[6 * li + r] = acc[r];
    }
}

// 6x6 SPD inverse via Cholesky; returns false if not positive definite.
__device__ bool inv6(const double* A, double* out) {
    double L[36];
    for (int k = 0; k < 36; ++k) L[k] = 0.0;
    for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
        if (!(s > 0.0)) return false;
        const double dj = sqrt(s);
        L[j * 6 + j] = dj;
        for (int i = j + 1; i < 6; ++i) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = t / dj;
        }
    }
    // inverse = L^-T L^-1 column by column
    for (int c = 0; c < 6; ++c) {
        double y[6];
        for (int i = 0; i < 6; ++i) {
            double t = (i == c) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) t -= L[i * 6 + k] * y[k];
            y[i] = t / L[i * 6 + i];
        }
        for (int i = 5; i >= 0; --i) {
            double t = y[i];
            for (int k = i + 1; k < 6; ++k) t -= L[k * 6 + i] * out[k * 6 + c];
            out[i * 6 + c] = t / L[i * 6 + i];
        }
    }
    return true;
}

__device__ __forceinline__ void bj_apply(const Work& w, const ScnCtx& sc, const double* r, double* z) {
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const double* M = w.Dinv + 36 * (sc.b0 + li);
        for (int a = 0; a < 6; ++a) {
            double s = 0.0;
            for (int c = 0; c < 6; ++c) s += M[a * 6 + c] * r[6 * li + c];
            z[6 * li + a] = s;
        }
    }
}

__device__ double wg_dot(const double* a, const double* b, int n, double* red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += a[i] * b[i];
    return wg_reduce(s, red, OpAdd());
}
__device__ double wg_absmax(const double* a, int n, double* red) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s = dmax(s, fabs(a[i]));
    return wg_reduce(s, red, OpMax());
}

// Block-Jacobi PCG for H dx = rhs. Returns false when not converged.
__device__ bool pcg_solve(const Dev& d, const Work& w, const ScnCtx& sc, const double* b2arr, const double* rhs,
                          double* xo, double* red, int* ok_sh, int* iters = nullptr) {
    const int nd = sc.nd;
    double* r = w.r + 6 * sc.b0;
    double* z = w.z + 6 * sc.b0;
    double* p = w.pv + 6 * sc.b0;
    double* Ap = w.Ap + 6 * sc.b0;
    if (threadIdx.x == 0) *ok_sh = 1;
    __syncthreads();
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x)
        if (!inv6(w.D + 36 * (sc.b0 + li), w.Dinv + 36 * (sc.b0 + li))) atomicAnd(ok_sh, 0);
    __syncthreads();
    if (!*ok_sh) return false;
    bj_apply(w, sc, rhs, xo);
    __syncthreads();
    spmv(d, w, sc, b2arr, xo, Ap);
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) r[i] = rhs[i] - Ap[i];
    __syncthreads();
    const double bn = wg_absmax(rhs, nd, red);
    const double tol = 1e-14 * bn;
    double rn = wg_absmax(r, nd, red);
    if (rn <= tol) return true;
    bj_apply(w, sc, r, z);
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) p[i] = z[i];
    __syncthreads();
    double rz = wg_dot(r, z, nd, red);
    // at least twice the rows (exact arithmetic ends within nd iterations), so stiff
    // chains reach the 1e-14 relative residual instead of the 1e-10 fallback
    const int max_it = max(400, 2 * nd);
    for (int it = 0; it < max_it; ++it) {
        if (iters) *iters = it + 1;
        spmv(d, w, sc, b2arr, p, Ap);
        __syncthreads();
        const double pAp = wg_dot(p, Ap, nd, red);
        if (!(pAp > 0.0)) return false;
        const double alpha = rz / pAp;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) {
            xo[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        __syncthreads();
        rn = wg_absmax(r, nd, red);
        if (rn <= tol) return true;
        bj_apply(w, sc, r, z);
        __syncthreads();
        const double rz2 = wg_dot(r, z, nd, red);
        const double beta = rz2 / rz;
        rz = rz2;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) p[i] = z[i] + beta * p[i];
        __syncthreads();
    }
    return rn <= 1e-10 * bn;
}

// ------------------------------------------------------------------ direct solve
// The Newton matrix is block diagonal up to the bodies coupled by active contacts and
// hinge joints. Bodies are grouped into connected components (union-find over active
// elements); every component is factored by LDL^T in natural (global body) order with
// the operation order of a row-oriented skyline LDL^T:
//   L_ij = (A_ij - sum_k (L_ik D_k) L_jk) / D_j,   D_i = A_ii - sum_k (L_ik L_ik) D_k,
// k ascending. Entries of other components are exact zeros in that factorization, so a
// component-wise factorization equals the factorization of the whole matrix. Singletons
// use the same recurrences on their 6x6 block (one lane each); larger components are
// factored right-looking with the trailing update in parallel (same per-entry order).
// Components above kDenseBodies bodies fall back to PCG.
struct CompLds {
    int lab[kMaxSceneBodies];   // component label of each body: its smallest member
    int cnt[kMaxSceneBodies];   // component size, indexed by label
    int off[kMaxSceneBodies];   // offset of a multi-body component in list, by label
    int list[kMaxSceneBodies];  // members of the multi-body components, ascending in each
    int tmp[kMaxSceneBodies];   // coupled bodies, ascending
    int scan[kBlk];
    int maxsize, changed, ncoupled, pad;
};
union alignas(16) NewtonLds {  // terms: 16-byte reads in lds_serial_sum
    CompLds comp;
    double terms[kTermLds];
};

// Calls f(a, b) (scene-local body indices) for every coupling element in parallel:
// active contacts and hinge joints.
template <typename F>
__device__ __forceinline__ void for_each_coupling(const Dev& d, const Work& w, const ScnCtx& sc, F&& f) {
    for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
        if (!w.cactive[(long long)sc.s * kMaxContacts + i]) continue;
        const int* ci = w.cint + 5 * (sc.cb + i);
        f(ci[0] - sc.b0, ci[1] - sc.b0);
    }
    for (int j = sc.j0 + threadIdx.x; j < sc.j1; j += blockDim.x)
        if (d.jtype[j] == 1) f(d.ja[j] - sc.b0, d.jb[j] - sc.b0);
}

// Connected components of the coupling graph, all lanes: min-label propagation with
// root hooking and pointer jumping, then the members of every multi-body component
// listed in ascending order. With no coupling element every body is a singleton.
__device__ void build_components(const Dev& d, const Work& w, const ScnCtx& sc, CompLds& L) {
    const int n = sc.n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        L.lab[i] = i;
        L.cnt[i] = 0;
    }
    if (threadIdx.x == 0) {
        L.maxsize = 1;
        L.ncoupled = 0;
    }
    int any = 0;
    for_each_coupling(d, w, sc, [&](int, int) { any = 1; });
    if (!__syncthreads_or(any)) return;
    while (true) {
        if (threadIdx.x == 0) L.changed = 0;
        __syncthreads();
        for_each_coupling(d, w, sc, [&](int a, int b) {
            const int la = L.lab[a], lb = L.lab[b];
            if (la != lb) {
                const int m = la < lb ? la : lb;
                atomicMin(&L.lab[a], m);
                atomicMin(&L.lab[b], m);
                atomicMin(&L.lab[la], m);
                atomicMin(&L.lab[lb], m);
                L.changed = 1;
            }
        });
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            int l = L.lab[i];
            while (L.lab[l] < l) l = L.lab[l];
            L.lab[i] = l;
        }
        __syncthreads();
        const int ch = L.changed;
        __syncthreads();
        if (!ch) break;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&L.cnt[L.lab[i]], 1);
    __syncthreads();
    int mx = 1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mx = max(mx, L.cnt[i]);
    atomicMax(&L.maxsize, mx);
    // offsets of the multi-body components (by root, ascending) and the coupled bodies in
    // ascending order; then one lane places each coupled body after the earlier members
    // of its component (the off[] of a root serves as its cursor meanwhile)
    // (each lane a contiguous run of bodies; one scan of both counts, packed: each sum is
    // at most n <= 4096)
    static_assert(kMaxSceneBodies < (1 << 15), "packed scan of two body counts");
    const int T = (n + blockDim.x - 1) / blockDim.x;
    const int i0 = min(n, (int)threadIdx.x * T), i1 = min(n, i0 + T);
    int sv = 0, ncp = 0;
    for (int i = i0; i < i1; ++i) {
        const int root = L.lab[i];
        if (root == i && L.cnt[i] > 1) sv += L.cnt[i];
        if (L.cnt[root] > 1) ++ncp;
    }
    int ex;
    const int tot = wg_exclusive_scan(sv | (ncp << 16), L.scan, &ex);
    int ov = ex & 0xFFFF, oc = ex >> 16;
    for (int i = i0; i < i1; ++i) {
        const int root = L.lab[i];
        if (root == i && L.cnt[i] > 1) {
            L.off[i] = ov;
            ov += L.cnt[i];
        }
        if (L.cnt[root] > 1) L.tmp[oc++] = i;
    }
    const int runc = tot >> 16;
    __syncthreads();
    if (threadIdx.x == 0) {
        L.ncoupled = runc;
        for (int p = 0; p < runc; ++p) {
            const int i = L.tmp[p];
            L.list[L.off[L.lab[i]]++] = i;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        if (L.lab[i] == i && L.cnt[i] > 1) L.off[i] -= L.cnt[i];
    __syncthreads();
}

// 6x6 LDL^T (skyline recurrences) and solve of one body block; false on a zero pivot.
// The 6x6 block and the right-hand side are read once into registers and x written
// once: x may alias neither in the compiler's view, so working on x in place would turn
// every substitution step into a global-memory round trip.
__device__ bool ldlt6_solve(const double* A, const double* b, double* x) {
    double a[36], L[36], y[6];
#pragma unroll
    for (int i = 0; i < 36; ++i) a[i] = A[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] = b[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double s = a[i * 6 + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[k * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = s / L[j * 6 + j];
        }
        double dd = a[i * 6 + i];
#pragma unroll
        for (int k = 0; k < i; ++k) dd -= L[i * 6 + k] * L[i * 6 + k] * L[k * 6 + k];
        if (dd == 0.0 || !isfinite(dd)) return false;
        L[i * 6 + i] = dd;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = y[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * y[k];
        y[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] /= L[i * 6 + i];
#pragma unroll
    for (int i = 5; i >= 0; --i)
#pragma unroll
        for (int k = 0; k < i; ++k) y[k] -= L[i * 6 + k] * y[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) x[i] = y[i];
    return true;
}

// Returns 0 solved, 1 zero pivot, 2 a component exceeds kDenseBodies.
constexpr int kSmallRows = 24;  // components of up to 4 bodies: one wavefront each
constexpr int kSmallElems = kSceneScratch / 4;  // element list of each wavefront
// Lower-triangle entries (i, j), j <= i, of an m x m matrix, m = 6, 12, 18, 24, as
// (i << 8) | j ordered by j descending (then i ascending): the entries with j > k are the
// first (m-k-1)(m-k)/2, the trailing triangle of elimination step k.
struct SmallTri {
    unsigned short v[21 + 78 + 171 + 300];
};
constexpr SmallTri make_small_tri() {
    SmallTri t{};
    int n = 0;
    for (int mb = 1; mb <= 4; ++mb) {
        const int m = 6 * mb;
        for (int j = m - 1; j >= 1; --j)
            for (int i = j; i < m; ++i) t.v[n++] = (unsigned short)((i << 8) | j);
    }
    return t;
}
__constant__ constexpr SmallTri kSmallTriTab = make_small_tri();
__constant__ constexpr int kSmallTriOff[5] = {0, 0, 15, 15 + 66, 15 + 66 + 153};
#define kSmallTri (kSmallTriTab.v)

// Orders the wavefront's LDS accesses (no work-group barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane l's double from lane `src` (a wavefront-uniform index), through two scalar reads.
__device__ __forceinline__ double readlane_d(double v, int src) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// LDL^T factorization and solve of a 12 x 12 component (two bodies) with the matrix in
// registers: lane i < 12 holds row i (its lower part, read from the LDS tile M), every
// pivot and multiplier is broadcast from its lane by scalar reads, and each entry
// receives the same subtractions in the same k-ascending order, with the same
// expressions, as the LDS path (so the factor and the solution are bitwise those). x:
// the right-hand side's row of the lane in, the solution's out. Returns false on a zero
// or non-finite pivot.
__device__ __noinline__ bool small12_ldlt_solve(const double* M, double& x) {
    constexpr int m = 12;
    const int lane = threadIdx.x & 63;
    double a[m];
#pragma unroll
    for (int j = 0; j < m; ++j) a[j] = (lane < m && j <= lane) ? M[lane * m + j] : 0.0;
#pragma unroll
    for (int k = 0; k < m; ++k) {
        const double Dk = readlane_d(a[k], k);
        if (Dk == 0.0 || !isfinite(Dk)) return false;
        const bool below = lane > k && lane < m;
        const double l = below ? a[k] / Dk : 0.0;
        if (below) a[k] = l;
#pragma unroll
        for (int j = k + 1; j < m; ++j) {
            const double ljk = readlane_d(a[k], j);
            if (lane == j) a[j] -= (l * l) * Dk;
            else if (lane > j && lane < m) a[j] -= (l * Dk) * ljk;
        }
    }
    // forward x_i -= L_ik x_k (k ascending), x_i /= D_i, backward x_k -= L_ik x_i
    // (i descending)
#pragma unroll
    for (int k = 0; k < m; ++k) {
        const double xk = readlane_d(x, k);
        if (lane > k && lane < m) x -= a[k] * xk;
    }
    double dd = 1.0;
#pragma unroll
    for (int j = 0; j < m; ++j)
        if (lane == j) dd = a[j];
    if (lane < m) x /= dd;
#pragma unroll
    for (int i = m - 1; i >= 0; --i) {
        const double xi = readlane_d(x, i);
#pragma unroll
        for (int c = 0; c < i; ++c) {
            const double lic = readlane_d(a[c], i);
            if (lane == c) x -= lic * xi;
        }
    }
    return true;
}

// The component solve below for one component of at most 4 bodies (m <= kSmallRows),
// by the calling wavefront alone: the same assembly (diagonal blocks, then the
// elements one at a time in the same order), the same right-looking LDL^T and the same
// substitutions, entry for entry and operation for operation, with wavefront
// synchronisation in place of work-group barriers; the element list is compacted
// with ballots. M: the wavefront's m x m LDS tile; el: its element list. Returns false
// on a zero or non-finite pivot.
__device__ bool small_component_solve(const Dev& d, const Work& w, const ScnCtx& sc, const double* b2arr,
                                      const double* rhs, double* xo, const CompLds& L, int c0, int c1, double* M,
                                      int* el) {
    const int lane = threadIdx.x & 63;
    const int c = L.lab[L.list[c0]];
    const int mb = c1 - c0, m = 6 * mb;
    auto lpos = [&](int lb) {
        int lo = c0, hi = c1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (L.list[mid] < lb) lo = mid + 1;
            else hi = mid;
        }
        return lo - c0;
    };
    for (int t = lane; t < m * m; t += 64) M[t] = 0.0;
    wave_sync();
    for (int t = lane; t < 36 * mb; t += 64) {
        const int tb = t / 36, r = (t % 36) / 6, q = t % 6;
        M[(6 * tb + r) * m + 6 * tb + q] = w.D[36 * (sc.b0 + L.list[c0 + tb]) + r * 6 + q];
    }
    // The component's elements: hinges (kind 0), active barrier (1) and friction (2)
    // contacts whose first body is a member, gathered from the members' incidence lists
    // (body -> joints, body -> contacts) and ordered by (kind, index) with a rank count:
    // the order of a scan over all elements of the scene.
    (void)c;
    int ne = 0;
    const int* coff = w.csr_off + sc.b0 + sc.s;
    const int* cent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    auto append = [&](bool f, int code) {
        const unsigned long long bal = __ballot(f);
        if (f) el[ne + __popcll(bal & ((1ull << lane) - 1ull))] = code;
        ne += __popcll(bal);
    };
    for (int t = 0; t < mb; ++t) {
        const int lb = L.list[c0 + t], b = sc.b0 + lb;
        const int jl = d.bj_off[b], jh = d.bj_off[b + 1];
        for (int base = jl; base < jh; base += 64) {
            const int e = base + lane;
            bool f = false;
            int code = 0;
            if (e < jh) {
                const int je = d.bj_ent[e], j = je >> 1;
                f = (je & 1) == 0 && d.jtype[j] == 1;
                code = j - sc.j0;
            }
            append(f, code);
        }
        const int el0 = coff[lb], eh = coff[lb + 1];
        for (int base = el0; base < eh; base += 64) {
            const int e = base + lane;
            int act = 0, i = 0;
            if (e < eh) {
                const int ce = cent[e];
                i = ce >> 1;
                if ((ce & 1) == 0) act = w.cactive[(long long)sc.s * kMaxContacts + i];
            }
            append((act & 1) != 0, (1 << 24) | i);
            append((act & 2) != 0, (2 << 24) | i);
        }
    }
    wave_sync();
    for (int e = lane; e < ne; e += 64) {
        const int code = el[e];
        int r = 0;
        for (int q = 0; q < ne; ++q) r += el[q] < code;
        el[ne + r] = code;
    }
    wave_sync();
    el += ne;
    // Entry (r, q) of every 6x6 block is owned by lane 6 r + q, which adds the elements'
    // terms to it in element order; the lanes share no entry, so the elements need no
    // synchronisation between them and each lane's loads are issued back to back.
    if (lane < 36) {
        const int r = lane / 6, q = lane % 6;
        for (int e = 0; e < ne; ++e) {
            const int code = el[e], kind = code >> 24, i = code & 0xffffff;
            if (kind == 0) {
                const int j = sc.j0 + i;
                const int pa = lpos(d.ja[j] - sc.b0), pb = lpos(d.jb[j] - sc.b0);
                const double* jb = w.jblk + 108 * j + 72;  // rows a, cols b
                if (pa > pb) M[(6 * pa + r) * m + 6 * pb + q] += jb[r * 6 + q];
                else M[(6 * pb + r) * m + 6 * pa + q] += jb[q * 6 + r];
            } else {
                const long long cc = sc.cb + i, ce = (long long)sc.s * kMaxContacts + i;
                const int pa = lpos(w.cint[5 * cc] - sc.b0), pb = lpos(w.cint[5 * cc + 1] - sc.b0);
                const int hi = pa > pb ? pa : pb, lo = pa > pb ? pb : pa;
                if (kind == 1) {
                    if (r < 3 && q < 3) {
                        const double* cd = w.cdbl + 7 * cc;
                        M[(6 * hi + r) * m + 6 * lo + q] += (-(b2arr[i] * (cd[3 + r] * cd[3 + q]))) * sc.h2;
                    }
                } else {
                    const double* blk = w.cblk + 144 * ce + (pa > pb ? 72 : 108);
                    M[(6 * hi + r) * m + 6 * lo + q] += blk[r * 6 + q];
                }
            }
        }
    }
    wave_sync();
    if (m == 12) {  // two bodies: factor and solve in registers
        double x = lane < m ? rhs[6 * L.list[c0 + lane / 6] + lane % 6] : 0.0;
        if (!small12_ldlt_solve(M, x)) return false;
        if (lane < m) xo[6 * L.list[c0 + lane / 6] + lane % 6] = x;
        wave_sync();
        return true;
    }
    // the trailing triangle of step k: the first (m-k-1)(m-k)/2 entries of the component
    // size's table (entries ordered by column, descending)
    const unsigned short* tri = kSmallTri + kSmallTriOff[mb];
    for (int k = 0; k < m; ++k) {
        const double Dk = M[k * m + k];
        if (Dk == 0.0 || !isfinite(Dk)) return false;
        const int i = k + 1 + lane;
        if (i < m) {
            const double l = M[i * m + k] / Dk;
            M[i * m + k] = l;
            M[k * m + i] = l;
        }
        wave_sync();
        const int cnt = (m - k - 1) * (m - k) / 2;
        for (int t = lane; t < cnt; t += 64) {
            const int ij = tri[t], ii = ij >> 8, j = ij & 255;
            const double lik = M[ii * m + k], ljk = M[j * m + k];
            if (ii == j) M[ii * m + ii] -= (lik * lik) * Dk;
            else M[ii * m + j] -= (lik * Dk) * ljk;
        }
        wave_sync();
    }
    // row `lane` of x in a register: forward x_i -= L_ik x_k (k ascending), x_i /= D_i,
    // backward x_k -= L_ik x_i (i descending)
    double x = lane < m ? rhs[6 * L.list[c0 + lane / 6] + lane % 6] : 0.0;
    for (int k = 0; k < m; ++k) {
        const double xk = __shfl(x, k);
        if (lane > k && lane < m) x -= M[k * m + lane] * xk;
    }
    if (lane < m) x /= M[lane * m + lane];
    for (int i = m - 1; i >= 0; --i) {
        const double xi = __shfl(x, i);
        if (lane < i) x -= M[i * m + lane] * xi;
    }
    if (lane < m) xo[6 * L.list[c0 + lane / 6] + lane % 6] = x;
    wave_sync();
    return true;
}

// Right-looking LDL^T of the m x m matrix M (row-major, lower triangle): at step k,
// D_k = A_kk, L_ik = A_ik / D_k, then the trailing update A_ij -= (L_ik D_k) L_jk
// (A_jj -= (L_jk L_jk) D_k). Every entry receives the same subtractions in the same
// (k ascending) order as in the row-oriented recurrences of the oracle's skyline
// factorization, so the factor is bitwise theirs. L_ik is mirrored to the upper
// triangle (never read by the factorization) so that the forward substitution reads
// rows. Blocked in panels of P columns: within a panel each column is finished in turn
// and only the panel's later columns are updated at once; the trailing columns then
// receive the panel's P updates, k ascending, on 2 x 2 register tiles, so the matrix
// in HBM is read and written once per panel instead of once per column. The panel's
// L_ik and L_ik D_k (the same rounded product the update uses) are staged in LDS. Not
// inlined: inside the Newton kernel it would raise register pressure into spills.
// Returns false on a zero or non-finite pivot.
template <int P>
__device__ __noinline__ bool dense_ldlt_panels(double* M, int m) {
    __shared__ double pl[kDenseRows * P], pld[kDenseRows * P], pd[P];  // the panel
    for (int k0 = 0; k0 < m; k0 += P) {
        const int k1 = min(m, k0 + P), np = k1 - k0;
        for (int k = k0; k < k1; ++k) {
            const int kk = k - k0;
            const double Dk = M[k * m + k];
            if (Dk == 0.0 || !isfinite(Dk)) return false;
            for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) {
                const double l = M[i * m + k] / Dk;
                M[i * m + k] = l;
                M[k * m + i] = l;
                pl[i * P + kk] = l;
                pld[i * P + kk] = l * Dk;
            }
            if (threadIdx.x == 0) pd[kk] = Dk;
            __syncthreads();
            const int nc = k1 - k - 1;  // the panel's later columns j, rows i >= j
            if (nc > 0) {
                for (int t = threadIdx.x; t < (m - k - 1) * nc; t += blockDim.x) {
                    const int i = k + 1 + t / nc, j = k + 1 + t % nc;
                    if (j > i) continue;
                    const double lik = pl[i * P + kk];
                    if (i == j) M[i * m + i] -= (lik * lik) * Dk;
                    else M[i * m + j] -= pld[i * P + kk] * pl[j * P + kk];
                }
                __syncthreads();
            }
        }
        const int r = m - k1;  // trailing rows and columns k1 .. m-1
        if (r <= 0) break;
        const int nt = (r + 1) >> 1;
        const int ntri = nt * (nt + 1) / 2;
        for (int t = threadIdx.x; t < ntri; t += blockDim.x) {
            int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
            while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
            while (ti * (ti + 1) / 2 > t) --ti;
            const int tj = t - ti * (ti + 1) / 2;
            const int i0 = k1 + 2 * ti, j0 = k1 + 2 * tj;
            const bool r1 = i0 + 1 < m, c1 = j0 + 1 < m;
            const bool diag = ti == tj;
            // a00 (i0, j0), a10 (i0+1, j0), a01 (i0, j0+1), a11 (i0+1, j0+1); upper entries of
            // a diagonal tile are not part of the lower triangle
            const bool h01 = c1 && !diag, h10 = r1, h11 = r1 && c1;
            double a00 = M[i0 * m + j0];
            double a10 = h10 ? M[(i0 + 1) * m + j0] : 0.0;
            double a01 = h01 ? M[i0 * m + j0 + 1] : 0.0;
            double a11 = h11 ? M[(i0 + 1) * m + j0 + 1] : 0.0;
            for (int q = 0; q < np; ++q) {
                const double d = pd[q];
                const double ld0 = pld[i0 * P + q], ld1 = r1 ? pld[(i0 + 1) * P + q] : 0.0;
                const double lj0 = pl[j0 * P + q], lj1 = c1 ? pl[(j0 + 1) * P + q] : 0.0;
                if (diag) {
                    a00 -= (lj0 * lj0) * d;
                    a10 -= ld1 * lj0;
                    a11 -= (lj1 * lj1) * d;
                } else {
                    a00 -= ld0 * lj0;
                    a10 -= ld1 * lj0;
                    a01 -= ld0 * lj1;
                    a11 -= ld1 * lj1;
                }
            }
            M[i0 * m + j0] = a00;
            if (h10) M[(i0 + 1) * m + j0] = a10;
            if (h01) M[i0 * m + j0 + 1] = a01;
            if (h11) M[(i0 + 1) * m + j0 + 1] = a11;
        }
        __syncthreads();
    }
    return true;
}

// The gradient (G, and NG = -G), the diagonal blocks (assemble_diag) and the singleton
// solves of component_solve in one pass over the bodies, with the gradient and the 6x6
// block in registers: the same values as separate passes (a singleton's right-hand side
// is -g). The components of this
// iteration must be built (build_components); *ok_sh is cleared if a singleton's block
// has a zero or non-finite pivot.
// Body li of grad_diag_singletons: G, NG, cg and the diagonal block, and with single
// the singleton solve into xo. Returns false if that 6x6 factor fails; gmax takes the
// body's max |g|.
__device__ __forceinline__ bool grad_diag_body(const Dev& d, const Work& w, const ScnCtx& sc, const double* X,
                                               double* G, double* NG, double* cgout, const double* b2arr, int li,
                                               const int* off, const int* ent, bool single, double* xo,
                                               double& gmax) {
    double g[6], cgv[6], D[36];
    body_grad_diag(d, w, sc, X, b2arr, li, off, ent, g, cgv, D);
    for (int k = 0; k < 6; ++k) {
        G[6 * li + k] = g[k];
        NG[6 * li + k] = -g[k];
        cgout[6 * li + k] = cgv[k];
        gmax = dmax(gmax, fabs(g[k]));
    }
    double* Dout = w.D + 36 * (sc.b0 + li);
#pragma unroll
    for (int k = 0; k < 36; ++k) Dout[k] = D[k];
    if (!single) return true;
    double nb[6];
    for (int k = 0; k < 6; ++k) nb[k] = -g[k];
    return ldlt6_solve(D, nb, xo + 6 * li);
}

__device__ double grad_diag_singletons(const Dev& d, const Work& w, const ScnCtx& sc, const double* X, double* G,
                                       double* NG, double* cgout, const double* b2arr, const CompLds& L, double* xo,
                                       double* red, int* ok_sh) {
    double gmax = 0.0;
    const int* off = w.csr_off + sc.b0 + sc.s;
    const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x)
        if (!grad_diag_body(d, w, sc, X, G, NG, cgout, b2arr, li, off, ent, L.cnt[L.lab[li]] <= 1, xo, gmax))
            atomicAnd(ok_sh, 0);
    return wg_reduce(gmax, red, OpMax());
}

// singles_done: the caller has built the components and solved the singletons
// (grad_diag_singletons), *ok_sh holding their outcome.
template <int P>
__device__ int component_solve(const Dev& d, const Work& w, const ScnCtx& sc, const double* b2arr, const double* rhs,
                               double* xo, CompLds& L, int* ok_sh, bool singles_done = false) {
    if (!singles_done) {
        build_components(d, w, sc, L);
        if (L.maxsize > kDenseBodies) return 2;
        if (threadIdx.x == 0) *ok_sh = 1;
        __syncthreads();
        // singletons (a body without coupling elements has cnt 0)
        for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
            if (L.cnt[L.lab[li]] > 1) continue;
            if (!ldlt6_solve(w.D + 36 * (sc.b0 + li), rhs + 6 * li, xo + 6 * li)) atomicAnd(ok_sh, 0);
        }
    } else if (L.maxsize > kDenseBodies) {
        return 2;
    }
    // components of up to 4 bodies: wavefront v takes the small components numbered
    // v, v + 4, ... in list order
    __shared__ double Msmall[kBlk / 64][kSmallRows * kSmallRows];
    {
        const int wv = threadIdx.x >> 6;
        int* el = w.sort_idx + (long long)sc.s * kSceneScratch + wv * kSmallElems;
        int nsmall = 0;
        bool wave_ok = true;
        for (int c0 = 0; c0 < L.ncoupled && wave_ok;) {
            const int mb = L.cnt[L.lab[L.list[c0]]];
            if (6 * mb <= kSmallRows) {
                if ((nsmall & (kBlk / 64 - 1)) == wv)
                    wave_ok = small_component_solve(d, w, sc, b2arr, rhs, xo, L, c0, c0 + mb, Msmall[wv], el);
                ++nsmall;
            }
            c0 += mb;
        }
        if (!wave_ok && (threadIdx.x & 63) == 0) atomicAnd(ok_sh, 0);
    }
    __syncthreads();
    if (!*ok_sh) return 1;
    double* Mg = w.dense + (long long)sc.s * kDenseRows * kDenseRows;
    for (int c0 = 0; c0 < L.ncoupled;) {
        const int c = L.lab[L.list[c0]];  // the component's label (its first member)
        const int mb = L.cnt[c];
        const int c1 = c0 + mb;
        const int m = 6 * mb;
        if (m <= kSmallRows) {
            c0 = c1;
            continue;
        }
        double* M = Mg;
        // local row of body b: 6 * (index of b in the component's ascending list)
        auto lpos = [&](int lb) {
            int lo = c0, hi = c1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (L.list[mid] < lb) lo = mid + 1;
                else hi = mid;
            }
            return lo - c0;
        };
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) M[t] = 0.0;
        __syncthreads();
        for (int t = threadIdx.x; t < mb; t += blockDim.x) {
            const double* D = w.D + 36 * (sc.b0 + L.list[c0 + t]);
            for (int r = 0; r < 6; ++r)
                for (int q = 0; q < 6; ++q) M[(6 * t + r) * m + 6 * t + q] = D[r * 6 + q];
        }
        __syncthreads();
        // Lower-triangle off-diagonal blocks: every entry receives the hinges, then the
        // barrier contacts, then the friction contacts that couple its two bodies, each in
        // element order (the order of a scan over the scene's elements). Restricted to the
        // block of members (hi, lo), hi > lo, that order is the incidence lists of member
        // hi: its joints (ascending), then its contacts (ascending) with barrier terms,
        // then the same contacts with friction terms. Seven groups of 36 lanes take every
        // seventh block row; in a group lane (r, q) owns entry (r, q) of the block an
        // element adds to, so no two lanes of the work-group touch the same entry and the
        // rows need no barriers.
        {
            const int* coff = w.csr_off + sc.b0 + sc.s;
            const int* cent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
            const int grp = threadIdx.x / 36, t = threadIdx.x % 36, r = t / 6, q = t % 6;
            for (int hi = grp; grp < 7 && hi < mb; hi += 7) {
                const int lb = L.list[c0 + hi], b = sc.b0 + lb;
                double* Mrow = M + (6 * hi + r) * m + q;
                for (int e = d.bj_off[b]; e < d.bj_off[b + 1]; ++e) {
                    const int j = d.bj_ent[e] >> 1, side = d.bj_ent[e] & 1;
                    if (d.jtype[j] != 1) continue;
                    const int oth = side ? d.ja[j] : d.jb[j];
                    if (oth < 0) continue;
                    const int lo = lpos(oth - sc.b0);
                    if (lo >= hi) continue;
                    const double* jb = w.jblk + 108 * j + 72;  // rows a, cols b
                    Mrow[6 * lo] += side == 0 ? jb[r * 6 + q] : jb[q * 6 + r];
                }
                for (int kind = 1; kind <= 2; ++kind) {
                    for (int e = coff[lb]; e < coff[lb + 1]; ++e) {
                        const int i = cent[e] >> 1, side = cent[e] & 1;
                        const long long cc = sc.cb + i, ce = (long long)sc.s * kMaxContacts + i;
                        if (!(w.cactive[ce] & kind)) continue;
                        const int lo = lpos(w.cint[5 * cc + (side ? 0 : 1)] - sc.b0);
                        if (lo >= hi) continue;
                        if (kind == 1) {
                            if (r < 3 && q < 3) {
                                const double* cd = w.cdbl + 7 * cc;
                                Mrow[6 * lo] += (-(b2arr[i] * (cd[3 + r] * cd[3 + q]))) * sc.h2;
                            }
                        } else {
                            Mrow[6 * lo] += w.cblk[144 * ce + (side == 0 ? 72 : 108) + r * 6 + q];
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (!dense_ldlt_panels<P>(M, m)) return 1;
        // substitutions by the first wavefront, rows lane + 64 q in registers: forward
        // x_i -= L_ik x_k (k ascending), x_i /= D_i, backward x_k -= L_ik x_i (i
        // descending), the per-row operation order of the column-oriented substitutions
        double* xs = w.r + 6 * sc.b0;  // scratch of this scene, m <= 6 n
        if (threadIdx.x < 64) {
            constexpr int QR = kDenseRows / 64;
            const int lane = threadIdx.x;
            double xr[QR];
#pragma unroll
            for (int q = 0; q < QR; ++q) {
                const int i = lane + 64 * q;
                xr[q] = i < m ? rhs[6 * L.list[c0 + i / 6] + i % 6] : 0.0;
            }
            auto bcast = [&](int k) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < QR; ++q)
                    if ((k >> 6) == q) v = xr[q];
                return __shfl(v, k & 63);
            };
            double cur[QR], nxt[QR];
#pragma unroll
            for (int q = 0; q < QR; ++q) {
                const int i = lane + 64 * q;
                cur[q] = (i > 0 && i < m) ? M[i] : 0.0;  // row 0 of the mirror: L_i0
            }
            for (int k = 0; k < m; ++k) {
#pragma unroll
                for (int q = 0; q < QR; ++q) {
                    const int i = lane + 64 * q;
                    nxt[q] = (k + 1 < m && i > k + 1 && i < m) ? M[(k + 1) * m + i] : 0.0;
                }
                const double xk = bcast(k);
#pragma unroll
                for (int q = 0; q < QR; ++q) {
                    const int i = lane + 64 * q;
                    if (i > k && i < m) xr[q] -= cur[q] * xk;
                    cur[q] = nxt[q];
                }
            }
#pragma unroll
            for (int q = 0; q < QR; ++q) {
                const int i = lane + 64 * q;
                if (i < m) xr[q] /= M[i * m + i];
            }
#pragma unroll
            for (int q = 0; q < QR; ++q) {
                const int k = lane + 64 * q;
                cur[q] = (k < m - 1) ? M[(m - 1) * m + k] : 0.0;  // row m-1 of L
            }
            for (int i = m - 1; i >= 0; --i) {
#pragma unroll
                for (int q = 0; q < QR; ++q) {
                    const int k = lane + 64 * q;
                    nxt[q] = (i >= 1 && k < i - 1) ? M[(i - 1) * m + k] : 0.0;
                }
                const double xi = bcast(i);
#pragma unroll
                for (int q = 0; q < QR; ++q) {
                    const int k = lane + 64 * q;
                    if (k < i) xr[q] -= cur[q] * xi;
                    cur[q] = nxt[q];
                }
            }
#pragma unroll
            for (int q = 0; q < QR; ++q) {
                const int i = lane + 64 * q;
                if (i < m) xs[i] = xr[q];
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m; t += blockDim.x) xo[6 * L.list[c0 + t / 6] + t % 6] = xs[t];
        __syncthreads();
        c0 = c1;
    }
    return 0;
}

// ------------------------------------------------------------------ line-search helpers
// In a small batch (S (1 + kHelpers) work-groups fit the CUs at one per CU) each scene
// lends its line search to kHelpers helper work-groups running beside it: the scene's
// work-group posts a batch of K trials (K, the first step length) in the scene's
// HelperBox and waits; helper h < kLsBatch evaluates trial h with objective_value_batch
// at K = 1, whose sums are those of the scene's own batch evaluation, helper kLsBatch,
// when asked, sums the slopes g.dx, cg.dx with serial_dot2, and helper kLsBatch + 1 the
// objective and contact values at x with objective_value, so the values are the same
// bits whichever work-group computes them. Every wait is bounded in time: a
// helper that hears nothing in time marks the box dead and leaves, and a scene
// work-group that waits too long for an acknowledgement marks it dead, waits out any
// helper already working on the batch, and does the work itself; results never depend
// on the helpers being there. Ticks: wall_clock64, 100 MHz on gfx950.
// Within a Newton launch a helper waits up to idle_ticks for the next batch (the
// scene's work-group computes in between); for a launch to open, epoch_ticks, since the
// helpers and the Newton launches run side by side. A scene work-group waits wait_ticks
// for an answer. When dispatches are serialised (a profiler collecting counters, say)
// the helpers never run beside the Newton kernel: whichever side times out first marks
// the box dead, and the step's remaining launches run in-group at once. Batch numbers
// increase through a step and every per-step field is cleared at the step's begin, so
// nothing a helper saw in one launch or step is mistaken for a later one.
// The default bounds are kHelper*Ticks (ndipc_internal.h); ndipc_helper_config sets them.

__device__ __forceinline__ void st_rel(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_rlx(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_rlx(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_rlx(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// sequentially consistent: the give-up handshake between a scene and its helpers
__device__ __forceinline__ void st_sc(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_sc(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_SEQ_CST, __HIP_MEMORY_SCOPE_AGENT);
}

// Scene work-group, lane 0: opens this launch for the helpers (before any batch).
__device__ void helpers_open(HelperBox* hb) {
    st_rlx(&hb->quit, 0);
    st_rlx(&hb->bopen, ld_rlx(&hb->bseq));
    __hip_atomic_fetch_add(&hb->epoch, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Scene work-group, first wavefront (lane h: helper h; due: an answer is expected from
// it): waits until every due helper has acknowledged batch bseq, or the box is dead, or
// wait_ticks passed. All answered: one acquire makes the answers visible. Otherwise the
// scene gives up: it marks the box dead and quit, then waits for every due helper that
// already took the batch up (busy == bseq). A helper stores busy before it re-checks
// quit and dead, and the scene stores those before it reads busy, all sequentially
// consistent, so either the helper sees the give-up and does nothing or the scene sees
// busy and waits for its acknowledgement: no helper store of the batch lands after the
// scene has moved on (unless a helper stays busy past straggle_ticks, which is
// counted). A helper that sees the give-up after taking the batch up withdraws (busy
// back to 0) instead of answering. Returns 1 (all answered) or 0, the same on every
// lane.
__device__ __forceinline__ int helpers_wait(HelperBox* hb, int bseq, bool due) {
    const int h = threadIdx.x;
    const unsigned long long bound = ld_rlx(&hb->wait_ticks);
    unsigned long long t0 = wall_clock64();
    bool mine = !due || ld_rlx(&hb->ack[h]) == bseq;
    int ok = 1;
    while (!__all(mine)) {
        if (ld_rlx(&hb->dead) || wall_clock64() - t0 > bound) {
            ok = 0;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
        if (!mine) mine = ld_rlx(&hb->ack[h]) == bseq;
    }
    ok = __all(ok);
    if (ok) {
        if (h == 0) atomicAdd(&hb->n_helped, 1);
    } else {
        st_sc(&hb->dead, 1);
        st_sc(&hb->quit, 1);
        int straggle = 0;
        if (due && ld_sc(&hb->busy[h]) == bseq) {
            const unsigned long long sb = ld_rlx(&hb->straggle_ticks);
            t0 = wall_clock64();
            // until it answers, or withdraws (it saw the give-up after taking the batch)
            while (ld_rlx(&hb->ack[h]) != bseq && ld_rlx(&hb->busy[h]) == bseq) {
                if (wall_clock64() - t0 > sb) {
                    straggle = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        straggle = __any(straggle);
        if (h == 0) {
            atomicAdd(&hb->n_giveup, 1);
            if (straggle) atomicAdd(&hb->n_straggle, 1);
        }
    }
    // one acquire for all the acknowledgements (an acquire invalidates caches, so the
    // polls above are relaxed)
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok;
}

// Scene work-group, lane 0: posts a batch unless the box is dead; returns 1 if posted.
__device__ __forceinline__ int helpers_post(HelperBox* hb, int bseq) {
    if (ld_rlx(&hb->dead)) return 0;
    st_rel(&hb->bseq, bseq);
    return 1;
}

// Scene work-group, all lanes: the gradient, NG = -G, the contact part of the gradient,
// the diagonal blocks and the singleton solves of every body by the helpers (body li on
// helper li % kHelpers), w.hsingle holding which bodies are singletons. *ginf: max |g|;
// *ok_sh cleared if a singleton's factor failed. Returns false, with the helpers sent
// away, when they did not answer in time (or were gone already).
__device__ __noinline__ bool helpers_grad(HelperBox* hb, int bseq, double* ginf, int* ok_sh, int* flag) {
    __threadfence();  // the work-group's element data and singleton flags, visible to the helpers
    __syncthreads();
    if (threadIdx.x == 0) {
        st_rlx(&hb->job, 1);
        *flag = helpers_post(hb, bseq);
    }
    __syncthreads();
    if (!*flag) return false;
    __syncthreads();  // every lane has read *flag
    if (threadIdx.x < 64) {
        const int h = threadIdx.x;
        const bool due = h < kHelpers;
        const int ok = helpers_wait(hb, bseq, due);
        if (ok) {
            double gm = due ? ld_rlx(&hb->gm[h]) : 0.0;
            int okh = due ? ld_rlx(&hb->okf[h]) : 1;
            for (int o = 32; o > 0; o >>= 1) {
                gm = dmax(gm, __shfl_xor(gm, o));
                okh &= __shfl_xor(okh, o);
            }
            if (h == 0) {
                *ginf = gm;
                if (!okh) *ok_sh = 0;
            }
        }
        if (h == 0) *flag = ok;
    }
    __syncthreads();
    return *flag != 0;
}

// Scene work-group, all lanes: trials alpha, alpha / 2, ... (K of them) by the helpers,
// values to vals / cvals (LDS), with slopes (LDS, 2) also g.dx and cg.dx, with fxc
// (LDS, 2) also the objective and contact values at x; bseq: this batch's number.
// Returns false, with the helpers sent away, when they did not answer in time (or were
// gone already).
__device__ __noinline__ bool helpers_batch(HelperBox* hb, int bseq, int K, double alpha, double* vals, double* cvals,
                              double* slopes, double* fxc, int* flag) {
    static_assert(kHelpers <= 64, "one lane of the first wavefront per helper");
    __threadfence();  // the work-group's stores of x, g, cg and dx, visible to the helpers
    __syncthreads();
    if (threadIdx.x == 0) {
        st_rlx(&hb->job, 0);
        st_rlx(&hb->K, K);
        st_rlx(&hb->alpha, alpha);
        st_rlx(&hb->want_slopes, slopes ? 1 : 0);
        st_rlx(&hb->want_f, fxc ? 1 : 0);
        *flag = helpers_post(hb, bseq);
    }
    __syncthreads();
    if (!*flag) return false;
    __syncthreads();  // every lane has read *flag
    if (threadIdx.x < 64) {  // lane h waits for helper h's acknowledgement, if one is due
        const int h = threadIdx.x;
        const int ok = helpers_wait(hb, bseq, h < K || (h == kLsBatch && slopes) || (h == kLsBatch + 1 && fxc));
        if (ok) {
            if (h < K) {
                vals[h] = ld_rlx(&hb->val[h]);
                cvals[h] = ld_rlx(&hb->cval[h]);
            }
            if (slopes && h < 2) slopes[h] = ld_rlx(&hb->slope[h]);
            if (fxc && h < 2) fxc[h] = ld_rlx(&hb->fx[h]);
        }
        if (h == 0) *flag = ok;
    }
    __syncthreads();
    return *flag != 0;
}

// Helper work-group blockIdx.x = s kHelpers + h serves trial h (h < kLsBatch), the
// slopes (h = kLsBatch) or f(x) (h = kLsBatch + 1) of each batch of scene s, and a share
// of the bodies in gradient jobs, through the Newton launches of a step (each opens an
// epoch and ends with quit) until the step closes; it also leaves, marking the box dead
// for the rest of the step, when it hears nothing in time. It runs beside the step's
// kernels, on a second stream forked from and joined to theirs.
__global__ void __launch_bounds__(kBlk) k_newton_helper(Dev d, Work w) {
    __shared__ double red[kBlk];
    __shared__ alignas(16) double T[kTermLds];
    __shared__ double lv[2];
    __shared__ int cmd;
    __shared__ int K_sh, slopes_sh, f_sh, job_sh, okh_sh;
    __shared__ double alpha_sh;
    const int s = blockIdx.x / kHelpers, h = blockIdx.x % kHelpers;
    HelperBox* hb = w.hbox + s;
    int served = 0;  // lane 0: the last batch taken up
    for (;;) {
        if (threadIdx.x == 0) {  // wait for the next Newton launch to open, or the step to close
            const int last = ld_rlx(&hb->hep[h]);
            const unsigned long long bound = ld_rlx(&hb->epoch_ticks);
            const unsigned long long t0 = wall_clock64();
            int ep = last, cl = 0;
            while ((ep = ld_rlx(&hb->epoch)) == last && !(cl = ld_rlx(&hb->closed)) && wall_clock64() - t0 <= bound)
                __builtin_amdgcn_s_sleep(8);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (ep != last) {
                st_rlx(&hb->hep[h], ep);
                served = ld_rlx(&hb->bopen);  // batches up to here belong to earlier launches
                cmd = 0;
            } else {
                if (!cl) st_sc(&hb->dead, 1);  // timed out
                cmd = -1;
            }
        }
        __syncthreads();
        if (cmd < 0) return;
        // the scene's parameters of this launch (contact count, barrier stiffness, ...),
        // set by the kernels before it
        ScnCtx sc;
        load_scn(d, s, sc);
        for (;;) {
            if (threadIdx.x == 0) {
                const unsigned long long bound = ld_rlx(&hb->idle_ticks);
                const unsigned long long t0 = wall_clock64();
                int c = -1;
                for (;;) {  // relaxed polls (an acquire invalidates caches), then one acquire
                    if (ld_rlx(&hb->quit)) break;
                    const int b = ld_rlx(&hb->bseq);
                    if (b != served) {
                        served = b;
                        // take the batch up, then re-check whether the scene gave up on it
                        st_sc(&hb->busy[h], b);
                        const int gone = ld_sc(&hb->dead), q = ld_sc(&hb->quit);
                        if (gone || q) {  // given up: withdraw, so the scene need not wait
                            st_sc(&hb->busy[h], 0);
                            if (gone) c = -2;
                            break;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                        c = b;
                        job_sh = ld_rlx(&hb->job);
                        K_sh = ld_rlx(&hb->K);
                        slopes_sh = ld_rlx(&hb->want_slopes);
                        f_sh = ld_rlx(&hb->want_f);
                        alpha_sh = ld_rlx(&hb->alpha);
                        break;
                    }
                    if (wall_clock64() - t0 > bound) {
                        st_sc(&hb->dead, 1);
                        c = -2;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(4);
                }
                cmd = c;
            }
            __syncthreads();
            const int c = cmd;
            if (c == -2) return;
            if (c < 0) break;  // this launch is closed
            // lane 0's acquire, then the barrier: x, dx of the batch are visible
            if (job_sh == 1) {  // gradient, diagonal blocks, singleton solves: bodies h, h + kHelpers, ...
                const int* off = w.csr_off + sc.b0 + sc.s;
                const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
                const double* b2arr = w.sort_key + (long long)s * kSceneScratch;
                if (threadIdx.x == 0) okh_sh = 1;
                __syncthreads();
                double gmax = 0.0;
                for (int li = h + kHelpers * (int)threadIdx.x; li < sc.n; li += kHelpers * (int)blockDim.x)
                    if (!grad_diag_body(d, w, sc, w.x, w.g + 6 * sc.b0, w.tmp + 6 * sc.b0, w.cg + 6 * sc.b0, b2arr, li,
                                        off, ent, w.hsingle[sc.b0 + li] != 0, w.dx + 6 * sc.b0, gmax))
                        atomicAnd(&okh_sh, 0);
                gmax = wg_reduce(gmax, red, OpMax());
                if (threadIdx.x == 0) {
                    st_rlx(&hb->gm[h], gmax);
                    st_rlx(&hb->okf[h], okh_sh);
                    st_rel(&hb->ack[h], c);
                }
            } else if (h < kLsBatch && h < K_sh) {
                double a = alpha_sh;  // alpha / 2^h, halved as the scene's batch halves it
                for (int q = 0; q < h; ++q) a *= 0.5;
                objective_value_batch(d, w, sc, 1, w.x, w.dx, a, red, lv, lv + 1, T);
                if (threadIdx.x == 0) {
                    st_rlx(&hb->val[h], lv[0]);
                    st_rlx(&hb->cval[h], lv[1]);
                    st_rel(&hb->ack[h], c);
                }
            } else if (h == kLsBatch && slopes_sh) {
                double dsc, cdsc;
                serial_dot2(w.g + 6 * sc.b0, w.cg + 6 * sc.b0, w.dx + 6 * sc.b0, sc.nd, red, &dsc, &cdsc, T);
                if (threadIdx.x == 0) {
                    st_rlx(&hb->slope[0], dsc);
                    st_rlx(&hb->slope[1], cdsc);
                    st_rel(&hb->ack[h], c);
                }
            } else if (h == kLsBatch + 1 && f_sh) {
                double fv, fcv;
                objective_value(d, w, sc, w.x, red, &fv, &fcv, T);
                if (threadIdx.x == 0) {
                    st_rlx(&hb->fx[0], fv);
                    st_rlx(&hb->fx[1], fcv);
                    st_rel(&hb->ack[h], c);
                }
            }
            __syncthreads();
        }
        __syncthreads();
    }
}

// One work-group per scene; with newton_helpers, k_newton_helper runs beside it.
__global__ void __launch_bounds__(kBlk) k_newton(Dev d, Work w) {
    __shared__ double red[kBlk];
    __shared__ int ok_sh;
    __shared__ NewtonLds nl;
    __shared__ double lsv[2 * kLsBatch];  // trial values, then contact values
    __shared__ double slp[2];             // slopes from a helper
    __shared__ double fxs[2];             // f(x) and the contact value from a helper
    __shared__ double gsh;                // max |g| from the helpers
    __shared__ int hflag;                 // the helpers answered
    __shared__ int hopen[2];              // batch number at this launch's open, helpers alive
    CompLds& comp = nl.comp;
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    HelperBox* hb = w.newton_helpers ? w.hbox + s : nullptr;
    if (hb && threadIdx.x == 0) {
        helpers_open(hb);
        hopen[0] = ld_rlx(&hb->bseq);
        hopen[1] = !ld_rlx(&hb->dead);
    }
    if (sp.alm_done) {
        if (hb && threadIdx.x == 0) st_rel(&hb->quit, 1);
        return;
    }
    __syncthreads();
    ScnCtx sc;
    load_scn(d, s, sc);
    double* X = w.x + 6 * sc.b0;
    double* G = w.g + 6 * sc.b0;
    double* DX = w.dx + 6 * sc.b0;
    double* CG = w.cg + 6 * sc.b0;
    double* NG = w.tmp + 6 * sc.b0;
    double* b2arr = w.sort_key + (long long)s * kSceneScratch;  // per-contact b'' scratch (<= 2000)
    int conv = 0;
    int it = 0;
    const double* Xg = w.x;
    const int maxit = sp.newton_max_iters;
    // The accepted trial point of a line search is bitwise the next iterate (both are
    // x + alpha dx with the same alpha), and objective_value_batch sums its terms as
    // objective_value does: the next iteration takes f(x) and the contact value from
    // the accepted trial instead of evaluating them again.
    double fx_next = 0.0, fc_next = 0.0;
    bool have_f = false;
    int kprev = kLsBatchLarge;            // trials the previous line search used
    int n_newton = 0, n_pcg = 0, n_pcg_iters = 0, n_direct = 0, comp_max = 0, n_trials = 0;  // launch counts
    bool helped = hb && hopen[1];         // the helpers answer (until they fail to)
    int bseq = hb ? hopen[0] : 0;         // batches posted to the helpers in this step
    const int Kb = 3 * (sc.n + (sc.n & 1) + max(sc.j1 - sc.j0, sc.nc) + 1) <= ((kTermLds / kLsBatch) & ~1)
                       ? kLsBatch
                       : kLsBatchLarge;
    for (; it < maxit; ++it) {
        eval_elements(d, w, sc, Xg, b2arr);
        // coupled components (active contacts and hinges of this iterate), then the
        // gradient, the diagonal blocks and the singleton solves in one pass
        build_components(d, w, sc, comp);
        if (threadIdx.x == 0) ok_sh = 1;
        __syncthreads();
        if (helped) {  // by the helpers, one body per lane
            for (int li = threadIdx.x; li < sc.n; li += blockDim.x)
                w.hsingle[sc.b0 + li] = comp.cnt[comp.lab[li]] <= 1;
            helped = helpers_grad(hb, ++bseq, &gsh, &ok_sh, &hflag);
        }
        const double ginf =
            helped ? gsh : grad_diag_singletons(d, w, sc, Xg, G, NG, CG, b2arr, comp, DX, red, &ok_sh);
        __syncthreads();
        if (ginf < sp.newton_tol) { conv = 1; break; }
        const int cs = comp.maxsize > kDenseBodies ? 2
                       : !ok_sh                   ? 1
                                                  : component_solve<8>(d, w, sc, b2arr, NG, DX, comp, &ok_sh, true);
        int pit = 0;
        const bool ok = cs == 0 ? true : cs == 2 ? pcg_solve(d, w, sc, b2arr, NG, DX, red, &ok_sh, &pit) : false;
        ++n_newton;
        if (cs == 2) {
            ++n_pcg;
            n_pcg_iters += pit;
        } else {
            ++n_direct;
        }
        comp_max = max(comp_max, comp.maxsize);
        __syncthreads();  // comp shares LDS with the term windows of the line search
        if (!ok) {
            if (threadIdx.x == 0) atomicOr(&sp.solve_fail, 1);
            break;
        }
        // line search (LineSearch.cpp:6-58), alpha0 = min(1, maxStep) = 1
        // f(x): carried over from the accepted trial, else by a helper beside the first
        // batch or here before it
        double fx = fx_next, fc = fc_next;
        // slopes g.dx and cg.dx: by a helper beside the first batch, or here before it;
        // an uphill direction (g.dx > 0) ends the Newton loop before any trial counts
        double descent = 0.0, cdesc = 0.0;
        bool have_desc = false;
        // The trials alpha = 1, 1/2, 1/4, ... (at most 20) are evaluated in batches: the
        // first as many as the previous line search needed, the next ones all that
        // remain, each at most Kb (20 where their terms fit one LDS round, 8
        // otherwise), or with helpers all remaining trials at once, one per helper.
        // The first that passes is taken, and only the trials up to it count.
        double alpha = 1.0;
        bool accepted = false;
        int ls_used = 0;
        for (int ls = 0; ls < 20 && !accepted;) {
            const int K = helped ? 20 - ls : min(Kb, ls == 0 ? kprev : 20 - ls);
            if (helped) {
                helped = helpers_batch(hb, ++bseq, K, alpha, lsv, lsv + kLsBatch, have_desc ? nullptr : slp,
                                       have_f ? nullptr : fxs, &hflag);
                if (helped && !have_desc) {
                    descent = slp[0];
                    cdesc = slp[1];
                    have_desc = true;
                }
                if (helped && !have_f) {
                    fx = fxs[0];
                    fc = fxs[1];
                    have_f = true;
                }
            }
            if (!have_desc) {
                serial_dot2(G, CG, DX, sc.nd, red, &descent, &cdesc, nl.terms);
                have_desc = true;
            }
            if (!have_f) {
                objective_value(d, w, sc, Xg, red, &fx, &fc, nl.terms);
                have_f = true;
            }
            if (descent > 0) break;  // uphill: not accepted, the Newton loop ends
            if (!helped) objective_value_batch(d, w, sc, K, w.x, w.dx, alpha, red, lsv, lsv + kLsBatch, nl.terms);
            int k = 0;
            for (; k < K; ++k) {
                const double fnk = lsv[k], fcnk = lsv[kLsBatch + k];
                const bool cok = (cdesc <= 0.0) ? (fcnk <= fc + 1e-4 * alpha * cdesc) : (fcnk <= fc);
                if (fnk <= fx + 1e-4 * alpha * descent && cok) {
                    accepted = true;
                    fx_next = fnk;
                    fc_next = fcnk;
                    have_f = true;
                    break;
                }
                alpha *= 0.5;
            }
            __syncthreads();  // lsv is rewritten by the next batch
            const int used = accepted ? k + 1 : K;
            n_trials += used;
            ls += used;
            ls_used = ls;
        }
        kprev = ls_used;
        if (!accepted) break;
        for (int i = threadIdx.x; i < sc.nd; i += blockDim.x) X[i] = X[i] + alpha * DX[i];
        __syncthreads();
    }
    if (hb && threadIdx.x == 0) st_rel(&hb->quit, 1);
    __syncthreads();
    // ALM bookkeeping (IPCSolver.cpp:1084-1110): violation, dual residual, hardening,
    // lambda += mu C with the hardened mu.
    double viol = 0.0, dual = 0.0;
    for (int j = sc.j0 + threadIdx.x; j < sc.j1; j += blockDim.x) {
        double C[6];
        joint_C(d, j, Xg, C);
        double inf = 0.0;
        for (int k = 0; k < 6; ++k) inf = dmax(inf, fabs(C[k]));
        viol = dmax(viol, inf);
        const double dc = d.jmu[j] * inf;
        dual = isfinite(dc) ? dmax(dual, dmin(dc, 1e12)) : dmax(dual, 1e12);
        if (inf > sp.alm_trig * sp.alm_ctol) {
            const double nm = d.jmu[j] * sp.alm_ratio;
            d.jmu[j] = isfinite(nm) ? dmin(nm, 1e12) : 1e12;
        }
        for (int k = 0; k < 6; ++k) d.jlambda[6 * j + k] += d.jmu[j] * C[k];
    }
    viol = wg_reduce(viol, red, OpMax());
    dual = wg_reduce(dual, red, OpMax());
    if (threadIdx.x == 0) {
        // this launch's counts (integers: the double totals are exact sums either way)
        sp.newton_total += double(n_newton);
        sp.solves_pcg += n_pcg;
        sp.pcg_iters += n_pcg_iters;
        sp.solves_direct += n_direct;
        sp.comp_max = max(sp.comp_max, comp_max);
        sp.ls_trials += n_trials;
        sp.trials_total += double(n_trials);
        sp.converged = conv;
        sp.newton_iters += (it < maxit ? it : maxit - 1) + 1;  // NewtonSolver reports the last index
        sp.viol = viol;
        sp.dual = dual;
    }
    // the poses of the new iterate for the narrow phase that follows
    // (detectContactsAtState(world, x_new), IPCSolver.cpp:1113)
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const int b = sc.b0 + li;
        st3(w.pose_p, b, v3(X[6 * li], X[6 * li + 1], X[6 * li + 2]));
        st4(w.pose_q, b, apply_theta(ld4(d.q, b), v3(X[6 * li + 3], X[6 * li + 4], X[6 * li + 5])));
    }
}

void launch_newton(const Dev& d, const Work& w, hipStream_t st) {
    timing_begin("newton", st);
    hipLaunchKernelGGL(k_newton, dim3(d.S), dim3(kBlk), 0, st, d, w);
    timing_end("newton", st);
}

void launch_newton_helpers(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_newton_helper, dim3(d.S * kHelpers), dim3(kBlk), 0, st, d, w);
}

// Ends the step's helper launch: once the step's last Newton launch has run, every
// scene's helpers leave (they wait for the next launch or for this flag).
__global__ void k_helpers_close(Dev d, Work w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < d.S) st_rel(&w.hbox[s].closed, 1);
}

void launch_helpers_close(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_helpers_close, dim3((d.S + 255) / 256), dim3(256), 0, st, d, w);
}

// ------------------------------------------------------------------ contacts setup
// Body -> contact CSR in contact order (each contact once under A, once under B).
// Scenes of up to 2 kMaxContacts bodies: per-body counts by LDS atomics, a scan, then
// one wavefront places 64 contacts per round in contact order: an entry's place in its
// body's list is the body's running count plus the entries of the same body in lower
// lanes (A before B within a contact), found with equality masks built from one ballot
// per key bit. Larger scenes: every body's lane scans the contacts (LDS broadcasts).
__device__ void build_csr(const Work& w, const ScnCtx& sc) {
    int* off = w.csr_off + sc.b0 + sc.s;
    int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    __shared__ int2 cab[kMaxContacts];
    __shared__ int csr_sh[kBlk / 64];
    if (sc.n <= 2 * kMaxContacts) {
        static_assert(2 * kMaxContacts <= (1 << 12), "body keys of 12 bits");
        int* cnt = reinterpret_cast<int*>(cab);  // per-body counts, then running places
        for (int li = threadIdx.x; li < sc.n; li += blockDim.x) cnt[li] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
            const int* ci = w.cint + 5 * (sc.cb + i);
            const int a = ci[0] - sc.b0, b = ci[1] - sc.b0;
            if (a >= 0 && a < sc.n) atomicAdd(&cnt[a], 1);
            if (b >= 0 && b < sc.n) atomicAdd(&cnt[b], 1);
        }
        __syncthreads();
        int run = 0;
        for (int base = 0; base < sc.n; base += blockDim.x) {  // counts -> offsets
            const int li = base + threadIdx.x;
            const int v = li < sc.n ? cnt[li] : 0;
            int ex;
            const int tot = wg_scan_shfl(v, csr_sh, &ex);
            if (li < sc.n) {
                off[li + 1] = run + ex + v;
                cnt[li] = run + ex;
            }
            run += tot;
        }
        if (threadIdx.x == 0) off[0] = 0;
        __syncthreads();
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const unsigned long long below = (1ull << lane) - 1;
            for (int base = 0; base < sc.nc; base += 64) {
                const int i = base + lane;
                int a = -1, b = -1;
                if (i < sc.nc) {
                    const int* ci = w.cint + 5 * (sc.cb + i);
                    a = ci[0] - sc.b0;
                    b = ci[1] - sc.b0;
                }
                const bool va = a >= 0 && a < sc.n, vb = b >= 0 && b < sc.n;
                // eqXY: lanes whose X-side body equals this lane's Y-side body
                unsigned long long eqAA = __ballot(va), eqBA = __ballot(vb), eqAB = eqAA, eqBB = eqBA;
                for (int k = 0; k < 12; ++k) {
                    const unsigned long long mA = __ballot(va && ((a >> k) & 1));
                    const unsigned long long mB = __ballot(vb && ((b >> k) & 1));
                    const bool ak = (a >> k) & 1, bk = (b >> k) & 1;
                    eqAA &= ak ? mA : ~mA;
                    eqBA &= ak ? mB : ~mB;
                    eqAB &= bk ? mA : ~mA;
                    eqBB &= bk ? mB : ~mB;
                }
                int pa = 0, pb = 0;
                if (va) pa = cnt[a] + __popcll(eqAA & below) + __popcll(eqBA & below);
                if (vb) pb = cnt[b] + __popcll(eqAB & below) + __popcll(eqBB & below) + (va && a == b);
                if (va) ent[pa] = (i << 1) | 0;
                if (vb) ent[pb] = (i << 1) | 1;
                __builtin_amdgcn_wave_barrier();  // every lane has read its places
                // the first entry of a body in this round advances the body's running place
                if (va && !((eqAA | eqBA) & below))
                    cnt[a] += __popcll(eqAA) + __popcll(eqBA);
                if (vb && !((eqAB | eqBB) & below) && !(va && a == b))
                    cnt[b] += __popcll(eqAB) + __popcll(eqBB);
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
        return;
    }
    // the contacts' body pairs in LDS: every lane then scans them in the same order, so
    // each read is a broadcast
    for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
        const int* ci = w.cint + 5 * (sc.cb + i);
        cab[i] = make_int2(ci[0], ci[1]);
    }
    __syncthreads();
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const int b = sc.b0 + li;
        int c = 0;
        for (int i = 0; i < sc.nc; ++i) {
            const int2 ab = cab[i];
            c += (ab.x == b) + (ab.y == b);
        }
        off[li + 1] = c;
    }
    __syncthreads();
    int run = 0;
    for (int base = 0; base < sc.n; base += blockDim.x) {  // counts -> inclusive prefix sums
        const int li = base + threadIdx.x;
        const int v = li < sc.n ? off[li + 1] : 0;
        int ex;
        const int tot = wg_scan_shfl(v, csr_sh, &ex);
        if (li < sc.n) off[li + 1] = run + ex + v;
        run += tot;
    }
    if (threadIdx.x == 0) off[0] = 0;
    __syncthreads();
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const int b = sc.b0 + li;
        int k = off[li];
        for (int i = 0; i < sc.nc; ++i) {
            const int2 ab = cab[i];
            if (ab.x == b) ent[k++] = (i << 1) | 0;
            if (ab.y == b) ent[k++] = (i << 1) | 1;
        }
    }
    __syncthreads();
}

// Contact forms of the step (IPCSolver.cpp:872-921): barrier on the step-start
// contacts, friction with mu_eff = mean of per-pair overrides and no normal forces yet.
__global__ void __launch_bounds__(kBlk) k_contacts_setup(Dev d, Work w) {
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    if (as_skip(w, sp)) return;
    const bool needs = sp.enable_adaptive || (sp.enable_friction && sp.mu > 0.0);
    if (threadIdx.x == 0) {
        if (!needs) sp.n_contacts = 0;
        sp.needs_forms = needs ? 1 : 0;
        sp.has_barrier = (sp.enable_adaptive && needs) ? 1 : 0;
        sp.has_friction = (sp.enable_friction && sp.mu > 0.0) ? 1 : 0;
        double emu = sp.mu;
        const bool any_override = d.allowed_off[s + 1] > d.allowed_off[s];
        if (sp.has_friction && sp.n_contacts > 0 && any_override) {
            bool has = false;
            for (int k = d.allowed_off[s]; k < d.allowed_off[s + 1]; ++k) has = has || d.allowed_mu[k] >= 0.0;
            if (has) {
                double sum = 0.0;
                int cnt = 0;
                const long long cb = (long long)s * kMaxRawContacts;
                for (int i = 0; i < sp.n_contacts; ++i) {
                    const int A = w.cint[5 * (cb + i)], B = w.cint[5 * (cb + i) + 1];
                    const double m = allowed_mu(d, s, d.body_id[A], d.body_id[B]);
                    sum += (m >= 0.0) ? m : sp.mu;
                    ++cnt;
                }
                if (cnt > 0) emu = sum / double(cnt);
            }
        }
        sp.f_mu = emu;
        sp.nf_count = 0;
        sp.alm_done = 0;
        sp.alm_iters = 0;
        sp.newton_iters = 0;
    }
    __syncthreads();
    ScnCtx sc;
    load_scn(d, s, sc);
    build_csr(w, sc);
    if (w.as_mode && threadIdx.x == 0) sp.as_state = 1;  // inside the step
}

void launch_contacts_setup(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_contacts_setup, dim3(d.S), dim3(kBlk), 0, st, d, w);
}

// After the narrow phase at x_new (IPCSolver.cpp:1112-1165): lagged normal forces from
// the new contacts, contact residual, ALM break test.
__device__ void alm_post_scene(const Dev& d, const Work& w, SceneParams& sp, int s, double* red) {
    ScnCtx sc;
    load_scn(d, s, sc);
    const double* X = w.x;  // global body indexing
    double cres = 0.0;
    if (sp.needs_forms) {
        for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
            const double dd = contact_dist(w, sc.cb + i, X);
            w.nf[(long long)s * kMaxContacts + i] =
                (dd < sc.dhat && dd > 0.0) ? fabs(bar_g(dd, sc.dhat, sc.kappa)) : 0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0 && sp.has_friction) sp.nf_count = sc.nc;
        __syncthreads();
        load_scn(d, s, sc);
        build_csr(w, sc);
        // contact gradient without h^2 (barrier + friction)
        double* b2arr = w.sort_key + (long long)s * kSceneScratch;
        ScnCtx sb = sc;
        sb.j1 = sb.j0;  // no joints here
        sb.has_barrier = 1;  // barrier form exists whenever contact forms are needed
        eval_elements<false>(d, w, sb, X, b2arr);  // gradients only (the blocks are not read)
        const int* off = w.csr_off + sc.b0 + sc.s;
        const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
        double cm = 0.0;
        for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
            double gb[6] = {0, 0, 0, 0, 0, 0}, gf[6] = {0, 0, 0, 0, 0, 0};
            const int e0 = off[li], e1 = off[li + 1];
            int en = e0 < e1 ? ent[e0] : 0;  // the next entry's index, fetched a step ahead
            for (int e = e0; e < e1; ++e) {
                const int c = en >> 1, side = en & 1;
                en = e + 1 < e1 ? ent[e + 1] : 0;
                const double* cg = w.cgrad + 18 * ((long long)sc.s * kMaxContacts + c);
                for (int k = 0; k < 3; ++k) gb[k] += cg[side * 3 + k];
                for (int k = 0; k < 6; ++k) gf[k] += cg[6 + side * 6 + k];
            }
            for (int k = 0; k < 6; ++k) {
                const double v = (0.0 + gb[k]) + (sp.has_friction ? gf[k] : 0.0);
                cm = isnan(v) ? 1.0 / 0.0 : dmax(cm, fabs(v));
            }
        }
        cres = wg_reduce(cm, red, OpMax());
        if (!isfinite(cres)) cres = 1e12;
    }
    if (threadIdx.x == 0) {
        sp.cres = cres;
        sp.alm_iters += 1;
        const bool brk = sp.converged && sp.viol <= sp.alm_ctol && sp.dual <= sp.alm_dtol && cres <= sp.alm_ctcol;
        if (brk || sp.alm_iters >= sp.alm_max_iters) sp.alm_done = 1;
    }
}

// End of one ALM round of a host-driven step: whether every scene's ALM loop has ended
// (IPCSolver.cpp:1160-1165 break, or the iteration cap), written to host-visible memory
// as (round number since upload, all ended), the number stored last with release, so
// the host can stop issuing rounds once every scene is done. Run by the last work-group
// of k_alm_post to finish (an agent-scope counter, reset by that work-group): every
// scene's flags are final by then.
__device__ void round_report(const Dev& d, const Work& w, unsigned int* hflag) {
    __shared__ int last, all;
    __syncthreads();  // this work-group's writes are done
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(w.post_count, 1u) == gridDim.x - 1;
        all = 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    int mine = 1;
    for (int s = threadIdx.x; s < d.S; s += blockDim.x) {
        const SceneParams& sp = d.sp[s];
        // asynchronous stepping: every scene done, or in the ALM loop end of its last step
        mine &= w.as_mode ? (sp.as_state == 2 || (sp.as_state == 1 && sp.alm_done && sp.as_steps_left <= 1))
                          : sp.alm_done;
    }
    if (!mine) atomicAnd(&all, 0);
    __syncthreads();
    if (threadIdx.x == 0) {
        *w.post_count = 0u;  // for the next round (no work-group of this launch adds again)
        const unsigned int seq = w.round_seq[0] + 1u;
        w.round_seq[0] = seq;
        __hip_atomic_store(hflag + 1, (unsigned int)all, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(hflag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void __launch_bounds__(kBlk) k_alm_post(Dev d, Work w, int gather, unsigned int* hflag) {
    __shared__ double red[kBlk];
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    if (!sp.alm_done) {
        if (gather) {
            np_cap_gather(w, sp, s);
            __syncthreads();  // n_contacts and the kept contacts, for load_scn
        }
        alm_post_scene(d, w, sp, s, red);
    }
    if (hflag) round_report(d, w, hflag);
}

void launch_alm_post(const Dev& d, const Work& w, hipStream_t st, bool gather, unsigned int* hflag) {
    timing_begin("alm_post", st);
    hipLaunchKernelGGL(k_alm_post, dim3(d.S), dim3(kBlk), 0, st, d, w, gather ? 1 : 0, hflag);
    timing_end("alm_post", st);
}

// ------------------------------------------------------------------ step prologue
__global__ void k_step_begin(Dev d, Work w, const double* dt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.S) return;
    SceneParams& sp = d.sp[s];
    if (as_skip(w, sp)) return;
    sp.dt = dt[s];
    sp.dhat_cand = sp.dhat;  // candidates use the barrier's dhat before its first initialisation
    sp.min_pair_dist = 1.0 / 0.0;
    sp.ccd_any = 0;
    sp.evals = 0.0;
    sp.kappa_flag = 0;
    sp.toi = 1.0;
    sp.solve_fail = 0;
    sp.alm_done = 0;
    sp.vv_needed = 0;
    sp.ls_trials = 0;
    sp.solves_direct = 0;
    sp.solves_pcg = 0;
    sp.pcg_iters = 0;
    sp.comp_max = 0;
    w.flag[s] = 0;
    if (w.newton_helpers) {  // the step's helper protocol starts afresh (no helper runs yet)
        HelperBox& hb = w.hbox[s];
        hb.epoch = 0;
        hb.bseq = 0;
        hb.bopen = 0;
        hb.quit = 0;
        hb.dead = 0;
        hb.closed = 0;
        for (int h = 0; h < kHelpers; ++h) {
            hb.hep[h] = 0;
            hb.ack[h] = 0;
            hb.busy[h] = 0;
        }
        hb.wait_ticks = w.hcfg->wait_ticks;
        hb.epoch_ticks = w.hcfg->epoch_ticks;
        hb.idle_ticks = w.hcfg->idle_ticks;
        hb.straggle_ticks = w.hcfg->straggle_ticks;
    }
}

void launch_step_begin(const Dev& d, const Work& w, const double* dt, hipStream_t st) {
    hipLaunchKernelGGL(k_step_begin, dim3((d.S + 255) / 256), dim3(256), 0, st, d, w, dt);
}

// IPCSolver::initializeAdaptiveBarrier + AdaptiveBarrier::initialize (IPCSolver.cpp:
// 327-353, AdaptiveBarrier.h:27-31) once per scene, then the max-speed CCD inflation
// (IPCSolver.cpp:360-365).
__global__ void __launch_bounds__(kBlk) k_barrier_init(Dev d, Work w) {
    __shared__ double red[kBlk];
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    const int b0 = d.body_off[s], b1 = d.body_off[s + 1];
    if (sp.enable_adaptive && !sp.barrier_initialized && b1 > b0) {
        if (threadIdx.x == 0) {
            double tm = 0.0;
            const double inf = 1.0 / 0.0;
            V3 mn = v3(inf, inf, inf), mx = v3(-inf, -inf, -inf);
            for (int b = b0; b < b1; ++b) {
                tm += d.mass[b];
                mn = vmin(mn, ld3(d.p, b));
                mx = vmax(mx, ld3(d.p, b));
            }
            const double avg = tm / double(b1 - b0);
            double diag = norm(mx - mn);
            if (diag < 1e-6) diag = 1.0;
            sp.kappa_min = 1e11 * avg / (diag * diag);
            sp.kappa = sp.kappa_min;
            sp.dhat = 1e-3;
            sp.barrier_initialized = 1;
        }
    }
    __syncthreads();
    double vm = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) vm = dmax(vm, norm(ld3(d.v, b)));
    vm = wg_reduce(vm, red, OpMax());
    if (threadIdx.x == 0) {
        const double dh = sp.dhat;
        sp.ccd_infl = dmax(1e-6, dh) + vm * sp.dt;
    }
}

// AdaptiveBarrier::updateStiffness (AdaptiveBarrier.h:33-37) when the broad-phase gap is
// below 2 dhat and some vertex-face plane distance is below 1e-9.
__global__ void k_kappa_update(Dev d, Work w) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.S) return;
    SceneParams& sp = d.sp[s];
    if (as_skip(w, sp)) return;
    if (!sp.enable_adaptive) return;
    if (!(sp.min_pair_dist < sp.dhat * 2.0)) return;
    if (w.flag[s] && sp.kappa < 1e20) {
        sp.kappa = dmin(sp.kappa * 2.0, 1e20);
        sp.kappa_flag = 1;
    }
}

void launch_kappa(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_kappa_update, dim3((d.S + 255) / 256), dim3(256), 0, st, d, w);
}

// IPCSolver::computeMaxStep result (IPCSolver.cpp:377-391).
// One work-group per scene. The candidate counts are sums of integer products, exact in
// any order, so the work-group sums them in 64-bit integers (the reference's double
// sums of the same integers are exact below 2^53).
__global__ void __launch_bounds__(kBlk) k_ccd_dt(Dev d, Work w) {
    const int s = blockIdx.x;
    SceneParams& sp = d.sp[s];
    if (as_skip(w, sp)) return;
    long long fv = 0, ee = 0;
    for (int p = w.pair_off[s] + threadIdx.x; p < w.pair_off[s + 1]; p += blockDim.x) {
        const int* k = w.lcnt + 6 * p;
        fv += (long long)k[4] * k[1] + (long long)k[5] * k[0];
        ee += (long long)k[2] * k[3];
    }
    for (int o = 32; o > 0; o >>= 1) {
        fv += __shfl_xor(fv, o);
        ee += __shfl_xor(ee, o);
    }
    __shared__ long long part[2][kBlk / 64];
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = fv;
        part[1][threadIdx.x >> 6] = ee;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    fv = ee = 0;
    for (int q = 0; q < kBlk / 64; ++q) {
        fv += part[0][q];
        ee += part[1][q];
    }
    const double nfv = double(fv), nee = double(ee);
    sp.n_fv = nfv;
    sp.n_ee = nee;
    double adt = sp.dt;
    if (sp.enable_ccd && d.body_off[s + 1] > d.body_off[s] && sp.ccd_any && nfv > 0.0 && sp.toi < 1.0)
        adt = sp.toi * sp.ccd_safety * sp.dt;
    sp.adt = adt;
}

void launch_ccd_dt(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_ccd_dt, dim3(d.S), dim3(kBlk), 0, st, d, w);
}

// Pack x, v, a and the implicit-Euler prediction x_hat = x + h v
// (IPCSolver.cpp:839-861, ImplicitEulerIntegrator.cpp:5-7).
__global__ void k_pack(Dev d, Work w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.N) return;
    if (as_skip(w, d.sp[d.body_scene[b]])) return;
    const double h = d.sp[d.body_scene[b]].adt;
    const double xv[6] = {d.p[3 * b], d.p[3 * b + 1], d.p[3 * b + 2], 0.0, 0.0, 0.0};
    const double vv[6] = {d.v[3 * b], d.v[3 * b + 1], d.v[3 * b + 2], d.w[3 * b], d.w[3 * b + 1], d.w[3 * b + 2]};
    for (int k = 0; k < 6; ++k) {
        w.xprev[6 * b + k] = xv[k];
        w.vprev[6 * b + k] = vv[k];
        const double xh = xv[k] + h * vv[k];
        w.xhat[6 * b + k] = xh;
        w.x[6 * b + k] = xh;
    }
}

void launch_pack(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3((d.N + 255) / 256), dim3(256), 0, st, d, w);
}

// ImplicitEuler::update_quantities + pose write-back (IPCSolver.cpp:1168-1195).
__global__ void k_finalize(Dev d, Work w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.N) return;
    if (w.as_mode) {
        const SceneParams& sp = d.sp[d.body_scene[b]];
        if (!(sp.as_state == 1 && sp.alm_done)) return;
    }
    const double h = d.sp[d.body_scene[b]].adt;
    double vn[6], an[6];
    for (int k = 0; k < 6; ++k) {
        vn[k] = (w.x[6 * b + k] - w.xprev[6 * b + k]) / h;
        an[k] = (vn[k] - w.vprev[6 * b + k]) / h;
    }
    st3(d.al, b, v3(an[0], an[1], an[2]));
    st3(d.aa, b, v3(an[3], an[4], an[5]));
    st3(d.v, b, v3(vn[0], vn[1], vn[2]));
    st3(d.w, b, v3(vn[3], vn[4], vn[5]));
    st3(d.p, b, v3(w.x[6 * b], w.x[6 * b + 1], w.x[6 * b + 2]));
    const V3 th = v3(w.x[6 * b + 3], w.x[6 * b + 4], w.x[6 * b + 5]);
    st4(d.q, b, apply_theta(ld4(d.q, b), th));
}

void launch_finalize(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3((d.N + 255) / 256), dim3(256), 0, st, d, w);
}

__global__ void k_async_begin(Dev d, int n_steps) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.S) return;
    d.sp[s].as_state = 0;
    d.sp[s].as_steps_left = n_steps;
}

void launch_async_begin(const Dev& d, const Work& w, int n_steps, hipStream_t st) {
    (void)w;
    hipLaunchKernelGGL(k_async_begin, dim3((d.S + 255) / 256), dim3(256), 0, st, d, n_steps);
}

// After the write-back of a scene whose ALM loop ended: its next step, or done.
__global__ void k_async_state(Dev d) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.S) return;
    SceneParams& sp = d.sp[s];
    if (sp.as_state == 1 && sp.alm_done) {
        sp.as_steps_left -= 1;
        sp.as_state = sp.as_steps_left > 0 ? 0 : 2;
    }
}

void launch_async_advance(const Dev& d, const Work& w, hipStream_t st) {
    launch_finalize(d, w, st);
    hipLaunchKernelGGL(k_async_state, dim3((d.S + 255) / 256), dim3(256), 0, st, d);
}

void launch_barrier_init_ccd(const Dev& d, const Work& w, hipStream_t st) {
    hipLaunchKernelGGL(k_barrier_init, dim3(d.S), dim3(kBlk), 0, st, d, w);
}

// ------------------------------------------------------------------ form evaluation API
// Dense evaluation of DistanceBarrierForm / FrictionForm for given contacts (one scene),
// using the same per-contact element code as the solver.
__global__ void __launch_bounds__(kBlk) k_forms_eval(Dev d, Work w, int s, int which, double dhat, double kappa,
                                                     double mu, double eps, double dt, const double* x, double* value,
                                                     double* grad, double* hess, double* nfo) {
    __shared__ double red[kBlk];
    ScnCtx sc;
    load_scn(d, s, sc);
    sc.h2 = 1.0;
    sc.dhat = dhat > 1e-9 ? dhat : 1e-9;
    sc.kappa = kappa > 0.0 ? kappa : 0.0;
    sc.fmu = mu;
    sc.feps = eps;
    sc.fdt = dt > 0.0 ? dt : 0.0;
    sc.has_barrier = which == 0;
    sc.has_friction = which == 1;
    sc.nfc = sc.nc;
    sc.j1 = sc.j0;
    const double* X = x - 6 * (long long)sc.b0;  // contacts carry global body indices
    double* b2arr = w.sort_key + (long long)s * kSceneScratch;
    build_csr(w, sc);
    eval_elements(d, w, sc, X, b2arr);
    double v = 0.0;
    for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
        const long long c = sc.cb + i;
        if (which == 0) {
            const double dd = contact_dist(w, c, X);
            v += bar_v(dd, sc.dhat, sc.kappa);
            if (nfo) nfo[i] = (dd < sc.dhat && dd > 0.0) ? fabs(bar_g(dd, sc.dhat, sc.kappa)) : 0.0;
        } else {
            const double N = w.nf[(long long)s * kMaxContacts + i];
            if (N > 0.0) {
                FKin k;
                fric_kin(d, w, c, X, sc.fdt, k);
                v += fric_v(k.u, N, sc.fmu, sc.feps);
            }
        }
    }
    v = wg_reduce(v, red, OpAdd());
    if (threadIdx.x == 0) *value = v;
    const int* off = w.csr_off + sc.b0 + sc.s;
    const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    const int nd = sc.nd;
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        double g[6] = {0, 0, 0, 0, 0, 0};
        for (int e = off[li]; e < off[li + 1]; ++e) {
            const int c = ent[e] >> 1, side = ent[e] & 1;
            const double* cg = w.cgrad + 18 * ((long long)sc.s * kMaxContacts + c);
            if (which == 0) for (int k = 0; k < 3; ++k) g[k] += cg[side * 3 + k];
            else for (int k = 0; k < 6; ++k) g[k] += cg[6 + side * 6 + k];
        }
        for (int k = 0; k < 6; ++k) grad[6 * li + k] = g[k];
    }
    if (hess) {
        __syncthreads();
        // sum in contact order per entry (one lane: dense test path)
        if (threadIdx.x == 0) {
            for (long long i = 0; i < (long long)nd * nd; ++i) hess[i] = 0.0;
            for (int i = 0; i < sc.nc; ++i) {
                const long long cc = sc.cb + i;
            const long long ce = (long long)sc.s * kMaxContacts + i;
                const int act = w.cactive[ce];
                const int a = w.cint[5 * cc] - sc.b0, b = w.cint[5 * cc + 1] - sc.b0;
                if (which == 0 && (act & 1)) {
                    const double* cd = w.cdbl + 7 * cc;
                    const double n[3] = {cd[3], cd[4], cd[5]};
                    for (int r = 0; r < 3; ++r)
                        for (int q = 0; q < 3; ++q) {
                            const double blk = b2arr[i] * (n[r] * n[q]);
                            hess[(long long)(6 * a + r) * nd + 6 * a + q] += blk;
                            hess[(long long)(6 * b + r) * nd + 6 * b + q] += blk;
                            hess[(long long)(6 * a + r) * nd + 6 * b + q] += -blk;
                            hess[(long long)(6 * b + r) * nd + 6 * a + q] += -blk;
                        }
                }
                if (which == 1 && (act & 2)) {
                    const double* blk = w.cblk + 144 * ce;
                    for (int r = 0; r < 6; ++r)
                        for (int q = 0; q < 6; ++q) {
                            hess[(long long)(6 * a + r) * nd + 6 * a + q] += blk[r * 6 + q];
                            hess[(long long)(6 * b + r) * nd + 6 * b + q] += blk[36 + r * 6 + q];
                            hess[(long long)(6 * a + r) * nd + 6 * b + q] += blk[72 + r * 6 + q];
                            hess[(long long)(6 * b + r) * nd + 6 * a + q] += blk[108 + r * 6 + q];
                        }
                }
            }
        }
    }
}

void launch_forms_eval(const Dev& d, const Work& w, int s, int which, double dhat, double kappa, double mu,
                       double eps, double dt, const double* x, double* value, double* grad, double* hess, double* nf,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_forms_eval, dim3(1), dim3(kBlk), 0, st, d, w, s, which, dhat, kappa, mu, eps, dt, x, value,
                       grad, hess, nf);
}

// ------------------------------------------------------------------ standalone PCG
// ndipc_pcg_solve: block-Jacobi PCG over caller-provided 6x6 blocks (one work-group).
__global__ void __launch_bounds__(kBlk) k_pcg(int nb, const double* diag, int noff, const int* offij,
                                              const double* offb, const double* b, double tol, int max_it, double* x,
                                              int* iters, double* rel, double* scr) {
    __shared__ double red[kBlk];
    const int nd = 6 * nb;
    double* r = scr;
    double* z = scr + nd;
    double* p = scr + 2 * nd;
    double* Ap = scr + 3 * nd;
    double* Dinv = scr + 4 * nd;
    auto mv = [&](const double* v, double* y) {
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int c = 0; c < 6; ++c) s += diag[36 * i + a * 6 + c] * v[6 * i + c];
                for (int o = 0; o < noff; ++o) {
                    const int bi = offij[2 * o], bj = offij[2 * o + 1];
                    const double* B = offb + 36 * o;
                    if (bi == i) for (int c = 0; c < 6; ++c) s += B[a * 6 + c] * v[6 * bj + c];
                    if (bj == i) for (int c = 0; c < 6; ++c) s += B[c * 6 + a] * v[6 * bi + c];
                }
                y[6 * i + a] = s;
            }
        }
        __syncthreads();
    };
    auto prec = [&](const double* v, double* y) {
        for (int i = threadIdx.x; i < nb; i += blockDim.x)
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int c = 0; c < 6; ++c) s += Dinv[36 * i + a * 6 + c] * v[6 * i + c];
                y[6 * i + a] = s;
            }
        __syncthreads();
    };
    for (int i = threadIdx.x; i < nb; i += blockDim.x) inv6(diag + 36 * i, Dinv + 36 * i);
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) { x[i] = 0.0; r[i] = b[i]; }
    __syncthreads();
    prec(r, z);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) p[i] = z[i];
    __syncthreads();
    double rz = wg_dot(r, z, nd, red);
    const double bn = sqrt(wg_dot(b, b, nd, red));
    double rn = bn;
    int it = 0;
    for (; it < max_it && rn > tol * bn; ++it) {
        mv(p, Ap);
        const double pAp = wg_dot(p, Ap, nd, red);
        const double alpha = rz / pAp;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        __syncthreads();
        rn = sqrt(wg_dot(r, r, nd, red));
        prec(r, z);
        const double rz2 = wg_dot(r, z, nd, red);
        const double beta = rz2 / rz;
        rz = rz2;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) p[i] = z[i] + beta * p[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *iters = it;
        *rel = bn > 0.0 ? rn / bn : 0.0;
    }
}

// The same solve (the same values in the same order of operations: each block row's
// product sums its diagonal block, then its off-diagonal blocks in list order) with each
// row's blocks indexed once per solve (row_off, ent: (block << 1) | transposed) and a
// lane per row entry of the products, for systems with many off-diagonal blocks.
__global__ void __launch_bounds__(kBlk) k_pcg_rows(int nb, const double* diag, int noff, const int* offij,
                                                   const double* offb, const double* b, double tol, int max_it,
                                                   double* x, int* iters, double* rel, double* scr, int* iscr) {
    __shared__ double red[kBlk];
    const int nd = 6 * nb;
    double* r = scr;
    double* z = scr + nd;
    double* p = scr + 2 * nd;
    double* Ap = scr + 3 * nd;
    double* Dinv = scr + 4 * nd;
    int* row_off = iscr;
    int* ent = iscr + nb + 1;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        int c = 0;
        for (int o = 0; o < noff; ++o) c += (offij[2 * o] == i ? 1 : 0) + (offij[2 * o + 1] == i ? 1 : 0);
        row_off[i + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        row_off[0] = 0;
        for (int i = 0; i < nb; ++i) row_off[i + 1] += row_off[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        int e = row_off[i];
        for (int o = 0; o < noff; ++o) {
            if (offij[2 * o] == i) ent[e++] = o << 1;
            if (offij[2 * o + 1] == i) ent[e++] = (o << 1) | 1;
        }
    }
    auto mv = [&](const double* v, double* y) {
        for (int idx = threadIdx.x; idx < nd; idx += blockDim.x) {
            const int i = idx / 6, a = idx - 6 * i;
            double s = 0.0;
            for (int c = 0; c < 6; ++c) s += diag[36 * i + a * 6 + c] * v[6 * i + c];
            for (int e = row_off[i]; e < row_off[i + 1]; ++e) {
                const int o = ent[e] >> 1;
                const double* B = offb + 36 * o;
                if ((ent[e] & 1) == 0) {
                    const int bj = offij[2 * o + 1];
                    for (int c = 0; c < 6; ++c) s += B[a * 6 + c] * v[6 * bj + c];
                } else {
                    const int bi = offij[2 * o];
                    for (int c = 0; c < 6; ++c) s += B[c * 6 + a] * v[6 * bi + c];
                }
            }
            y[idx] = s;
        }
        __syncthreads();
    };
    auto prec = [&](const double* v, double* y) {
        for (int idx = threadIdx.x; idx < nd; idx += blockDim.x) {
            const int i = idx / 6, a = idx - 6 * i;
            double s = 0.0;
            for (int c = 0; c < 6; ++c) s += Dinv[36 * i + a * 6 + c] * v[6 * i + c];
            y[idx] = s;
        }
        __syncthreads();
    };
    for (int i = threadIdx.x; i < nb; i += blockDim.x) inv6(diag + 36 * i, Dinv + 36 * i);
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) { x[i] = 0.0; r[i] = b[i]; }
    __syncthreads();
    prec(r, z);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) p[i] = z[i];
    __syncthreads();
    double rz = wg_dot(r, z, nd, red);
    const double bn = sqrt(wg_dot(b, b, nd, red));
    double rn = bn;
    int it = 0;
    for (; it < max_it && rn > tol * bn; ++it) {
        mv(p, Ap);
        const double pAp = wg_dot(p, Ap, nd, red);
        const double alpha = rz / pAp;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) { x[i] += alpha * p[i]; r[i] -= alpha * Ap[i]; }
        __syncthreads();
        rn = sqrt(wg_dot(r, r, nd, red));
        prec(r, z);
        const double rz2 = wg_dot(r, z, nd, red);
        const double beta = rz2 / rz;
        rz = rz2;
        for (int i = threadIdx.x; i < nd; i += blockDim.x) p[i] = z[i] + beta * p[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *iters = it;
        *rel = bn > 0.0 ? rn / bn : 0.0;
    }
}

void launch_pcg_rows(int nb, const double* diag, int noff, const int* offij, const double* offb, const double* b,
                     double tol, int max_it, double* x, int* iters, double* rel, double* scratch, int* iscratch,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_pcg_rows, dim3(1), dim3(kBlk), 0, st, nb, diag, noff, offij, offb, b, tol, max_it, x, iters,
                       rel, scratch, iscratch);
}

void launch_pcg(int nb, const double* diag, int noff, const int* offij, const double* offb, const double* b,
                double tol, int max_it, double* x, int* iters, double* rel, double* scratch, hipStream_t st) {
    hipLaunchKernelGGL(k_pcg, dim3(1), dim3(kBlk), 0, st, nb, diag, noff, offij, offb, b, tol, max_it, x, iters, rel,
                       scratch);
}

// ------------------------------------------------------------------ adjoint
// Recording: before each Newton launch the objective of the solve (scenes whose ALM loop
// has finished mark the slot unused), after it the converged variable and the hardened
// joint stiffness the multiplier update used.
__global__ void __launch_bounds__(kBlk) k_adj_snapshot(Dev d, Work w, Adj a, int slot) {
    const int s = blockIdx.x;
    const SceneParams& sp = d.sp[s];
    const long long S = d.S;
    const int J = d.joint_off[d.S];
    if (sp.alm_done) {
        if (threadIdx.x == 0) a.s_sp[slot * S + s].alm_done = 1;
        return;
    }
    if (threadIdx.x == 0) a.s_sp[slot * S + s] = sp;
    const int nc = sp.n_contacts;
    const long long cb = (long long)s * kMaxRawContacts;
    const long long so = ((long long)slot * S + s) * kMaxContacts;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        for (int k = 0; k < 5; ++k) a.s_cint[5 * (so + i) + k] = w.cint[5 * (cb + i) + k];
        for (int k = 0; k < 7; ++k) a.s_cdbl[7 * (so + i) + k] = w.cdbl[7 * (cb + i) + k];
        for (int k = 0; k < 2; ++k) a.s_cext[2 * (so + i) + k] = w.cext[2 * (cb + i) + k];
        a.s_nf[so + i] = w.nf[(long long)s * kMaxContacts + i];
    }
    for (int j = d.joint_off[s] + threadIdx.x; j < d.joint_off[s + 1]; j += blockDim.x) {
        for (int k = 0; k < 6; ++k) a.s_jl[(long long)slot * 6 * J + 6 * j + k] = d.jlambda[6 * j + k];
        a.s_jm[(long long)slot * J + j] = d.jmu[j];
    }
}

__global__ void __launch_bounds__(kBlk) k_adj_snapshot_post(Dev d, Work w, Adj a, int slot) {
    const int s = blockIdx.x;
    const long long S = d.S;
    const int J = d.joint_off[d.S];
    if (a.s_sp[slot * S + s].alm_done) return;
    const int b0 = d.body_off[s], b1 = d.body_off[s + 1];
    if (threadIdx.x == 0)  // Newton iterations of this solve (the step's count before it + this + 1)
        a.s_nit[slot * S + s] = d.sp[s].newton_iters - a.s_sp[slot * S + s].newton_iters - 1;
    for (int i = 6 * b0 + threadIdx.x; i < 6 * b1; i += blockDim.x) a.s_x[(long long)slot * 6 * d.N + i] = w.x[i];
    for (int j = d.joint_off[s] + threadIdx.x; j < d.joint_off[s + 1]; j += blockDim.x)
        a.s_jm_post[(long long)slot * J + j] = d.jmu[j];
}

void launch_adj_snapshot(const Dev& d, const Work& w, const Adj& a, int slot, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_snapshot, dim3(d.S), dim3(kBlk), 0, st, d, w, a, slot);
}
void launch_adj_snapshot_post(const Dev& d, const Work& w, const Adj& a, int slot, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_snapshot_post, dim3(d.S), dim3(kBlk), 0, st, d, w, a, slot);
}

// Diagonal blocks of the parts of A = dg/dx that equal the assembled Hessian: inertia,
// fixed joints (C is linear in x) and the barrier (its b'' n n^T; zero where the
// distance sits at its clamp, see k_adj_step). Hinge and friction elements enter A
// through their own exact Jacobians instead.
__device__ void assemble_diag_base(const Dev& d, const Work& w, const Adj& a, const ScnCtx& sc,
                                   const double* b2arr) {
    const int* off = w.csr_off + sc.b0 + sc.s;
    const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const int b = sc.b0 + li;
        double D[36];
        for (int k = 0; k < 36; ++k) D[k] = 0.0;
        const double m = d.mass[b];
        D[0] = m; D[7] = m; D[14] = m;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) D[(3 + r) * 6 + 3 + c] = d.inertia[9 * b + 3 * r + c];
        for (int j = sc.j0; j < sc.j1; ++j) {
            if (d.jtype[j] != 0 || d.ja[j] != b) continue;
            const double* jb = w.jblk + 108 * j;
            for (int k = 0; k < 36; ++k) D[k] += jb[k];
        }
        for (int e = off[li]; e < off[li + 1]; ++e) {
            const int c = ent[e] >> 1;
            if (!(w.cactive[(long long)sc.s * kMaxContacts + c] & 1)) continue;
            const double* cd = w.cdbl + 7 * (sc.cb + c);
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) D[r * 6 + q] += (b2arr[c] * (cd[3 + r] * cd[3 + q])) * sc.h2;
        }
        for (int k = 0; k < 36; ++k) a.dbase[36 * (long long)b + k] = D[k];
    }
}

// y = A^T v for the scene, A = dg/dx the exact Jacobian of the Newton gradient at the
// converged point: the base diagonal blocks, the barrier's off-diagonal blocks and the
// exact (transposed) Jacobians of the hinge and friction elements. With ya != nullptr
// also ya = |A^T| |v| (entrywise), the scale of the componentwise backward error.
__device__ void adj_matvec_t(const Dev& d, const Work& w, const Adj& a, const ScnCtx& sc, const double* b2arr,
                             const double* v, double* y, double* ya) {
    const double* vg = v - 6 * (long long)sc.b0;  // global-indexed view
    for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
        const long long ce = (long long)sc.s * kMaxContacts + i;
        if (!(w.cactive[ce] & 2)) continue;
        const long long cc = sc.cb + i;
        const int A = w.cint[5 * cc], B = w.cint[5 * cc + 1];
        const double* E = a.ejac_c + 144 * ce;
        double e[12];
        for (int k = 0; k < 6; ++k) {
            e[k] = vg[6 * A + k];
            e[6 + k] = vg[6 * B + k];
        }
        for (int r = 0; r < 12; ++r) {
            double acc = 0.0, aa = 0.0;
            for (int c = 0; c < 12; ++c) {
                acc += E[12 * r + c] * e[c];
                aa += fabs(E[12 * r + c]) * fabs(e[c]);
            }
            a.ye_c[12 * ce + r] = acc;
            a.yea_c[12 * ce + r] = aa;
        }
    }
    for (int j = sc.j0 + threadIdx.x; j < sc.j1; j += blockDim.x) {
        if (d.jtype[j] != 1) continue;
        const double* E = a.ejac_j + 144 * (long long)j;
        double e[12];
        for (int k = 0; k < 6; ++k) {
            e[k] = vg[6 * d.ja[j] + k];
            e[6 + k] = vg[6 * d.jb[j] + k];
        }
        for (int r = 0; r < 12; ++r) {
            double acc = 0.0, aa = 0.0;
            for (int c = 0; c < 12; ++c) {
                acc += E[12 * r + c] * e[c];
                aa += fabs(E[12 * r + c]) * fabs(e[c]);
            }
            a.ye_j[12 * (long long)j + r] = acc;
            a.yea_j[12 * (long long)j + r] = aa;
        }
    }
    __syncthreads();
    const int* off = w.csr_off + sc.b0 + sc.s;
    const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        const int b = sc.b0 + li;
        const double* D = a.dbase + 36 * (long long)b;
        double acc[6], aa[6];
        for (int r = 0; r < 6; ++r) {
            double s_ = 0.0, t_ = 0.0;
            for (int c = 0; c < 6; ++c) {
                s_ += D[r * 6 + c] * v[6 * li + c];
                t_ += fabs(D[r * 6 + c]) * fabs(v[6 * li + c]);
            }
            acc[r] = s_;
            aa[r] = t_;
        }
        for (int e = off[li]; e < off[li + 1]; ++e) {
            const int c = ent[e] >> 1, side = ent[e] & 1;
            const long long cc = sc.cb + c;
            const long long ce = (long long)sc.s * kMaxContacts + c;
            const int act = w.cactive[ce];
            if (act & 1) {
                const int other = (side == 0 ? w.cint[5 * cc + 1] : w.cint[5 * cc]) - sc.b0;
                const double* cd = w.cdbl + 7 * cc;
                for (int r = 0; r < 3; ++r)
                    for (int q = 0; q < 3; ++q) {
                        const double hv = (-(b2arr[c] * (cd[3 + r] * cd[3 + q]))) * sc.h2;
                        acc[r] += hv * v[6 * other + q];
                        aa[r] += fabs(hv) * fabs(v[6 * other + q]);
                    }
            }
            if (act & 2)
                for (int k = 0; k < 6; ++k) {
                    acc[k] += a.ye_c[12 * ce + 6 * side + k];
                    aa[k] += a.yea_c[12 * ce + 6 * side + k];
                }
        }
        for (int j = sc.j0; j < sc.j1; ++j) {
            if (d.jtype[j] != 1) continue;
            const int side = d.ja[j] == b ? 0 : (d.jb[j] == b ? 1 : -1);
            if (side < 0) continue;
            for (int k = 0; k < 6; ++k) {
                acc[k] += a.ye_j[12 * (long long)j + 6 * side + k];
                aa[k] += a.yea_j[12 * (long long)j + 6 * side + k];
            }
        }
        for (int r = 0; r < 6; ++r) {
            y[6 * li + r] = acc[r];
            if (ya) ya[6 * li + r] = aa[r];
        }
    }
    __syncthreads();
}

// Componentwise backward error of eta as a solution of A^T eta = xbar:
// max_i |xbar - A^T eta|_i / (|A^T| |eta| + |xbar|)_i (Oettli-Prager). r, ra: scratch.
__device__ double adj_backward_error(const Dev& d, const Work& w, const Adj& a, const ScnCtx& sc,
                                     const double* b2arr, const double* xb, const double* eta, double* r, double* ra,
                                     double* red) {
    adj_matvec_t(d, w, a, sc, b2arr, eta, r, ra);
    double om = 0.0;
    for (int i = threadIdx.x; i < sc.nd; i += blockDim.x) {
        const double den = ra[i] + fabs(xb[i]);
        const double num = fabs(xb[i] - r[i]);
        const double q = den > 0.0 ? num / den : (num > 0.0 ? 1.0 / 0.0 : 0.0);
        om = isnan(q) ? 1.0 / 0.0 : dmax(om, q);
    }
    return wg_reduce(om, red, OpMax());
}

// Acceptance bound of a transposed solve: componentwise backward error (Oettli-Prager).
// A normwise relative residual is no test here: fixed joints (mu = 1e8 and hardened up
// to 1e12) and stiction blocks make A's rows differ in scale by up to 1e20.
constexpr double kAdjBackwardTol = 1e-12;

// One ALM solve of the reverse sweep: eta = A^-T xbar, A the exact Jacobian of the
// Newton gradient at the converged point, by restarted GMRES with right preconditioning
// by the assembled Hessian's direct solve (A and H differ in the second-order hinge and
// friction terms only, so few iterations are needed). The fallback of the direct solve
// for components above kDenseBodies bodies. Returns whether the componentwise backward
// error (in *rres; -1 if the preconditioner failed) is within kAdjBackwardTol.
constexpr int kGmresM = 24;
__device__ bool adj_solve_gmres(const Dev& d, const Work& w, const Adj& a, const ScnCtx& sc, const double* b2arr,
                                double* red, int* ok_sh, NewtonLds& nl, double* rres) {
    __shared__ double Hh[(kGmresM + 1) * kGmresM], cs[kGmresM], sn[kGmresM], gg[kGmresM + 1], yy[kGmresM];
    const long long o = 6 * (long long)sc.b0;
    const long long stride = 6 * (long long)d.N;
    const double* xb = a.xbar + o;
    double* eta = a.eta + o;
    double* r = a.res + o;
    double* z = a.tmp + o;
    auto V = [&](int j) { return a.kry + j * stride + o; };
    const int nd = sc.nd;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) eta[i] = 0.0;
    __syncthreads();
    const double bn = sqrt(wg_dot(xb, xb, nd, red));
    *rres = 0.0;
    if (!(bn > 0.0)) return true;
    const double tol = 1e-13 * bn;
    auto precond = [&](const double* in, double* out) -> bool {
        const int c = component_solve<4>(d, w, sc, b2arr, in, out, nl.comp, ok_sh);
        const bool ok = c == 0 ? true : c == 2 ? pcg_solve(d, w, sc, b2arr, in, out, red, ok_sh) : false;
        __syncthreads();
        return ok;
    };
    double rn = bn;
    for (int restart = 0; restart < 8; ++restart) {
        adj_matvec_t(d, w, a, sc, b2arr, eta, z, nullptr);
        for (int i = threadIdx.x; i < nd; i += blockDim.x) r[i] = xb[i] - z[i];
        __syncthreads();
        const double beta = sqrt(wg_dot(r, r, nd, red));
        rn = beta;
        *rres = rn / bn;
        if (beta <= tol) {
            const double om = adj_backward_error(d, w, a, sc, b2arr, xb, eta, r, V(1), red);
            *rres = om;
            return om <= kAdjBackwardTol;
        }
        double* v0 = V(0);
        for (int i = threadIdx.x; i < nd; i += blockDim.x) v0[i] = r[i] / beta;
        if (threadIdx.x == 0) {
            gg[0] = beta;
            for (int k = 1; k <= kGmresM; ++k) gg[k] = 0.0;
        }
        __syncthreads();
        int jend = 0;
        for (int j = 0; j < kGmresM; ++j) {
            if (!precond(V(j), z)) {
                *rres = -1.0;
                return false;
            }
            double* wv = V(j + 1);
            adj_matvec_t(d, w, a, sc, b2arr, z, wv, nullptr);
            for (int i = 0; i <= j; ++i) {
                const double hij = wg_dot(wv, V(i), nd, red);
                const double* vi = V(i);
                for (int t = threadIdx.x; t < nd; t += blockDim.x) wv[t] -= hij * vi[t];
                if (threadIdx.x == 0) Hh[i * kGmresM + j] = hij;
                __syncthreads();
            }
            const double hn = sqrt(wg_dot(wv, wv, nd, red));
            if (threadIdx.x == 0) Hh[(j + 1) * kGmresM + j] = hn;
            if (hn > 0.0)
                for (int t = threadIdx.x; t < nd; t += blockDim.x) wv[t] /= hn;
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int i = 0; i < j; ++i) {  // previous rotations
                    const double h0 = Hh[i * kGmresM + j], h1 = Hh[(i + 1) * kGmresM + j];
                    Hh[i * kGmresM + j] = cs[i] * h0 + sn[i] * h1;
                    Hh[(i + 1) * kGmresM + j] = -sn[i] * h0 + cs[i] * h1;
                }
                const double h0 = Hh[j * kGmresM + j], h1 = Hh[(j + 1) * kGmresM + j];
                const double den = sqrt(h0 * h0 + h1 * h1);
                cs[j] = den > 0.0 ? h0 / den : 1.0;
                sn[j] = den > 0.0 ? h1 / den : 0.0;
                Hh[j * kGmresM + j] = den;
                Hh[(j + 1) * kGmresM + j] = 0.0;
                gg[j + 1] = -sn[j] * gg[j];
                gg[j] = cs[j] * gg[j];
            }
            __syncthreads();
            jend = j + 1;
            if (fabs(gg[j + 1]) <= tol || !(hn > 0.0)) break;
        }
        if (threadIdx.x == 0) {
            for (int i = jend - 1; i >= 0; --i) {
                double s_ = gg[i];
                for (int k = i + 1; k < jend; ++k) s_ -= Hh[i * kGmresM + k] * yy[k];
                yy[i] = s_ / Hh[i * kGmresM + i];
            }
        }
        __syncthreads();
        // eta += H^-1 (V y)
        for (int t = threadIdx.x; t < nd; t += blockDim.x) {
            double acc = 0.0;
            for (int k = 0; k < jend; ++k) acc += yy[k] * V(k)[t];
            r[t] = acc;
        }
        __syncthreads();
        if (!precond(r, z)) {
            *rres = -1.0;
            return false;
        }
        for (int t = threadIdx.x; t < nd; t += blockDim.x) eta[t] += z[t];
        __syncthreads();
    }
    const double om = adj_backward_error(d, w, a, sc, b2arr, xb, eta, r, V(1), red);
    *rres = om;
    return om <= kAdjBackwardTol;
}

// One ALM solve of the reverse sweep by a direct method. A^T is block diagonal over the
// coupled components of H (hinges and active contacts couple bodies), so each
// component's A^T is assembled densely (base diagonal blocks, then the barrier's
// off-diagonal blocks and the exact hinge / friction Jacobians in element order) and
// factored by LU without pivoting (right-looking, the trailing update in parallel; A
// differs from the symmetric positive definite H only in second-order hinge and friction
// terms, so the pivots stay those of a positive definite matrix). Two refinement steps
// against the unfactored copy follow. Singletons are their base 6x6 blocks. Returns 0
// when solved; 1 when a component exceeds kDenseBodies, 2 / 3 when a singleton /
// component pivot vanishes, 4 when the componentwise backward error (*rres) exceeds
// kAdjBackwardTol. The caller then runs GMRES.
__device__ int adj_solve_direct(const Dev& d, const Work& w, const Adj& a, const ScnCtx& sc, const double* b2arr,
                                double* red, int* ok_sh, CompLds& L, double* rres) {
    const long long o = 6 * (long long)sc.b0;
    const double* xb = a.xbar + o;
    double* eta = a.eta + o;
    double* xs = a.res + o;  // component right-hand side, then its residual
    double* ys = a.tmp + o;  // component solution of one triangular pass
    const int nd = sc.nd;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) eta[i] = 0.0;
    __syncthreads();
    const double bn = sqrt(wg_dot(xb, xb, nd, red));
    *rres = 0.0;
    if (!(bn > 0.0)) return 0;
    build_components(d, w, sc, L);
    if (L.maxsize > kDenseBodies) return 1;
    if (threadIdx.x == 0) *ok_sh = 1;
    __syncthreads();
    for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
        if (L.cnt[L.lab[li]] > 1) continue;
        if (!ldlt6_solve(a.dbase + 36 * (sc.b0 + li), xb + 6 * li, eta + 6 * li)) atomicAnd(ok_sh, 0);
    }
    __syncthreads();
    if (!*ok_sh) return 2;
    double* Mo = a.dense + (long long)sc.s * 2 * kDenseRows * kDenseRows;
    double* Mf = Mo + kDenseRows * kDenseRows;
    for (int c0 = 0; c0 < L.ncoupled;) {
        const int c = L.lab[L.list[c0]];
        const int mb = L.cnt[c];
        const int c1 = c0 + mb;
        const int m = 6 * mb;
        auto lpos = [&](int lb) {
            int lo = c0, hi = c1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (L.list[mid] < lb) lo = mid + 1;
                else hi = mid;
            }
            return lo - c0;
        };
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) Mo[t] = 0.0;
        __syncthreads();
        for (int t = threadIdx.x; t < 36 * mb; t += blockDim.x) {
            const int bi = t / 36, e = t % 36;
            Mo[(6 * bi + e / 6) * m + 6 * bi + e % 6] = a.dbase[36 * (sc.b0 + L.list[c0 + bi]) + e];
        }
        __syncthreads();
        // one element at a time, its 12x12 contribution to A^T (rows/cols [A | B]) by 144 lanes
        auto add_element = [&](int A, int B, auto&& val) {
            const int pa = lpos(A), pb = lpos(B);
            if (threadIdx.x < 144) {
                const int r = threadIdx.x / 12, q = threadIdx.x % 12;
                const int pr = r < 6 ? pa : pb, pq = q < 6 ? pa : pb;
                Mo[(6 * pr + r % 6) * m + 6 * pq + q % 6] += val(r, q);
            }
            __syncthreads();
        };
        for (int j = sc.j0; j < sc.j1; ++j) {
            if (d.jtype[j] != 1) continue;
            const int A = d.ja[j] - sc.b0, B = d.jb[j] - sc.b0;
            if (L.lab[A] != c) continue;
            const double* E = a.ejac_j + 144 * (long long)j;
            add_element(A, B, [&](int r, int q) { return E[12 * r + q]; });
        }
        for (int i = 0; i < sc.nc; ++i) {
            const long long cc = sc.cb + i, ce = (long long)sc.s * kMaxContacts + i;
            const int act = w.cactive[ce];
            if (!act) continue;
            const int A = w.cint[5 * cc] - sc.b0, B = w.cint[5 * cc + 1] - sc.b0;
            if (L.lab[A] != c) continue;
            const double* cd = w.cdbl + 7 * cc;
            const double* E = a.ejac_c + 144 * ce;
            const double bb = b2arr[i];
            add_element(A, B, [&](int r, int q) {
                double v = (act & 2) ? E[12 * r + q] : 0.0;
                const bool off = (r < 6) != (q < 6);
                if (off && (act & 1) && r % 6 < 3 && q % 6 < 3)
                    v += (-(bb * (cd[3 + r % 6] * cd[3 + q % 6]))) * sc.h2;
                return v;
            });
        }
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) Mf[t] = Mo[t];
        __syncthreads();
        for (int k = 0; k < m; ++k) {
            const double piv = Mf[k * m + k];
            if (piv == 0.0 || !isfinite(piv)) return 3;
            for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) Mf[i * m + k] /= piv;
            __syncthreads();
            const int r = m - k - 1;
            for (int t = threadIdx.x; t < r * r; t += blockDim.x) {
                const int i = k + 1 + t / r, j = k + 1 + t % r;
                Mf[i * m + j] -= Mf[i * m + k] * Mf[k * m + j];
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < m; t += blockDim.x) {
            xs[t] = xb[6 * L.list[c0 + t / 6] + t % 6];
            ys[t] = 0.0;
        }
        __syncthreads();
        for (int pass = 0; pass < 3; ++pass) {
            if (pass > 0) {  // residual of the component against the unfactored copy
                for (int i = threadIdx.x; i < m; i += blockDim.x) {
                    double acc = xb[6 * L.list[c0 + i / 6] + i % 6];
                    for (int j = 0; j < m; ++j) acc -= Mo[i * m + j] * ys[j];
                    xs[i] = acc;
                }
                __syncthreads();
            }
            for (int k = 0; k < m; ++k) {  // unit lower
                const double xk = xs[k];
                for (int i = k + 1 + threadIdx.x; i < m; i += blockDim.x) xs[i] -= Mf[i * m + k] * xk;
                __syncthreads();
            }
            for (int k = m - 1; k >= 0; --k) {  // upper, solution added to ys
                const double xk = xs[k] / Mf[k * m + k];
                for (int i = threadIdx.x; i < k; i += blockDim.x) xs[i] -= Mf[i * m + k] * xk;
                if (threadIdx.x == 0) ys[k] += xk;
                __syncthreads();
            }
        }
        for (int t = threadIdx.x; t < m; t += blockDim.x) eta[6 * L.list[c0 + t / 6] + t % 6] = ys[t];
        __syncthreads();
        c0 = c1;
    }
    const double om = adj_backward_error(d, w, a, sc, b2arr, xb, eta, xs, ys, red);
    *rres = om;
    return om <= kAdjBackwardTol ? 0 : 4;
}

// Reverse of step `step` of the recorded segment, every scene. ALM iteration a solved
// g_a(x; lambda_{a-1}, z) = 0 (z: step-start pose and velocity, masses, friction
// coefficient) and then set lambda_a = lambda_{a-1} + mu_a C(x*_a). Going back over the
// iterations: xbar_a gets the write-back's cotangent (last iteration) and the multiplier
// update's mu_a (dC/dx)^T lbar; eta = A_a^-T xbar_a; z and lambda_{a-1} receive -eta^T
// dg_a/dz and -eta^T dg_a/dlambda. The contact set, lagged normal forces, barrier
// stiffness and the hardening decisions are the solves' data and are held fixed.
// b'(d) over dual numbers: its derivative is b''(d) d'.
template <int NV>
__device__ du::Dn<NV> bar_g_dn(const du::Dn<NV>& dd, double dh, double k) {
    du::Dn<NV> r;
    r.v = bar_g(dd.v, dh, k);
    const double h = bar_h(dd.v, dh, k);
    for (int i = 0; i < NV; ++i) r.d[i] = h * dd.d[i];
    return r;
}

// Pose of body b as dual numbers seeded at slots O.. : at state xp (x_p at O..O+2, theta at
// O+3..O+5; q = q_step_start updated by theta) or, with xp == nullptr, the step-start pose
// (p at O..O+2, q at O+3..O+6). The slots are compile-time constants.
template <int O>
__device__ __forceinline__ void adj_geom_seed(const Dev& d, const double* xp, int b, du::V3T<du::Dn<14>>& P,
                                              du::QT<du::Dn<14>>& Qr) {
    if (xp) {
        P = du::mk3(du::var<14>(xp[6 * b], O), du::var<14>(xp[6 * b + 1], O + 1), du::var<14>(xp[6 * b + 2], O + 2));
        const du::V3T<du::Dn<14>> th = du::mk3(du::var<14>(xp[6 * b + 3], O + 3), du::var<14>(xp[6 * b + 4], O + 4),
                                               du::var<14>(xp[6 * b + 5], O + 5));
        Qr = du::rot_update(du::liftq<du::Dn<14>>(d.q + 4 * b), th, true);
    } else {
        P = du::mk3(du::var<14>(d.p[3 * b], O), du::var<14>(d.p[3 * b + 1], O + 1), du::var<14>(d.p[3 * b + 2], O + 2));
        Qr = du::mkq(du::var<14>(d.q[4 * b], O + 3), du::var<14>(d.q[4 * b + 1], O + 4),
                     du::var<14>(d.q[4 * b + 2], O + 5), du::var<14>(d.q[4 * b + 3], O + 6));
    }
}

// Cotangent of a solve's objective through the geometry of its contact i. The contact's
// point and normal, and in a solve after the first its lagged normal force, are functions
// of the state the narrow phase ran at (the contact set itself is discrete and held
// fixed): for a solve after the first the previous solve's x* (xp; 12 variables, the
// Newton variables of A then B: out[7 side + k], k < 6), for the first the step-start
// poses (xp == nullptr; 14 variables, p and q of A then B: out[7 side + k]). out = -h^2
// d(eta_A . g_A + eta_B . g_B)/d(variables), g the contact's barrier and friction gradient
// at the solve's x* (w.x).
__device__ void adj_contact_geom(const Dev& d, const Work& w, const ScnCtx& sc, int i, const double* xp,
                                 const double* ETA, int b0, double* out) {
    typedef du::Dn<14> D;
    const long long cc = sc.cb + i;
    const long long ce = (long long)sc.s * kMaxContacts + i;
    const int* ci = w.cint + 5 * cc;
    const int A = ci[0], B = ci[1], type = ci[2], pa = ci[3], pb = ci[4];
    const int pa2 = w.cext[2 * cc], pb2 = w.cext[2 * cc + 1];
    const int bod[2] = {A, B};
    du::V3T<D> P[2];
    du::QT<D> Qr[2];
    adj_geom_seed<0>(d, xp, A, P[0], Qr[0]);
    adj_geom_seed<7>(d, xp, B, P[1], Qr[1]);
    auto wv = [&](int side, int lv) {
        const double* vl = d.verts + 3 * (d.v_off[bod[side]] + lv);
        return du::add(P[side], du::rotate(Qr[side], du::lift3<D>(vl[0], vl[1], vl[2])));
    };
    du::V3T<D> v[4];
    int kind = 0;
    if (type == 1) {
        kind = 1;
        v[0] = wv(0, pa); v[1] = wv(0, pa2); v[2] = wv(1, pb); v[3] = wv(1, pb2);
    } else if (pa2 == -2) {
        kind = 2;
        v[0] = wv(0, pa); v[1] = wv(1, pb);
    } else {
        const int* f = d.faces + 3 * (d.f_off[A] + pa);
        v[0] = wv(0, f[0]); v[1] = wv(0, f[1]); v[2] = wv(0, f[2]); v[3] = wv(1, pb);
    }
    du::V3T<D> pt, nr;
    du::contact_geom_t<D>(kind, v, P[0], P[1], pt, nr);
    const double* x = w.x;
    const int act = w.cactive[ce];
    D g[12];
    for (int k = 0; k < 12; ++k) g[k] = du::cst<14>(0.0);
    if (act & 1) {  // barrier: -b'(d) n on A, b'(d) n on B, d = max(1e-9, (pB - pA).n) at x*
        const D pr = (x[6 * B] - x[6 * A]) * nr.x + (x[6 * B + 1] - x[6 * A + 1]) * nr.y +
                     (x[6 * B + 2] - x[6 * A + 2]) * nr.z;
        const D dd = pr.v > 1e-9 ? pr : du::cst<14>(1e-9);
        const D b1 = bar_g_dn(dd, sc.dhat, sc.kappa);
        g[0] = -(b1 * nr.x); g[1] = -(b1 * nr.y); g[2] = -(b1 * nr.z);
        g[6] = b1 * nr.x; g[7] = b1 * nr.y; g[8] = b1 * nr.z;
    }
    if (act & 2) {  // friction, N = |b'(d)| at the narrow phase's state (computeNormalForces)
        D N = du::cst<14>(w.nf[ce]);
        if (xp) {
            const du::V3T<D> ab = du::sub(P[1], P[0]);
            const D pr = du::dot(ab, nr);
            const D dd = pr.v > 1e-9 ? pr : du::cst<14>(1e-9);
            N = (dd.v < sc.dhat && dd.v > 0.0) ? du::dfabs(bar_g_dn(dd, sc.dhat, sc.kappa)) : du::cst<14>(0.0);
        }
        D xa[6], xb[6], gf[12];
        for (int k = 0; k < 6; ++k) {
            xa[k] = du::cst<14>(x[6 * A + k]);
            xb[k] = du::cst<14>(x[6 * B + k]);
        }
        du::friction_grad_g<D>(xa, xb, du::lift3<D>(d.p[3 * A], d.p[3 * A + 1], d.p[3 * A + 2]),
                               du::liftq<D>(d.q + 4 * A), du::lift3<D>(d.p[3 * B], d.p[3 * B + 1], d.p[3 * B + 2]),
                               du::liftq<D>(d.q + 4 * B), pt, nr, N, du::cst<14>(sc.fmu), sc.feps, sc.fdt, gf);
        for (int k = 0; k < 12; ++k) g[k] = g[k] + gf[k];
    }
    const double* lA = ETA + 6 * (A - b0);
    const double* lB = ETA + 6 * (B - b0);
    for (int k = 0; k < 14; ++k) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += lA[j] * g[j].d[k] + lB[j] * g[6 + j].d[k];
        out[k] = -(acc * sc.h2);
    }
}

__global__ void __launch_bounds__(kBlk) k_adj_step(Dev d, Work w, Adj a, int step, int gstep) {
    __shared__ double red[kBlk];
    __shared__ int ok_sh;
    __shared__ NewtonLds nl;
    const int s = blockIdx.x;
    const long long S = d.S;
    const int N = d.N, J = d.joint_off[d.S];
    const int b0 = d.body_off[s], b1 = d.body_off[s + 1];
    const double* stt = a.s_state + (long long)step * 13 * N;
    const double* sp_ = stt;
    const double* sq_ = stt + 3 * (long long)N;
    const double* sv_ = stt + 7 * (long long)N;
    int alast = -1;
    for (int k = a.A - 1; k >= 0; --k)
        if (!a.s_sp[((long long)step * a.A + k) * S + s].alm_done) { alast = k; break; }
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        for (int k = 0; k < 3; ++k) d.p[3 * b + k] = sp_[3 * b + k];
        for (int k = 0; k < 4; ++k) d.q[4 * b + k] = sq_[4 * b + k];
    }
    if (alast < 0) return;  // nothing recorded for this scene
    const double* xfin = a.s_x + ((long long)step * a.A + alast) * 6 * N;
    const double h = a.s_sp[((long long)step * a.A + alast) * S + s].adt;
    // write-back: p' = x_p, v' = (x_p - p)/h, w' = x_th/h, q' = apply_theta(q, x_th)
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) {
        du::Dn<7> th[3], qq[4];
        for (int k = 0; k < 3; ++k) th[k] = du::var<7>(xfin[6 * b + 3 + k], k);
        for (int k = 0; k < 4; ++k) qq[k] = du::var<7>(sq_[4 * b + k], 3 + k);
        const du::QT<du::Dn<7>> qn =
            du::rot_update(du::mkq(qq[0], qq[1], qq[2], qq[3]), du::mk3(th[0], th[1], th[2]), true);
        const du::Dn<7> qo[4] = {qn.w, qn.x, qn.y, qn.z};
        double c[7];
        for (int k = 0; k < 7; ++k) {
            double acc = 0.0;
            for (int r = 0; r < 4; ++r) acc += a.aq[4 * b + r] * qo[r].d[k];
            c[k] = acc;
        }
        for (int k = 0; k < 3; ++k) {
            a.bp[3 * b + k] = -a.av[3 * b + k] / h;
            a.bv[3 * b + k] = 0.0;
            a.bw[3 * b + k] = 0.0;
            a.tmp[6 * b + k] = a.ap[3 * b + k] + a.av[3 * b + k] / h;  // final xbar (kept in tmp)
            a.tmp[6 * b + 3 + k] = a.aw[3 * b + k] / h + c[k];
        }
        for (int k = 0; k < 4; ++k) a.bq[4 * b + k] = c[3 + k];
    }
    __syncthreads();
    for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
        for (int k = 0; k < 6; ++k) a.res[6 * b + k] = a.tmp[6 * b + k];  // res holds the final xbar
    __syncthreads();
    for (int it = alast; it >= 0; --it) {
        const long long slot = (long long)step * a.A + it;
        if (a.s_sp[slot * S + s].alm_done) continue;
        const double* sx = a.s_x + slot * 6 * N;
        // restore the objective of this solve
        if (threadIdx.x == 0) d.sp[s] = a.s_sp[slot * S + s];
        __syncthreads();
        SceneParams& sp = d.sp[s];
        for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
            for (int k = 0; k < 6; ++k) {
                w.x[6 * b + k] = sx[6 * b + k];
                const double cv = a.carry[6 * b + k];  // from the next solve, if it took no step
                a.carry[6 * b + k] = 0.0;
                a.xbar[6 * b + k] = it == alast ? a.res[6 * b + k] : cv;
            }
        {
            const int nc = sp.n_contacts;
            const long long cb = (long long)s * kMaxRawContacts;
            const long long so = (slot * S + s) * kMaxContacts;
            for (int i = threadIdx.x; i < nc; i += blockDim.x) {
                for (int k = 0; k < 5; ++k) w.cint[5 * (cb + i) + k] = a.s_cint[5 * (so + i) + k];
                for (int k = 0; k < 7; ++k) w.cdbl[7 * (cb + i) + k] = a.s_cdbl[7 * (so + i) + k];
                for (int k = 0; k < 2; ++k) w.cext[2 * (cb + i) + k] = a.s_cext[2 * (so + i) + k];
                w.nf[(long long)s * kMaxContacts + i] = a.s_nf[so + i];
            }
            for (int j = d.joint_off[s] + threadIdx.x; j < d.joint_off[s + 1]; j += blockDim.x) {
                for (int k = 0; k < 6; ++k) d.jlambda[6 * j + k] = a.s_jl[slot * 6 * J + 6 * j + k];
                d.jmu[j] = a.s_jm[slot * J + j];
            }
        }
        __syncthreads();
        ScnCtx sc;
        load_scn(d, s, sc);
        // multiplier update lambda_a = lambda_{a-1} + mu_a C(x*_a): lbar flows to
        // lambda_{a-1} unchanged and to x*_a and the reference rotations
        for (int j = sc.j0 + threadIdx.x; j < sc.j1; j += blockDim.x) {
            const int A = d.ja[j], B = d.jb[j], ty = d.jtype[j];
            const double mpost = a.s_jm_post[slot * J + j];
            const double* lb = a.lbar + 6 * (long long)j;
            typedef du::Dn<12> D12;
            D12 xa[6], xbv[6];
            for (int k = 0; k < 6; ++k) {
                xa[k] = du::var<12>(w.x[6 * A + k], k);
                xbv[k] = du::var<12>(B >= 0 ? w.x[6 * B + k] : 0.0, 6 + k);
            }
            const du::QT<D12> qa12 = du::liftq<D12>(d.q + 4 * A);
            const du::QT<D12> qb12 = du::liftq<D12>(d.q + 4 * (B >= 0 ? B : A));
            D12 g12[12], C12[6];
            du::joint_grad_t<D12>(ty, d.jdata + 16 * j, xa, xbv, qa12, qb12, d.jmu[j], d.jlambda + 6 * j, g12, C12);
            double* xo = a.jxbar + 12 * (long long)j;
            for (int k = 0; k < 12; ++k) {
                double acc = 0.0;
                for (int r = 0; r < 6; ++r) acc += lb[r] * C12[r].d[k];
                xo[k] = mpost * acc;
            }
            typedef du::Dn<8> D8;
            D8 xa8[6], xb8[6];
            for (int k = 0; k < 6; ++k) {
                xa8[k] = du::cst<8>(w.x[6 * A + k]);
                xb8[k] = du::cst<8>(B >= 0 ? w.x[6 * B + k] : 0.0);
            }
            const du::QT<D8> qa8 = du::mkq(du::var<8>(d.q[4 * A], 0), du::var<8>(d.q[4 * A + 1], 1),
                                           du::var<8>(d.q[4 * A + 2], 2), du::var<8>(d.q[4 * A + 3], 3));
            const int Bq = B >= 0 ? B : A;
            const du::QT<D8> qb8 = du::mkq(du::var<8>(d.q[4 * Bq], 4), du::var<8>(d.q[4 * Bq + 1], 5),
                                           du::var<8>(d.q[4 * Bq + 2], 6), du::var<8>(d.q[4 * Bq + 3], 7));
            D8 g8[12], C8[6];
            du::joint_grad_t<D8>(ty, d.jdata + 16 * j, xa8, xb8, qa8, qb8, d.jmu[j], d.jlambda + 6 * j, g8, C8);
            double* qo = a.jvjp + 8 * (long long)j;
            for (int k = 0; k < 8; ++k) {
                double acc = 0.0;
                for (int r = 0; r < 6; ++r) acc += lb[r] * C8[r].d[k];
                qo[k] = mpost * acc;
            }
            if (ty == 0)
                for (int k = 4; k < 8; ++k) qo[k] = 0.0;  // fixed joint: qb8 aliases qa8
            // exact Jacobian of a hinge's gradient, transposed: E[r][c] = dg_c / dx_r
            if (ty == 1) {
                double* E = a.ejac_j + 144 * (long long)j;
                for (int r = 0; r < 12; ++r)
                    for (int c = 0; c < 12; ++c) E[12 * r + c] = g12[c].d[r];
            }
        }
        __syncthreads();
        for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
            const int b = b0 + li;
            for (int j = sc.j0; j < sc.j1; ++j) {
                if (d.ja[j] == b) for (int k = 0; k < 6; ++k) a.xbar[6 * b + k] += a.jxbar[12 * (long long)j + k];
                if (d.jb[j] == b) for (int k = 0; k < 6; ++k) a.xbar[6 * b + k] += a.jxbar[12 * (long long)j + 6 + k];
                if (d.ja[j] == b) for (int k = 0; k < 4; ++k) a.bq[4 * b + k] += a.jvjp[8 * (long long)j + k];
                if (d.jb[j] == b) for (int k = 0; k < 4; ++k) a.bq[4 * b + k] += a.jvjp[8 * (long long)j + 4 + k];
            }
        }
        __syncthreads();
        // a solve whose Newton loop accepted no step returned its starting point: the
        // previous solve's x*, or x_hat = [p + h v, h w] for the first
        if (a.s_nit[slot * S + s] == 0) {
            for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x)
                for (int k = 0; k < 6; ++k) {
                    const double v = a.xbar[6 * b + k];
                    if (it > 0) a.carry[6 * b + k] = v;
                    else if (k < 3) {
                        a.bp[3 * b + k] += v;
                        a.bv[3 * b + k] += h * v;
                    } else {
                        a.bw[3 * b + k - 3] += h * v;
                    }
                }
            __syncthreads();
            continue;
        }
        // Hessian of the solve at x*, then the hinge and contact corrections
        build_csr(w, sc);
        double* b2arr = w.sort_key + (long long)s * kSceneScratch;
        eval_elements(d, w, sc, w.x, b2arr);
        // d = max(1e-9, (pB - pA).n): at the clamp the distance, and so the barrier
        // gradient, does not move with x. Its Hessian term (b'' at 1e-9, enormous) is not
        // part of the exact Jacobian; it is left out of the assembly rather than added and
        // subtracted, which would cancel every smaller term of the block.
        for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
            const long long ce = (long long)s * kMaxContacts + i;
            if (!(w.cactive[ce] & 1)) continue;
            const long long cc = sc.cb + i;
            const int A = w.cint[5 * cc], B = w.cint[5 * cc + 1];
            const double* cd = w.cdbl + 7 * cc;
            const double pr = (w.x[6 * B] - w.x[6 * A]) * cd[3] + (w.x[6 * B + 1] - w.x[6 * A + 1]) * cd[4] +
                              (w.x[6 * B + 2] - w.x[6 * A + 2]) * cd[5];
            if (!(pr > 1e-9)) b2arr[i] = 0.0;
        }
        __syncthreads();
        assemble_diag(d, w, sc, b2arr);  // H, for the GMRES fallback's preconditioner
        assemble_diag_base(d, w, a, sc, b2arr);
        for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
            const long long ce = (long long)s * kMaxContacts + i;
            const long long cc = sc.cb + i;
            double* Dc = a.ejac_c + 144 * ce;
            const int act = w.cactive[ce];
            const int A = w.cint[5 * cc], B = w.cint[5 * cc + 1];
            const double* cd = w.cdbl + 7 * cc;
            if (act & 2) {
                typedef du::Dn<12> D12;
                D12 xa[6], xbv[6];
                for (int k = 0; k < 6; ++k) {
                    xa[k] = du::var<12>(w.x[6 * A + k], k);
                    xbv[k] = du::var<12>(w.x[6 * B + k], 6 + k);
                }
                D12 g[12];
                du::friction_grad_t<D12>(xa, xbv, du::lift3<D12>(d.p[3 * A], d.p[3 * A + 1], d.p[3 * A + 2]),
                                         du::liftq<D12>(d.q + 4 * A),
                                         du::lift3<D12>(d.p[3 * B], d.p[3 * B + 1], d.p[3 * B + 2]),
                                         du::liftq<D12>(d.q + 4 * B), cd, cd + 3, w.nf[ce], du::cst<12>(sc.fmu),
                                         sc.feps, sc.fdt, g);
                // exact Jacobian of the friction gradient (scaled by h^2), transposed
                for (int r = 0; r < 12; ++r)
                    for (int c = 0; c < 12; ++c) Dc[12 * r + c] = g[c].d[r] * sc.h2;
            }
        }
        __syncthreads();
        double rdir = 0.0, rgm = -2.0;
        const int dcode = adj_solve_direct(d, w, a, sc, b2arr, red, &ok_sh, nl.comp, &rdir);
        bool ok = dcode == 0;
        if (!ok) ok = adj_solve_gmres(d, w, a, sc, b2arr, red, &ok_sh, nl, &rgm);
        if (!ok && threadIdx.x == 0 && a.fail[s] == 0) {
            a.fail[s] = 1;
            double* fi = a.fail_info + 5 * (long long)s;
            fi[0] = gstep;
            fi[1] = it;
            fi[2] = dcode;
            fi[3] = rdir;
            fi[4] = rgm;
        }
        const double* ETA = a.eta + 6 * (long long)b0;
        // contact geometry and lagged normal forces: to x*_{a-1} or the step-start poses
        {
            const double* xp = it > 0 ? a.s_x + (slot - 1) * 6 * N : nullptr;
            for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
                const long long ce = (long long)s * kMaxContacts + i;
                double* out = a.gvjp + 14 * ce;
                if (!(w.cactive[ce] & 3)) {
                    for (int k = 0; k < 14; ++k) out[k] = 0.0;
                    continue;
                }
                adj_contact_geom(d, w, sc, i, xp, ETA, b0, out);
            }
        }
        // friction: lagged poses of both bodies and mu
        for (int i = threadIdx.x; i < sc.nc; i += blockDim.x) {
            const long long ce = (long long)s * kMaxContacts + i;
            double* out = a.cvjp + 15 * ce;
            for (int k = 0; k < 15; ++k) out[k] = 0.0;
            if (!(w.cactive[ce] & 2)) continue;
            const long long cc = sc.cb + i;
            const int A = w.cint[5 * cc], B = w.cint[5 * cc + 1];
            const double* cd = w.cdbl + 7 * cc;
            typedef du::Dn<15> D;
            D xa[6], xbv[6];
            for (int k = 0; k < 6; ++k) {
                xa[k] = du::cst<15>(w.x[6 * A + k]);
                xbv[k] = du::cst<15>(w.x[6 * B + k]);
            }
            const du::V3T<D> pA0 = du::mk3(du::var<15>(d.p[3 * A], 0), du::var<15>(d.p[3 * A + 1], 1),
                                           du::var<15>(d.p[3 * A + 2], 2));
            const du::QT<D> qA0 = du::mkq(du::var<15>(d.q[4 * A], 3), du::var<15>(d.q[4 * A + 1], 4),
                                          du::var<15>(d.q[4 * A + 2], 5), du::var<15>(d.q[4 * A + 3], 6));
            const du::V3T<D> pB0 = du::mk3(du::var<15>(d.p[3 * B], 7), du::var<15>(d.p[3 * B + 1], 8),
                                           du::var<15>(d.p[3 * B + 2], 9));
            const du::QT<D> qB0 = du::mkq(du::var<15>(d.q[4 * B], 10), du::var<15>(d.q[4 * B + 1], 11),
                                          du::var<15>(d.q[4 * B + 2], 12), du::var<15>(d.q[4 * B + 3], 13));
            D g[12];
            du::friction_grad_t<D>(xa, xbv, pA0, qA0, pB0, qB0, cd, cd + 3, w.nf[ce], du::var<15>(sc.fmu, 14),
                                   sc.feps, sc.fdt, g);
            const double* lA = ETA + 6 * (A - b0);
            const double* lB = ETA + 6 * (B - b0);
            for (int k = 0; k < 15; ++k) {
                double acc = 0.0;
                for (int j = 0; j < 6; ++j) acc += lA[j] * g[j].d[k] + lB[j] * g[6 + j].d[k];
                out[k] = -(acc * sc.h2);
            }
        }
        // joints: reference rotations, and lambda_{a-1} through g = J^T (mu C + lambda)
        for (int j = sc.j0 + threadIdx.x; j < sc.j1; j += blockDim.x) {
            const int A = d.ja[j], B = d.jb[j], ty = d.jtype[j];
            const double* lA = ETA + 6 * (A - b0);
            const double* lB = B >= 0 ? ETA + 6 * (B - b0) : lA;
            typedef du::Dn<8> D8;
            D8 xa8[6], xb8[6];
            for (int k = 0; k < 6; ++k) {
                xa8[k] = du::cst<8>(w.x[6 * A + k]);
                xb8[k] = du::cst<8>(B >= 0 ? w.x[6 * B + k] : 0.0);
            }
            const int Bq = B >= 0 ? B : A;
            const du::QT<D8> qa8 = du::mkq(du::var<8>(d.q[4 * A], 0), du::var<8>(d.q[4 * A + 1], 1),
                                           du::var<8>(d.q[4 * A + 2], 2), du::var<8>(d.q[4 * A + 3], 3));
            const du::QT<D8> qb8 = du::mkq(du::var<8>(d.q[4 * Bq], 4), du::var<8>(d.q[4 * Bq + 1], 5),
                                           du::var<8>(d.q[4 * Bq + 2], 6), du::var<8>(d.q[4 * Bq + 3], 7));
            D8 g8[12], C8[6];
            du::joint_grad_t<D8>(ty, d.jdata + 16 * j, xa8, xb8, qa8, qb8, d.jmu[j], d.jlambda + 6 * j, g8, C8);
            double* qo = a.jvjp + 8 * (long long)j;
            for (int k = 0; k < 8; ++k) {
                double acc = 0.0;
                for (int r = 0; r < 6; ++r) acc += lA[r] * g8[r].d[k] + (ty == 1 ? lB[r] * g8[6 + r].d[k] : 0.0);
                qo[k] = ty == 0 && k >= 4 ? 0.0 : -acc;
            }
            // dg/dlambda = J^T: lbar_{a-1} = lbar_a - J eta (J the joint Jacobian at x*)
            JEval e;
            joint_eval(d, j, w.x, e);
            for (int r = 0; r < 6; ++r) {
                double acc = 0.0;
                for (int c = 0; c < 6; ++c) acc += e.J[r][c] * lA[c];
                if (ty == 1)
                    for (int c = 0; c < 6; ++c) acc += e.J[r][6 + c] * lB[c];
                a.lbar[6 * (long long)j + r] -= acc;
            }
        }
        __syncthreads();
        const int* off = w.csr_off + sc.b0 + sc.s;
        const int* ent = w.csr_ent + (long long)sc.s * 2 * kMaxContacts;
        for (int li = threadIdx.x; li < sc.n; li += blockDim.x) {
            const int b = b0 + li;
            const double m = d.mass[b];
            const double* I = d.inertia + 9 * b;
            const double* lp = ETA + 6 * li;
            const double gvec[3] = {0.0, kG, 0.0};
            double gm = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double xh = sp_[3 * b + k] + h * sv_[3 * b + k];
                gm -= lp[k] * ((sx[6 * b + k] - xh) - sc.h2 * gvec[k]);
                a.bp[3 * b + k] += m * lp[k];
                a.bv[3 * b + k] += h * (m * lp[k]);
                a.bw[3 * b + k] += h * (I[k] * lp[3] + I[3 + k] * lp[4] + I[6 + k] * lp[5]);
            }
            a.g_mass[b] += gm;
            for (int e = off[li]; e < off[li + 1]; ++e) {
                const int c = ent[e] >> 1, side = ent[e] & 1;
                const double* cv = a.cvjp + 15 * ((long long)s * kMaxContacts + c) + 7 * side;
                for (int k = 0; k < 3; ++k) a.bp[3 * b + k] += cv[k];
                for (int k = 0; k < 4; ++k) a.bq[4 * b + k] += cv[3 + k];
                const double* gv = a.gvjp + 14 * ((long long)s * kMaxContacts + c) + 7 * side;
                if (it > 0) {
                    for (int k = 0; k < 6; ++k) a.carry[6 * b + k] += gv[k];
                } else {
                    for (int k = 0; k < 3; ++k) a.bp[3 * b + k] += gv[k];
                    for (int k = 0; k < 4; ++k) a.bq[4 * b + k] += gv[3 + k];
                }
            }
            for (int j = sc.j0; j < sc.j1; ++j) {
                if (d.ja[j] == b) for (int k = 0; k < 4; ++k) a.bq[4 * b + k] += a.jvjp[8 * (long long)j + k];
                if (d.jb[j] == b) for (int k = 0; k < 4; ++k) a.bq[4 * b + k] += a.jvjp[8 * (long long)j + 4 + k];
            }
        }
        if (threadIdx.x == 0) {
            double gmu = 0.0;
            for (int i = 0; i < sc.nc; ++i) gmu += a.cvjp[15 * ((long long)s * kMaxContacts + i) + 14];
            a.g_mu[s] += gmu;
        }
        __syncthreads();
    }
}

void launch_adj_step(const Dev& d, const Work& w, const Adj& a, int step, int gstep, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_step, dim3(d.S), dim3(kBlk), 0, st, d, w, a, step, gstep);
}

}  // namespace nd
