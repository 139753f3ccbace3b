// ORACLE — test infrastructure only. Nothing in the product (nexdynipc_amd/,
// include/) links, loads or calls this file; only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg use it, as the checker.
//
// Dependency-free C++20 restatement of the CPU path of NexDynIPC's contact stage of
// the implicit-Euler Newton step (W-Moorer/NexDynIPC). Each function cites the
// reference file:line it follows. The reference itself cannot be built here (Eigen,
// Boost, TBB, nlohmann_json, spdlog and Catch2 are absent), so this restatement is the
// oracle. Reference quirks are kept on purpose:
//   - the barrier acts on body-centre distance along the contact normal
//     (DistanceBarrierForm.cpp:89-98);
//   - face-vertex CCD never reports an impact (TimeOfImpact.cpp:130-147);
//   - the body BVH splits by bin counts without partitioning (BVH.cpp:289-362);
//   - top-K ties are decided by std::nth_element (ContactAssembler.cpp:67-88);
//   - the adaptive barrier is initialised once per solver (IPCSolver.cpp:817-822;
//     per process in the reference, i.e. per scene here);
//   - body-frame inertia is used unrotated (InertiaForm.cpp:29-30).
// Parity pinning of each function is listed in DESIGN.md.
#include "ndipc_oracle.h"
#include "ndo_interval.h"
#include "ndo_math.h"
#include "ipc_contact.h"

#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <stack>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace ndo {

thread_local bool g_interval_nan = false;

using std::vector;

struct Mesh {
    vector<V3> V;
    vector<std::array<int, 3>> F;
    vector<std::array<int, 2>> E;
    V3 lmin, lmax;
};

struct Body {
    int id = -1;
    V3 p, v, w, al, aa;
    Q q;
    double mass = 1.0;
    M3 I = M3::identity();
    int mesh = -1;
};

struct Joint {
    int type = 0;  // 0 fixed, 1 hinge
    int a = -1, b = -1;
    V3 tp;
    Q tq;
    V3 anchorA, anchorB, axisA, axisB;
    double mu = 1e8;
    double lambda[6] = {0, 0, 0, 0, 0, 0};
    int gA = -1, gB = -1;
    V3 pA_ref, pB_ref;
    Q qA_ref, qB_ref;
};

struct Contact {
    int A = 0, B = 0, type = 0, pa = 0, pb = 0;
    V3 point, normal;
    double dist = 0;
    // primitives beyond the reference's record, for the CPU adjoint's contact geometry:
    // the second vertex of each edge (EE), or -1 (FV); vv marks the vertex-vertex fallback
    int pa2 = -1, pb2 = -1;
    bool vv = false;
    // The CPU adjoint's differences of the contact geometry keep each contact on the side
    // of the 1e-9 distance clamp it was recorded on (-1: evaluate the clamp; 0: never
    // clamp; 1: always clamped): clamp_x for the barrier at the solve's x, clamp_n for the
    // lagged normal force at the previous x.
    signed char clamp_x = -1, clamp_n = -1;
};

struct CandSet {
    vector<std::array<int, 4>> vv;
    vector<std::array<int, 5>> ev;
    vector<std::array<int, 6>> ee;
    vector<std::array<int, 4>> fv;
    void clear() { vv.clear(); ev.clear(); ee.clear(); fv.clear(); }
};

// ContactAssembler.cpp:15-20
static uint64_t pair_key(int a, int b) {
    const int lo = std::min(a, b), hi = std::max(a, b);
    return (uint64_t(uint32_t(lo)) << 32) | uint32_t(hi);
}

// ContactAssembler.cpp:90-120 — unique edges in face order (v0v1, v1v2, v2v0), stored (min, max).
static vector<std::array<int, 2>> build_unique_edges(const vector<std::array<int, 3>>& F) {
    std::unordered_set<uint64_t> seen;
    vector<std::array<int, 2>> E;
    E.reserve(F.size() * 3);
    for (const auto& f : F) {
        const int e[3][2] = {{f[0], f[1]}, {f[1], f[2]}, {f[2], f[0]}};
        for (auto& ab : e) {
            if (seen.insert(pair_key(ab[0], ab[1])).second)
                E.push_back({std::min(ab[0], ab[1]), std::max(ab[0], ab[1])});
        }
    }
    return E;
}

// ---------------------------------------------------------------------------------
// BVH: Physics::Geometry::BVH (BVH.cpp:19-34 init, :57-90 intersectBox,
// :218-287 buildRecursive, :289-362 findSAHSplit).
struct BVHNode {
    AABB box;
    int left = -1, right = -1, parent = -1, prim = -1;
    bool leaf() const { return prim >= 0; }
};

struct BVH {
    vector<BVHNode> nodes;
    int root = -1;

    void init(const vector<AABB>& boxes) {
        nodes.clear();
        root = -1;
        if (boxes.empty()) return;
        vector<int> prims(boxes.size());
        std::iota(prims.begin(), prims.end(), 0);
        nodes.reserve(2 * boxes.size() - 1);
        root = build(prims, boxes);
    }

    int sah_split(const vector<int>& prims, const vector<AABB>& boxes, int axis, double& split) {
        const int numBins = 16;
        const int n = int(prims.size());
        double minC = std::numeric_limits<double>::max();
        double maxC = std::numeric_limits<double>::lowest();
        for (int p : prims) {
            const double c = boxes[p].center()[axis];
            minC = std::min(minC, c);
            maxC = std::max(maxC, c);
        }
        if (maxC - minC < 1e-10) { split = minC; return n / 2; }
        vector<AABB> binBox(numBins);
        vector<int> binCount(numBins, 0);
        const double binWidth = (maxC - minC) / numBins;
        for (int p : prims) {
            const double c = boxes[p].center()[axis];
            const int bin = std::min(numBins - 1, int((c - minC) / binWidth));
            binBox[bin].expand(boxes[p]);
            binCount[bin]++;
        }
        vector<double> leftArea(numBins, 0), rightArea(numBins, 0);
        vector<int> leftCount(numBins, 0), rightCount(numBins, 0);
        AABB lb; int lt = 0;
        for (int i = 0; i < numBins; ++i) { lb.expand(binBox[i]); lt += binCount[i]; leftArea[i] = lb.surfaceArea(); leftCount[i] = lt; }
        AABB rb; int rt = 0;
        for (int i = numBins - 1; i >= 0; --i) { rb.expand(binBox[i]); rt += binCount[i]; rightArea[i] = rb.surfaceArea(); rightCount[i] = rt; }
        double bestCost = std::numeric_limits<double>::max();
        int best = n / 2;
        split = minC + (maxC - minC) / 2;
        for (int i = 0; i < numBins - 1; ++i) {
            if (leftCount[i] == 0 || rightCount[i + 1] == 0) continue;
            const double cost = leftArea[i] * leftCount[i] + rightArea[i + 1] * rightCount[i + 1];
            if (cost < bestCost) { bestCost = cost; best = leftCount[i]; split = minC + (i + 1) * binWidth; }
        }
        return best;
    }

    int build(vector<int>& prims, const vector<AABB>& boxes) {
        if (prims.empty()) return -1;
        const int id = int(nodes.size());
        nodes.push_back(BVHNode());
        if (prims.size() == 1) {
            nodes[id].box = boxes[prims[0]];
            nodes[id].prim = prims[0];
            return id;
        }
        AABB cb;
        for (int p : prims) cb.expand(boxes[p].center());
        const int axis = cb.maxExtentAxis();
        int mid;
        if (cb.extent()[axis] < 1e-10) {
            mid = int(prims.size()) / 2;
        } else {
            double sp;
            mid = sah_split(prims, boxes, axis, sp);
            if (mid <= 0 || mid >= int(prims.size())) {
                mid = int(prims.size()) / 2;
                std::nth_element(prims.begin(), prims.begin() + mid, prims.end(), [&](int a, int b) {
                    return boxes[a].center()[axis] < boxes[b].center()[axis];
                });
            }
        }
        vector<int> L(prims.begin(), prims.begin() + mid), R(prims.begin() + mid, prims.end());
        const int l = build(L, boxes);
        const int r = build(R, boxes);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].box = nodes[l].box.merged(nodes[r].box);
        if (l >= 0) nodes[l].parent = id;
        if (r >= 0) nodes[r].parent = id;
        return id;
    }

    void query(const AABB& q, vector<unsigned>& out) const {
        out.clear();
        if (root < 0) return;
        std::stack<int> st;
        st.push(root);
        while (!st.empty()) {
            const int id = st.top();
            st.pop();
            const BVHNode& n = nodes[id];
            if (!n.box.intersects(q)) continue;
            if (n.leaf()) {
                out.push_back(unsigned(n.prim));
            } else {
                if (n.left >= 0) st.push(n.left);
                if (n.right >= 0) st.push(n.right);
            }
        }
    }
};

// IncrementalBVHBroadPhase::detect after markNeedsRebuild (BroadPhase.cpp:84-123).
static vector<std::pair<int, int>> broad_phase_detect(const vector<AABB>& boxes) {
    vector<std::pair<int, int>> res;
    if (boxes.size() < 2) return res;
    BVH bvh;
    bvh.init(boxes);
    vector<unsigned> hits;
    for (size_t i = 0; i < boxes.size(); ++i) {
        bvh.query(boxes[i], hits);
        for (unsigned j : hits)
            if (size_t(j) > i) res.emplace_back(int(i), int(j));
    }
    return res;
}

// ContactAssembler.cpp:45-55
static double aabb_distance(const AABB& a, const AABB& b) {
    const double dx = std::max({0.0, a.mn.x - b.mx.x, b.mn.x - a.mx.x});
    const double dy = std::max({0.0, a.mn.y - b.mx.y, b.mn.y - a.mx.y});
    const double dz = std::max({0.0, a.mn.z - b.mx.z, b.mn.z - a.mx.z});
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ContactAssembler.cpp:57-65
static double point_aabb_distance(const V3& p, const AABB& box) {
    const double dx = std::max({box.mn.x - p.x, 0.0, p.x - box.mx.x});
    const double dy = std::max({box.mn.y - p.y, 0.0, p.y - box.mx.y});
    const double dz = std::max({box.mn.z - p.z, 0.0, p.z - box.mx.z});
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ContactAssembler.cpp:67-88 — the scores are pure functions of the index, so they
// are evaluated once; the comparator sees the same values as the reference's.
static vector<int> select_top_k(const vector<double>& score, int k) {
    const int n = int(score.size());
    if (n <= 0 || k <= 0) return {};
    vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    auto cmp = [&](int a, int b) { return score[a] < score[b]; };
    if (k < n) {
        std::nth_element(idx.begin(), idx.begin() + k, idx.end(), cmp);
        idx.resize(k);
    }
    std::sort(idx.begin(), idx.end(), cmp);
    return idx;
}

// ContactAssembler.cpp:22-43
static AABB transform_local_aabb(const V3& mn, const V3& mx, const V3& p, const Q& q) {
    AABB box;
    for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
            for (int iz = 0; iz < 2; ++iz) {
                const V3 c(ix ? mx.x : mn.x, iy ? mx.y : mn.y, iz ? mx.z : mn.z);
                box.expand(p + rotate(q, c));
            }
    return box;
}

struct BodyContactMesh {
    int body_id, world_index;
    const Mesh* mesh;
    V3 pos;
    Q ori;
    AABB world_aabb;
};

// IPCSolver.cpp:32-90
static void closest_points_on_segments(const V3& p1, const V3& q1, const V3& p2, const V3& q2,
                                       V3& c1, V3& c2) {
    const double eps = 1e-12;
    const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double s, t;
    if (a <= eps && e <= eps) { c1 = p1; c2 = p2; return; }
    if (a <= eps) {
        s = 0.0;
        t = std::clamp(f / e, 0.0, 1.0);
    } else {
        const double c = dot(d1, r);
        if (e <= eps) {
            t = 0.0;
            s = std::clamp(-c / a, 0.0, 1.0);
        } else {
            const double b = dot(d1, d2);
            const double denom = a * e - b * b;
            s = (std::abs(denom) > eps) ? std::clamp((b * f - c * e) / denom, 0.0, 1.0) : 0.0;
            t = (b * s + f) / e;
            if (t < 0.0) { t = 0.0; s = std::clamp(-c / a, 0.0, 1.0); }
            else if (t > 1.0) { t = 1.0; s = std::clamp((b - c) / a, 0.0, 1.0); }
        }
    }
    c1 = p1 + d1 * s;
    c2 = p2 + d2 * t;
}

// IPCSolver.cpp:92-105
static V3 safe_normalized(const V3& v, const V3& fb) {
    const double n = norm(v);
    if (n > 1e-12) return v / n;
    const double fn = norm(fb);
    if (fn > 1e-12) return fb / fn;
    return V3(0, 1, 0);
}

// IPCSolver.cpp:107-134
static bool point_in_triangle(const V3& p, const V3& a, const V3& b, const V3& c) {
    const double eps = 1e-8;
    const V3 v0 = b - a, v1 = c - a, v2 = p - a;
    const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
    const double d20 = dot(v2, v0), d21 = dot(v2, v1);
    const double denom = d00 * d11 - d01 * d01;
    if (std::abs(denom) < 1e-16) return false;
    const double v = (d11 * d20 - d01 * d21) / denom;
    const double w = (d00 * d21 - d01 * d20) / denom;
    const double u = 1.0 - v - w;
    return u >= -eps && v >= -eps && w >= -eps;
}

// Physics::BarrierPotential (Barrier.cpp:11-46)
static double barrier_value(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    if (d <= 0.0) return std::numeric_limits<double>::infinity();
    const double r = d / dh;
    return -k * (d - dh) * (d - dh) * nd::ef::log(r);
}
static double barrier_gradient(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    if (d <= 0.0) return -std::numeric_limits<double>::infinity();
    const double diff = d - dh;
    const double lr = nd::ef::log(d / dh);
    return -k * (2 * diff * lr + diff * diff / d);
}
static double barrier_hessian(double d, double dh, double k) {
    if (d >= dh) return 0.0;
    if (d <= 0.0) return std::numeric_limits<double>::infinity();
    const double diff = d - dh;
    const double lr = nd::ef::log(d / dh);
    return -k * (2 * lr + 4 * diff / d - (diff * diff) / (d * d));
}

// Physics::Contact::FrictionPotential (Friction.cpp:5-62)
static double friction_value(const V3& u, double N, double mu, double eps) {
    if (N <= 0) return 0.0;
    const double un = norm(u);
    const double sn = std::sqrt(un * un + eps * eps);
    return mu * N * (sn - eps);
}
static V3 friction_grad(const V3& u, double N, double mu, double eps) {
    if (N <= 0) return V3();
    const double un = norm(u);
    if (un < 1e-12) return V3();
    const double sn = std::sqrt(un * un + eps * eps);
    const double scale = mu * N / sn;
    return scale * u;
}
static M3 friction_hess(const V3& u, double N, double mu, double eps) {
    if (N <= 0) return M3();
    const double un = norm(u);
    if (un < 1e-12) {
        M3 r = M3::identity();
        const double s = mu * N / eps;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r(i, j) = s * r(i, j);
        return r;
    }
    const double sn = std::sqrt(un * un + eps * eps);
    const double scale = mu * N / sn;
    const double s2 = sn * sn;
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = scale * ((i == j ? 1.0 : 0.0) - (u[i] * u[j]) / s2);
    return r;
}

// AdaptiveBarrier::initialize / updateStiffness (AdaptiveBarrier.h:27-37) with the
// default AdaptiveBarrierParams (initial dhat 1e-3, dhat_epsilon 1e-9, min stiffness
// scale 1e11, growth 2, max stiffness 1e20 unless given).
struct AdaptiveBarrierO {
    double kappa = 0.0, dhat = 1e-3, kappa_min = 0.0, max_stiffness = 1e20;
    void initialize(double average_mass, double bbox_diagonal) {
        kappa_min = 1e11 * average_mass / (bbox_diagonal * bbox_diagonal);
        kappa = kappa_min;
        dhat = 1e-3;
    }
    bool update(double min_distance) {
        if (min_distance < 1e-9 && kappa < max_stiffness) {
            kappa = std::min(kappa * 2.0, max_stiffness);
            return true;
        }
        return false;
    }
};

// Simple triplet accumulator (row, col) -> value, summed in insertion order.
struct SymAccum {
    int n = 0;
    vector<std::unordered_map<int, double>> rows;  // full matrix rows
    void reset(int n_) { n = n_; rows.assign(n_, {}); }
    void add(int r, int c, double v) {
        auto it = rows[r].find(c);
        if (it == rows[r].end()) rows[r].emplace(c, v);
        else it->second = it->second + v;
    }
};

// Symmetric skyline (profile) LDL^T with natural ordering. Stands in for Eigen's
// SimplicialLDLT (LinearSolver.cpp:11-38); the ordering differs (AMD there), so the
// last bits of the solution are "parity unpinned".
struct SkylineLDLT {
    int n = 0;
    vector<int> first, off;
    vector<double> L;  // row i: columns first[i]..i (diag slot holds D)
    bool ok = false;
    bool factor(const SymAccum& A) {
        n = A.n;
        first.assign(n, 0);
        off.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            int f = i;
            for (auto& kv : A.rows[i]) if (kv.first < f) f = kv.first;
            first[i] = f;
            off[i + 1] = off[i] + (i - f + 1);
        }
        L.assign(off[n], 0.0);
        for (int i = 0; i < n; ++i)
            for (auto& kv : A.rows[i])
                if (kv.first <= i) L[off[i] + kv.first - first[i]] = kv.second;
        for (int i = 0; i < n; ++i) {
            double* Li = &L[off[i]];
            for (int j = first[i]; j < i; ++j) {
                const double* Lj = &L[off[j]];
                const int k0 = std::max(first[i], first[j]);
                double s = Li[j - first[i]];
                for (int k = k0; k < j; ++k) s -= Li[k - first[i]] * L[off[k] + (k - first[k])] * Lj[k - first[j]];
                Li[j - first[i]] = s / L[off[j] + (j - first[j])];
            }
            double d = Li[i - first[i]];
            for (int k = first[i]; k < i; ++k) {
                const double lik = Li[k - first[i]];
                d -= lik * lik * L[off[k] + (k - first[k])];
            }
            if (d == 0.0 || !std::isfinite(d)) { ok = false; return false; }
            Li[i - first[i]] = d;
        }
        ok = true;
        return true;
    }
    void solve(const vector<double>& b, vector<double>& x) const {
        x = b;
        for (int i = 0; i < n; ++i) {
            const double* Li = &L[off[i]];
            double s = x[i];
            for (int k = first[i]; k < i; ++k) s -= Li[k - first[i]] * x[k];
            x[i] = s;
        }
        for (int i = 0; i < n; ++i) x[i] /= L[off[i] + (i - first[i])];
        for (int i = n - 1; i >= 0; --i) {
            const double* Li = &L[off[i]];
            for (int k = first[i]; k < i; ++k) x[k] -= Li[k - first[i]] * x[i];
        }
    }
};

static V3 seg3(const vector<double>& x, int i) { return V3(x[i], x[i + 1], x[i + 2]); }

// ---------------------------------------------------------------------------------
struct Solver {
    vector<Body> bodies;
    vector<Mesh> meshes;
    vector<Joint> joints;

    // IPCSolver settings (IPCSolver.h:94-124)
    bool enable_ccd = true;
    double ccd_safety = 0.8;
    bool enable_adaptive_barrier = true;
    bool enable_friction = false;
    double friction_coeff = 0.0;
    int alm_max_iters = 10;
    double newton_tol = 1e-6;     // NewtonSolver (NewtonSolver.cpp:10-95)
    int newton_max_iters = 20;
    double alm_constraint_tol = 1e-4, alm_dual_tol = 1e-3, alm_contact_tol = 1e-3;
    double alm_hardening_trigger = 2.0, alm_hardening_ratio = 2.0;
    std::unordered_set<uint64_t> allowed;
    std::unordered_map<uint64_t, double> overrides;

    // AdaptiveBarrier (AdaptiveBarrier.h:22-87) + the once-per-process init flag.
    double kappa = 0.0, dhat = 1e-3, kappa_min = 0.0;
    bool barrier_initialized = false;

    // ContactAssembler state
    vector<BodyContactMesh> bcm;
    CandSet cands;

    // forms of the current step
    bool has_barrier = false, has_friction = false;
    vector<Contact> b_contacts;
    double b_dhat = 1e-3, b_kappa = 0;
    vector<Contact> f_contacts;
    vector<double> f_normal;
    double f_mu = 0, f_eps = 1e-8, f_dt = 0.01;

    // integrator (ImplicitEuler)
    vector<double> x_prev, v_prev, a_prev, x_hat;
    double idt = 0;

    // Recording for the CPU adjoint (adjoint_sweep): every Newton solve of every step
    // with the data its objective reads, the converged x, and the joint stiffness after
    // the hardening that follows it.
    // Optional record of a Newton solve's iterations (tests pin the Newton loop, the
    // linear solve and the line search against it): per iteration the iterate, the
    // solved direction, f(x), the slopes, the contact value and every trial.
    struct LsTrial {
        double alpha, fn, fcn;
    };
    struct NewtonRec {
        vector<double> x, dx;
        double fx = 0, descent = 0, fc = 0, cdesc = 0;
        vector<LsTrial> trials;
    };
    mutable vector<NewtonRec>* newton_rec = nullptr;
    struct AdjSolve {
        vector<NewtonRec> iters;  // filled when newton_rec recording is on
        vector<double> x_star;
        vector<Joint> joints;  // multipliers before the solve, stiffness of the solve
        vector<double> mu_post;
        vector<Contact> b_contacts, f_contacts;
        vector<double> f_normal;
        double b_dhat = 0, b_kappa = 0, f_mu = 0, f_eps = 0, f_dt = 0;
        bool has_barrier = false, has_friction = false;
        bool converged = false;
        int newton_iters = 0;
    };
    struct AdjStep {
        vector<Body> bodies;  // step-start state
        vector<double> x_hat;
        double adt = 0;
        double dhat = 0;  // the narrow phase's dhat in this step
        CandSet cands;    // the step's candidates (built once at the step start)
        vector<AdjSolve> solves;
    };
    vector<AdjStep>* adj_rec = nullptr;
    bool record_newton = false;  // with adj_rec: also record the Newton iterations

    // CPU-baseline sampling of computeSceneMinDistance: only every k-th vertex of the
    // outer loop is evaluated and the measured time is scaled by k (bench only).
    int sample_stride = 1;
    bool prune_scene_min = false;  // exact for the 1e-9 comparison (scene_min_distance)
    bool narrow_raw = false;       // test hook: narrow_phase skips the 2000-contact cap
    double t_scene_min = 0, t_step = 0;

    // statistics of the last step
    double st_alm = 0, st_newton = 0, st_contacts = 0, st_fv = 0, st_ee = 0, st_evals = 0, st_kappa = 0, st_toi = 1;

    int n() const { return int(bodies.size()); }

    bool pair_allowed(int ida, int idb) const {
        if (allowed.empty()) return true;
        return allowed.count(pair_key(ida, idb)) != 0;
    }

    // ContactAssembler::build (ContactAssembler.cpp:124-154)
    void assembler_build() {
        bcm.clear();
        id_bcm_n = size_t(-1);
        for (int i = 0; i < n(); ++i) {
            const Body& b = bodies[i];
            if (b.mesh < 0) continue;
            BodyContactMesh m;
            m.body_id = b.id;
            m.world_index = i;
            m.mesh = &meshes[b.mesh];
            m.pos = b.p;
            m.ori = b.q;
            m.world_aabb = transform_local_aabb(m.mesh->lmin, m.mesh->lmax, b.p, b.q);
            bcm.push_back(m);
        }
    }

    // ContactAssembler::detectBodyPairs (ContactAssembler.cpp:156-178)
    vector<std::pair<int, int>> detect_body_pairs(double r) const {
        vector<std::pair<int, int>> out;
        if (bcm.size() < 2) return out;
        vector<AABB> boxes;
        for (auto& m : bcm) boxes.push_back(m.world_aabb.inflated(std::max(0.0, r)));
        for (auto& p : broad_phase_detect(boxes)) out.emplace_back(bcm[p.first].world_index, bcm[p.second].world_index);
        return out;
    }

    const BodyContactMesh* mesh_by_world(int wi) const {
        for (auto& m : bcm) if (m.world_index == wi) return &m;
        return nullptr;
    }

    // ContactAssembler::estimateMinBodyPairDistance (ContactAssembler.cpp:180-212)
    double min_body_pair_distance(double r) const {
        if (bcm.size() < 2) return std::numeric_limits<double>::infinity();
        auto pairs = detect_body_pairs(r);
        if (pairs.empty()) return std::numeric_limits<double>::infinity();
        double md = std::numeric_limits<double>::infinity();
        for (auto& wp : pairs) {
            const BodyContactMesh* A = mesh_by_world(wp.first);
            const BodyContactMesh* B = mesh_by_world(wp.second);
            if (!A || !B) continue;
            md = std::min(md, aabb_distance(A->world_aabb, B->world_aabb));
        }
        return md;
    }

    static V3 wv(const BodyContactMesh* m, int vid) { return m->pos + rotate(m->ori, m->mesh->V[vid]); }

    // The six top-K lists of a close pair (ContactAssembler.cpp:249-302): vertices of A
    // and of B, edges of A and of B, faces of A and of B nearest the other body's box.
    void pair_lists(const BodyContactMesh* A, const BodyContactMesh* B, vector<int> (&L)[6]) const {
        const int K = 32;
        auto vscores = [&](const BodyContactMesh* X, const BodyContactMesh* Y) {
            vector<double> s(X->mesh->V.size());
            for (size_t i = 0; i < s.size(); ++i) s[i] = point_aabb_distance(wv(X, int(i)), Y->world_aabb);
            return s;
        };
        auto escores = [&](const BodyContactMesh* X, const BodyContactMesh* Y) {
            vector<double> s(X->mesh->E.size());
            for (size_t i = 0; i < s.size(); ++i) {
                const V3 a = wv(X, X->mesh->E[i][0]), b = wv(X, X->mesh->E[i][1]);
                s[i] = std::min(point_aabb_distance(a, Y->world_aabb), point_aabb_distance(b, Y->world_aabb));
            }
            return s;
        };
        auto fscores = [&](const BodyContactMesh* X, const BodyContactMesh* Y) {
            vector<double> s(X->mesh->F.size());
            for (size_t i = 0; i < s.size(); ++i) {
                const auto& f = X->mesh->F[i];
                const V3 a = wv(X, f[0]), b = wv(X, f[1]), c = wv(X, f[2]);
                s[i] = std::min({point_aabb_distance(a, Y->world_aabb), point_aabb_distance(b, Y->world_aabb),
                                 point_aabb_distance(c, Y->world_aabb)});
            }
            return s;
        };
        L[0] = select_top_k(vscores(A, B), std::min<int>(K, int(A->mesh->V.size())));
        L[1] = select_top_k(vscores(B, A), std::min<int>(K, int(B->mesh->V.size())));
        L[2] = select_top_k(escores(A, B), std::min<int>(K, int(A->mesh->E.size())));
        L[3] = select_top_k(escores(B, A), std::min<int>(K, int(B->mesh->E.size())));
        L[4] = select_top_k(fscores(A, B), std::min<int>(K, int(A->mesh->F.size())));
        L[5] = select_top_k(fscores(B, A), std::min<int>(K, int(B->mesh->F.size())));
    }

    // ContactAssembler::buildCandidateSet (ContactAssembler.cpp:214-343)
    void build_candidates(double act, CandSet& out) const {
        out.clear();
        if (bcm.size() < 2) return;
        auto pairs = detect_body_pairs(act);
        for (auto& p : pairs) {
            const BodyContactMesh* A = mesh_by_world(p.first);
            const BodyContactMesh* B = mesh_by_world(p.second);
            if (!A || !B) continue;
            if (aabb_distance(A->world_aabb, B->world_aabb) > act) continue;
            vector<int> L[6];
            pair_lists(A, B, L);
            const auto &vA = L[0], &vB = L[1], &eA = L[2], &eB = L[3], &fA = L[4], &fB = L[5];
            for (int i : vA) for (int j : vB) out.vv.push_back({A->body_id, B->body_id, i, j});
            for (int i : eA) for (int j : vB) out.ev.push_back({A->body_id, B->body_id, A->mesh->E[i][0], A->mesh->E[i][1], j});
            for (int i : eB) for (int j : vA) out.ev.push_back({B->body_id, A->body_id, B->mesh->E[i][0], B->mesh->E[i][1], j});
            for (int i : eA) for (int j : eB)
                out.ee.push_back({A->body_id, B->body_id, A->mesh->E[i][0], A->mesh->E[i][1], B->mesh->E[j][0], B->mesh->E[j][1]});
            for (int i : fA) for (int j : vB) out.fv.push_back({A->body_id, B->body_id, i, j});
            for (int i : fB) for (int j : vA) out.fv.push_back({B->body_id, A->body_id, i, j});
        }
    }

    // ipc mode: the point-triangle and edge-edge pair records of every close pair's top-K
    // lists (the candidates above): faces of A x vertices of B, faces of B x vertices of
    // A, edges of A x edges of B, in that order; pairs of two static bodies skipped.
    void ipc_pairs_from_lists(double act, const vector<char>& stat, vector<std::array<int32_t, 5>>& out) const {
        out.clear();
        if (bcm.size() < 2) return;
        for (auto& p : detect_body_pairs(act)) {
            const BodyContactMesh* A = mesh_by_world(p.first);
            const BodyContactMesh* B = mesh_by_world(p.second);
            if (!A || !B) continue;
            if (aabb_distance(A->world_aabb, B->world_aabb) > act) continue;
            if (stat[p.first] && stat[p.second]) continue;
            vector<int> L[6];
            pair_lists(A, B, L);
            const int a = p.first, b = p.second;
            for (int f : L[4]) for (int v : L[1]) out.push_back({0, b, a, v, f});
            for (int f : L[5]) for (int v : L[0]) out.push_back({0, a, b, v, f});
            for (int i : L[2]) for (int j : L[3]) out.push_back({1, a, b, i, j});
        }
    }

    // id -> world index / contact mesh (the first body with the id; ids are unique in
    // every scene here), hashed like the reference's maps (IPCSolver.cpp:401-410)
    mutable std::unordered_map<int, int> id_index, id_bcm;
    mutable size_t id_index_n = size_t(-1), id_bcm_n = size_t(-1);
    int world_index_of_id(int id) const {
        if (id_index_n != bodies.size()) {
            id_index.clear();
            for (int i = 0; i < n(); ++i) id_index.emplace(bodies[i].id, i);
            id_index_n = bodies.size();
        }
        const auto it = id_index.find(id);
        return it == id_index.end() ? -1 : it->second;
    }
    const BodyContactMesh* mesh_by_id(int id) const {
        if (id_bcm_n != bcm.size()) {
            id_bcm.clear();
            for (size_t k = 0; k < bcm.size(); ++k) id_bcm.emplace(bcm[k].body_id, int(k));
            id_bcm_n = bcm.size();
        }
        const auto it = id_bcm.find(id);
        return it == id_bcm.end() ? nullptr : &bcm[it->second];
    }

    // Body pose used by detectContactsAtState / FrictionForm::getBodyState.
    void pose_at(const vector<double>* x, int wi, V3& p, Q& q) const {
        if (!x) { p = bodies[wi].p; q = bodies[wi].q; return; }
        p = seg3(*x, wi * 6);
        q = apply_theta(bodies[wi].q, seg3(*x, wi * 6 + 3));
    }

    // IPCSolver::detectContacts (IPCSolver.cpp:394-580) for x == nullptr and
    // IPCSolver::detectContactsAtState (IPCSolver.cpp:582-806) otherwise.
    vector<Contact> narrow_phase(const vector<double>* x, double* evals = nullptr) const {
        vector<Contact> contacts;
        if (bodies.empty()) return contacts;
        const double dh = dhat;
        const size_t kMaxRaw = 40000, kMax = 2000;
        // world vertices, each computed once per call (the same expression the reference
        // evaluates per fetch, so the same values)
        vector<vector<V3>> wv(bodies.size());
        vector<vector<char>> wv_ok(bodies.size());
        auto world_vertex = [&](int body_id, int lv, V3& out) -> bool {
            const BodyContactMesh* m = mesh_by_id(body_id);
            const int wi = world_index_of_id(body_id);
            if (!m || wi < 0) return false;
            if (lv < 0 || lv >= int(m->mesh->V.size())) return false;
            if (x && (wi * 6 + 5 >= int(x->size()))) return false;
            if (wv[wi].empty()) {
                wv[wi].resize(m->mesh->V.size());
                wv_ok[wi].assign(m->mesh->V.size(), 0);
            }
            if (!wv_ok[wi][lv]) {
                V3 p; Q q;
                pose_at(x, wi, p, q);
                wv[wi][lv] = p + rotate(q, m->mesh->V[lv]);
                wv_ok[wi][lv] = 1;
            }
            out = wv[wi][lv];
            return true;
        };
        auto append = [&](int ida, int idb, int type, int pa, int pb, const V3& pt, const V3& nrm, double d,
                          int pa2 = -1, int pb2 = -1, bool vv = false) {
            if (contacts.size() >= kMaxRaw) return;
            const int ia = world_index_of_id(ida), ib = world_index_of_id(idb);
            if (ia < 0 || ib < 0) return;
            if (!pair_allowed(ida, idb)) return;
            V3 pA, pB; Q qA, qB;
            pose_at(x, ia, pA, qA);
            pose_at(x, ib, pB, qB);
            V3 nn = safe_normalized(nrm, pB - pA);
            if (dot(nn, pB - pA) < 0.0) nn = -nn;
            Contact c;
            c.A = ia; c.B = ib; c.type = type; c.pa = pa; c.pb = pb;
            c.point = pt; c.normal = nn; c.dist = std::max(0.0, d);
            c.pa2 = pa2; c.pb2 = pb2; c.vv = vv;
            contacts.push_back(c);
        };
        double ev = 0;
        for (auto& c : cands.fv) {
            const BodyContactMesh* fm = mesh_by_id(c[0]);
            if (!fm) continue;
            if (c[2] < 0 || c[2] >= int(fm->mesh->F.size())) continue;
            const auto& f = fm->mesh->F[c[2]];
            V3 v0, v1, v2, p;
            if (!world_vertex(c[0], f[0], v0) || !world_vertex(c[0], f[1], v1) || !world_vertex(c[0], f[2], v2) ||
                !world_vertex(c[1], c[3], p))
                continue;
            ev += 1;
            const V3 raw = cross(v1 - v0, v2 - v0);
            const V3 nn = safe_normalized(raw, V3(0, 1, 0));
            const double sd = dot(p - v0, nn);
            const double dist = std::abs(sd);
            if (dist >= dh) continue;
            const V3 proj = p - sd * nn;
            if (!point_in_triangle(proj, v0, v1, v2)) continue;
            append(c[0], c[1], 0, c[2], c[3], 0.5 * (proj + p), (sd >= 0.0) ? nn : -nn, dist);
        }
        for (auto& c : cands.ee) {
            V3 a0, a1, b0, b1;
            if (!world_vertex(c[0], c[2], a0) || !world_vertex(c[0], c[3], a1) || !world_vertex(c[1], c[4], b0) ||
                !world_vertex(c[1], c[5], b1))
                continue;
            ev += 1;
            V3 c0, c1;
            closest_points_on_segments(a0, a1, b0, b1, c0, c1);
            const V3 delta = c1 - c0;
            const double dist = norm(delta);
            if (dist >= dh) continue;
            append(c[0], c[1], 1, c[2], c[4], 0.5 * (c0 + c1), delta, dist, c[3], c[5]);
        }
        if (contacts.empty()) {
            for (auto& c : cands.vv) {
                V3 a, b;
                if (!world_vertex(c[0], c[2], a) || !world_vertex(c[1], c[3], b)) continue;
                ev += 1;
                const V3 d = b - a;
                const double dist = norm(d);
                if (dist >= dh) continue;
                append(c[0], c[1], 0, c[2], c[3], 0.5 * (a + b), d, dist, -1, -1, true);
            }
        }
        if (contacts.size() > kMax && !narrow_raw) {
            std::sort(contacts.begin(), contacts.end(), [](const Contact& a, const Contact& b) { return a.dist < b.dist; });
            contacts.resize(kMax);
        }
        if (evals) *evals += ev;
        return contacts;
    }

    // Point and normal of a recorded contact recomputed from its own primitives at state x
    // (nullptr: the world poses), with narrow_phase's formulas; the adjoint differentiates
    // the contact geometry through this.
    void contact_geom(const Contact& c, const vector<double>* x, V3& point, V3& normal) const {
        auto wvert = [&](int wi, int lv) {
            V3 p; Q q;
            pose_at(x, wi, p, q);
            return p + rotate(q, meshes[bodies[wi].mesh].V[lv]);
        };
        V3 nrm;
        if (c.type == 0 && !c.vv) {
            const auto& f = meshes[bodies[c.A].mesh].F[c.pa];
            const V3 v0 = wvert(c.A, f[0]), v1 = wvert(c.A, f[1]), v2 = wvert(c.A, f[2]), p = wvert(c.B, c.pb);
            const V3 nn = safe_normalized(cross(v1 - v0, v2 - v0), V3(0, 1, 0));
            const double sd = dot(p - v0, nn);
            const V3 proj = p - sd * nn;
            point = 0.5 * (proj + p);
            nrm = (sd >= 0.0) ? nn : -nn;
        } else if (c.type == 1) {
            V3 c0, c1;
            closest_points_on_segments(wvert(c.A, c.pa), wvert(c.A, c.pa2), wvert(c.B, c.pb), wvert(c.B, c.pb2), c0, c1);
            point = 0.5 * (c0 + c1);
            nrm = c1 - c0;
        } else {
            const V3 a = wvert(c.A, c.pa), b = wvert(c.B, c.pb);
            point = 0.5 * (a + b);
            nrm = b - a;
        }
        V3 pA, pB; Q qA, qB;
        pose_at(x, c.A, pA, qA);
        pose_at(x, c.B, pB, qB);
        normal = safe_normalized(nrm, pB - pA);
        if (dot(normal, pB - pA) < 0.0) normal = -normal;
    }

    // DistanceBarrierForm::contactDistance (DistanceBarrierForm.cpp:89-98)
    static double contact_distance(const Contact& c, const vector<double>& x, signed char clamp = -1) {
        const V3 pA = seg3(x, c.A * 6), pB = seg3(x, c.B * 6);
        const double pr = dot(pB - pA, c.normal);
        if (clamp == 1) return 1e-9;
        if (clamp == 0) return pr;
        return std::max(1e-9, pr);
    }
    // DistanceBarrierForm::value/gradient/hessian/computeNormalForces (:11-87)
    static double bar_value(const vector<Contact>& cs, const vector<double>& x, double dh, double k) {
        double t = 0.0;
        for (auto& c : cs) t += barrier_value(contact_distance(c, x, c.clamp_x), dh, k);
        return t;
    }
    static void bar_grad(const vector<Contact>& cs, const vector<double>& x, double dh, double k, vector<double>& g) {
        g.assign(x.size(), 0.0);
        for (auto& c : cs) {
            const double d = contact_distance(c, x, c.clamp_x);
            if (d >= dh) continue;
            const double b1 = barrier_gradient(d, dh, k);
            for (int i = 0; i < 3; ++i) {
                g[c.A * 6 + i] += -b1 * c.normal[i];
                g[c.B * 6 + i] += b1 * c.normal[i];
            }
        }
    }
    static void bar_hess(const vector<Contact>& cs, const vector<double>& x, double dh, double k, double scale,
                         SymAccum& H) {
        for (auto& c : cs) {
            const double d = contact_distance(c, x, c.clamp_x);
            if (d >= dh) continue;
            const double b2 = barrier_hessian(d, dh, k);
            const M3 nn = outer(c.normal, c.normal);
            for (int r = 0; r < 3; ++r)
                for (int col = 0; col < 3; ++col) {
                    const double blk = b2 * nn(r, col);
                    H.add(c.A * 6 + r, c.A * 6 + col, blk * scale);
                    H.add(c.B * 6 + r, c.B * 6 + col, blk * scale);
                    H.add(c.A * 6 + r, c.B * 6 + col, -blk * scale);
                    H.add(c.B * 6 + r, c.A * 6 + col, -blk * scale);
                }
        }
    }
    static vector<double> bar_normal_forces(const vector<Contact>& cs, const vector<double>& x, double dh, double k) {
        vector<double> nf;
        for (auto& c : cs) {
            const double d = contact_distance(c, x, c.clamp_n);
            if (d < dh && d > 0.0) nf.push_back(std::abs(barrier_gradient(d, dh, k)));
            else nf.push_back(0.0);
        }
        return nf;
    }

    // FrictionForm (FrictionForm.cpp:19-224)
    struct FrictionKin {
        V3 u;
        M3 P;
        V3 rA, rB;
    };
    FrictionKin fric_kin(const Contact& c, const vector<double>& x, double dt) const {
        V3 pA, pB; Q qA, qB;
        pose_at(&x, c.A, pA, qA);
        pose_at(&x, c.B, pB, qB);
        const Body& bA = bodies[c.A];
        const Body& bB = bodies[c.B];
        const V3 dispA = pA - bA.p, dispB = pB - bB.p;
        const V3 lA = rotate(conj(bA.q), c.point - bA.p);
        const V3 lB = rotate(conj(bB.q), c.point - bB.p);
        FrictionKin k;
        k.rA = rotate(qA, lA);
        k.rB = rotate(qB, lB);
        V3 wA, wB, vA, vB;
        if (dt > 1e-12) {
            wA = seg3(x, c.A * 6 + 3) / dt;
            wB = seg3(x, c.B * 6 + 3) / dt;
            vA = dispA / dt;
            vB = dispB / dt;
        }
        vA += cross(wA, k.rA);
        vB += cross(wB, k.rB);
        const V3 rel = vB - vA;
        const V3 nrm = (sqnorm(c.normal) > 1e-16) ? normalized(c.normal) : V3(0, 1, 0);
        const V3 tv = rel - dot(rel, nrm) * nrm;
        k.u = tv * dt;
        M3 P = M3::identity();
        const M3 nn = outer(nrm, nrm);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) P(i, j) = P(i, j) - nn(i, j);
        k.P = P;
        return k;
    }
    // J_A = [-P, P [rA]x], J_B = [P, -P [rB]x]   (FrictionForm.cpp:81-87)
    static void fric_jac(const FrictionKin& k, double JA[3][6], double JB[3][6]) {
        const M3 PA = mul(k.P, skew(k.rA));
        const M3 PB = mul(k.P, skew(k.rB));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                JA[i][j] = -k.P(i, j);
                JB[i][j] = k.P(i, j);
                JA[i][3 + j] = PA(i, j);
                JB[i][3 + j] = -PB(i, j);
            }
    }
    double fric_value(const vector<Contact>& cs, const vector<double>& nf, const vector<double>& x, double mu,
                      double eps, double dt) const {
        double t = 0.0;
        for (size_t i = 0; i < cs.size(); ++i) {
            if (i >= nf.size()) continue;
            if (nf[i] <= 0) continue;
            t += friction_value(fric_kin(cs[i], x, dt).u, nf[i], mu, eps);
        }
        return t;
    }
    void fric_grad(const vector<Contact>& cs, const vector<double>& nf, const vector<double>& x, double mu, double eps,
                   double dt, vector<double>& g) const {
        g.assign(x.size(), 0.0);
        for (size_t i = 0; i < cs.size(); ++i) {
            if (i >= nf.size()) continue;
            if (nf[i] <= 0) continue;
            const FrictionKin k = fric_kin(cs[i], x, dt);
            const V3 fg = friction_grad(k.u, nf[i], mu, eps);
            double JA[3][6], JB[3][6];
            fric_jac(k, JA, JB);
            for (int j = 0; j < 6; ++j) {
                g[cs[i].A * 6 + j] += JA[0][j] * fg.x + JA[1][j] * fg.y + JA[2][j] * fg.z;
                g[cs[i].B * 6 + j] += JB[0][j] * fg.x + JB[1][j] * fg.y + JB[2][j] * fg.z;
            }
        }
    }
    void fric_hess(const vector<Contact>& cs, const vector<double>& nf, const vector<double>& x, double mu, double eps,
                   double dt, double scale, SymAccum& H) const {
        for (size_t i = 0; i < cs.size(); ++i) {
            if (i >= nf.size()) continue;
            if (nf[i] <= 0) continue;
            const FrictionKin k = fric_kin(cs[i], x, dt);
            const M3 fh = friction_hess(k.u, nf[i], mu, eps);
            double JA[3][6], JB[3][6];
            fric_jac(k, JA, JB);
            auto JtHJ = [&](double X[3][6], double Y[3][6], double out[6][6]) {
                double T[6][3];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 3; ++c) T[r][c] = X[0][r] * fh(0, c) + X[1][r] * fh(1, c) + X[2][r] * fh(2, c);
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) out[r][c] = T[r][0] * Y[0][c] + T[r][1] * Y[1][c] + T[r][2] * Y[2][c];
            };
            double HAA[6][6], HBB[6][6], HAB[6][6], HBA[6][6];
            JtHJ(JA, JA, HAA);
            JtHJ(JB, JB, HBB);
            JtHJ(JA, JB, HAB);
            JtHJ(JB, JA, HBA);
            const int a = cs[i].A * 6, b = cs[i].B * 6;
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) {
                    H.add(a + r, a + c, HAA[r][c] * scale);
                    H.add(b + r, b + c, HBB[r][c] * scale);
                    H.add(a + r, b + c, HAB[r][c] * scale);
                    H.add(b + r, a + c, HBA[r][c] * scale);
                }
        }
    }

    // ---------------- joints (Joint.cpp:5-71, FixedJoint.cpp, HingeJoint.cpp,
    // JointPrimitives.h:11-108,208-245,281-324)
    struct JRow {
        int col;
        double v;
    };
    void joint_C(const Joint& j, const vector<double>& x, double C[6]) const {
        if (j.type == 0) {
            if (j.gA < 0) { for (int i = 0; i < 6; ++i) C[i] = 0; return; }
            const V3 p = seg3(x, j.gA);
            const V3 d = p - j.tp;
            const V3 tt = to_rotation_vector(mul(j.tq, conj(j.qA_ref)));
            const V3 th = seg3(x, j.gA + 3);
            const V3 e = th - tt;
            C[0] = d.x; C[1] = d.y; C[2] = d.z; C[3] = e.x; C[4] = e.y; C[5] = e.z;
        } else {
            V3 pA = j.pA_ref, pB = j.pB_ref;
            Q qA = j.qA_ref, qB = j.qB_ref;
            if (j.gA >= 0) { pA = seg3(x, j.gA); qA = update_rotation(j.qA_ref, seg3(x, j.gA + 3)); }
            if (j.gB >= 0) { pB = seg3(x, j.gB); qB = update_rotation(j.qB_ref, seg3(x, j.gB + 3)); }
            const V3 rA = rotate(qA, j.anchorA), rB = rotate(qB, j.anchorB);
            const V3 pv = (pA + rA) - (pB + rB);
            const V3 nA = rotate(qA, j.axisA), nB = rotate(qB, j.axisB);
            const V3 ac = cross(nA, nB);
            C[0] = pv.x; C[1] = pv.y; C[2] = pv.z; C[3] = ac.x; C[4] = ac.y; C[5] = ac.z;
        }
    }
    // Jacobian as 6 sparse rows (triplets summed per (row, col)).
    void joint_J(const Joint& j, const vector<double>& x, vector<JRow> rows[6]) const {
        for (int r = 0; r < 6; ++r) rows[r].clear();
        auto put = [&](int r, int c, double v) {
            for (auto& e : rows[r]) if (e.col == c) { e.v = e.v + v; return; }
            rows[r].push_back({c, v});
        };
        if (j.type == 0) {
            if (j.gA < 0) return;
            for (int i = 0; i < 3; ++i) put(i, j.gA + i, 1.0);
            for (int i = 0; i < 3; ++i) put(3 + i, j.gA + 3 + i, 1.0);
            return;
        }
        Q qA = j.qA_ref, qB = j.qB_ref;
        if (j.gA >= 0) qA = update_rotation(j.qA_ref, seg3(x, j.gA + 3));
        if (j.gB >= 0) qB = update_rotation(j.qB_ref, seg3(x, j.gB + 3));
        const V3 rA = rotate(qA, j.anchorA), rB = rotate(qB, j.anchorB);
        if (j.gA >= 0) {
            for (int i = 0; i < 3; ++i) put(i, j.gA + i, 1.0);
            const M3 s = skew(rA);
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) put(i, j.gA + 3 + k, -s(i, k));
        }
        if (j.gB >= 0) {
            for (int i = 0; i < 3; ++i) put(i, j.gB + i, -1.0);
            const M3 s = skew(rB);
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) put(i, j.gB + 3 + k, s(i, k));
        }
        const V3 nA = rotate(qA, j.axisA), nB = rotate(qB, j.axisB);
        if (j.gA >= 0) {
            const M3 dA = mul(skew(nB), skew(nA));
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) put(3 + i, j.gA + 3 + k, dA(i, k));
        }
        if (j.gB >= 0) {
            M3 mnA = skew(nA);
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) mnA(i, k) = -mnA(i, k);
            const M3 dB = mul(mnA, skew(nB));
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) put(3 + i, j.gB + 3 + k, dB(i, k));
        }
    }
    double joint_value(const Joint& j, const vector<double>& x) const {
        // Joint.cpp:5-17: lambda.dot(C) and C.squaredNorm() over dim() = 6 entries are
        // Eigen reductions of dynamic-size vectors (vecsum.h)
        double C[6];
        joint_C(j, x, C);
        const double t1 = nd::eigen_dot6(j.lambda, C);
        const double sq = nd::eigen_sqnorm6(C);
        return t1 + 0.5 * j.mu * sq;
    }
    void joint_grad(const Joint& j, const vector<double>& x, vector<double>& g) const {
        double C[6];
        joint_C(j, x, C);
        vector<JRow> rows[6];
        joint_J(j, x, rows);
        double f[6];
        for (int i = 0; i < 6; ++i) f[i] = j.mu * C[i] + j.lambda[i];
        // J^T f accumulated column by column over rows in increasing order.
        std::unordered_map<int, double> out;
        vector<int> order;
        for (int r = 0; r < 6; ++r)
            for (auto& e : rows[r]) {
                if (!out.count(e.col)) { out[e.col] = 0.0; order.push_back(e.col); }
                out[e.col] = out[e.col] + e.v * f[r];
            }
        for (int c : order) g[c] += out[c];
    }
    void joint_hess(const Joint& j, const vector<double>& x, SymAccum& H) const {
        vector<JRow> rows[6];
        joint_J(j, x, rows);
        std::unordered_map<int64_t, double> acc;
        vector<int64_t> order;
        for (int r = 0; r < 6; ++r)
            for (auto& a : rows[r])
                for (auto& b : rows[r]) {
                    const int64_t key = int64_t(a.col) * (1 << 30) + b.col;
                    if (!acc.count(key)) { acc[key] = 0.0; order.push_back(key); }
                    acc[key] = acc[key] + a.v * b.v;
                }
        for (auto key : order) H.add(int(key / (1 << 30)), int(key % (1 << 30)), acc[key] * j.mu);
    }

    // ---------------- SimulationProblem (IPCSolver.cpp:137-262)
    double obj_value(const vector<double>& x) const {
        const double h2 = idt * idt;
        double val = 0.0;
        for (int i = 0; i < n(); ++i) {  // InertiaForm.cpp:14-35
            const Body& b = bodies[i];
            const V3 dp = seg3(x, i * 6) - seg3(x_hat, i * 6);
            val += 0.5 * b.mass * sqnorm(dp);
            const V3 dt = seg3(x, i * 6 + 3) - seg3(x_hat, i * 6 + 3);
            const V3 r = 0.5 * dt;
            const double s0 = r.x * b.I(0, 0) + r.y * b.I(1, 0) + r.z * b.I(2, 0);
            const double s1 = r.x * b.I(0, 1) + r.y * b.I(1, 1) + r.z * b.I(2, 1);
            const double s2 = r.x * b.I(0, 2) + r.y * b.I(1, 2) + r.z * b.I(2, 2);
            val += s0 * dt.x + s1 * dt.y + s2 * dt.z;
        }
        double gv = 0.0;  // GravityForm.cpp:8-17
        const V3 g(0, -9.81, 0);
        for (int i = 0; i < n(); ++i) gv -= bodies[i].mass * dot(g, seg3(x, i * 6));
        val += h2 * gv;
        double cv = 0.0;
        for (auto& j : joints) cv += joint_value(j, x);
        val += cv;
        if (has_barrier) val += h2 * bar_value(b_contacts, x, b_dhat, b_kappa);
        if (has_friction) val += h2 * fric_value(f_contacts, f_normal, x, f_mu, f_eps, f_dt);
        return val;
    }
    double contact_value(const vector<double>& x) const {
        const double h2 = idt * idt;
        double val = 0.0;
        if (has_barrier) val += h2 * bar_value(b_contacts, x, b_dhat, b_kappa);
        if (has_friction) val += h2 * fric_value(f_contacts, f_normal, x, f_mu, f_eps, f_dt);
        return val;
    }
    void contact_grad(const vector<double>& x, vector<double>& g) const {
        const double h2 = idt * idt;
        g.assign(x.size(), 0.0);
        vector<double> t;
        if (has_barrier) { bar_grad(b_contacts, x, b_dhat, b_kappa, t); for (size_t i = 0; i < g.size(); ++i) g[i] += h2 * t[i]; }
        if (has_friction) { fric_grad(f_contacts, f_normal, x, f_mu, f_eps, f_dt, t); for (size_t i = 0; i < g.size(); ++i) g[i] += h2 * t[i]; }
    }
    void obj_grad(const vector<double>& x, vector<double>& g) const {
        const double h2 = idt * idt;
        g.assign(x.size(), 0.0);
        vector<double> t(x.size(), 0.0);
        for (int i = 0; i < n(); ++i) {  // InertiaForm.cpp:37-53
            const Body& b = bodies[i];
            const V3 dp = seg3(x, i * 6) - seg3(x_hat, i * 6);
            const V3 gp = b.mass * dp;
            const V3 dt = seg3(x, i * 6 + 3) - seg3(x_hat, i * 6 + 3);
            const V3 gr = mul(b.I, dt);
            t[i * 6 + 0] = gp.x; t[i * 6 + 1] = gp.y; t[i * 6 + 2] = gp.z;
            t[i * 6 + 3] = gr.x; t[i * 6 + 4] = gr.y; t[i * 6 + 5] = gr.z;
        }
        for (size_t i = 0; i < g.size(); ++i) g[i] += t[i];
        const V3 gvec(0, -9.81, 0);
        for (int i = 0; i < n(); ++i) {  // GravityForm.cpp:19-28
            const V3 f = (-bodies[i].mass) * gvec;
            t[i * 6 + 0] = f.x; t[i * 6 + 1] = f.y; t[i * 6 + 2] = f.z;
            t[i * 6 + 3] = 0; t[i * 6 + 4] = 0; t[i * 6 + 5] = 0;
        }
        for (size_t i = 0; i < g.size(); ++i) g[i] += h2 * t[i];
        std::fill(t.begin(), t.end(), 0.0);
        for (auto& j : joints) joint_grad(j, x, t);
        for (size_t i = 0; i < g.size(); ++i) g[i] += t[i];
        if (has_barrier) { bar_grad(b_contacts, x, b_dhat, b_kappa, t); for (size_t i = 0; i < g.size(); ++i) g[i] += h2 * t[i]; }
        if (has_friction) { fric_grad(f_contacts, f_normal, x, f_mu, f_eps, f_dt, t); for (size_t i = 0; i < g.size(); ++i) g[i] += h2 * t[i]; }
    }
    void obj_hess(const vector<double>& x, SymAccum& H) const {
        const double h2 = idt * idt;
        H.reset(int(x.size()));
        for (int i = 0; i < n(); ++i) {  // InertiaForm.cpp:55-74
            const Body& b = bodies[i];
            for (int k = 0; k < 3; ++k) H.add(i * 6 + k, i * 6 + k, b.mass);
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) H.add(i * 6 + 3 + r, i * 6 + 3 + c, b.I(r, c));
        }
        for (auto& j : joints) joint_hess(j, x, H);
        if (has_barrier) bar_hess(b_contacts, x, b_dhat, b_kappa, h2, H);
        if (has_friction) fric_hess(f_contacts, f_normal, x, f_mu, f_eps, f_dt, h2, H);
    }

    // BacktrackingLineSearch::search (LineSearch.cpp:6-58)
    double line_search(const vector<double>& x, const vector<double>& dir, const vector<double>& g,
                       const vector<double>& cg, double alpha0) const {
        double alpha = alpha0;
        const double fx = obj_value(x);
        // grad.dot(dir), contact_grad->dot(dir): Eigen dot products of VectorXd (vecsum.h)
        const double descent = nd::eigen_vec_sum(int(g.size()), [&](int i) { return g[i] * dir[i]; });
        const double fc = contact_value(x);
        const double cdesc = nd::eigen_vec_sum(int(cg.size()), [&](int i) { return cg[i] * dir[i]; });
        NewtonRec* nr = newton_rec && !newton_rec->empty() ? &newton_rec->back() : nullptr;
        if (nr) { nr->fx = fx; nr->descent = descent; nr->fc = fc; nr->cdesc = cdesc; }
        if (descent > 0) return 0.0;
        vector<double> xn(x.size());
        for (int it = 0; it < 20; ++it) {
            for (size_t i = 0; i < x.size(); ++i) xn[i] = x[i] + alpha * dir[i];
            const double fn = obj_value(xn);
            const double fcn = contact_value(xn);
            if (nr) nr->trials.push_back({alpha, fn, fcn});
            const bool cok = (cdesc <= 0.0) ? (fcn <= fc + 1e-4 * alpha * cdesc) : (fcn <= fc);
            if (fn <= fx + 1e-4 * alpha * descent && cok) return alpha;
            alpha *= 0.5;
        }
        return 0.0;
    }

    // NewtonSolver::minimize (NewtonSolver.cpp:10-95)
    bool newton(vector<double>& x, int& iters) const {
        vector<double> g, cg, dx, ng;
        SymAccum H;
        SkylineLDLT ldlt;
        for (int it = 0; it < newton_max_iters; ++it) {
            iters = it;
            obj_grad(x, g);
            double ginf = 0.0;
            for (double v : g) ginf = std::max(ginf, std::abs(v));
            if (ginf < newton_tol) return true;
            obj_hess(x, H);
            if (!ldlt.factor(H)) return false;
            ng.resize(g.size());
            for (size_t i = 0; i < g.size(); ++i) ng[i] = -g[i];
            ldlt.solve(ng, dx);
            contact_grad(x, cg);
            if (newton_rec) {
                NewtonRec nr;
                nr.x = x;
                nr.dx = dx;
                newton_rec->push_back(std::move(nr));
            }
            const double alpha = line_search(x, dx, g, cg, std::min(1.0, 1.0));
            if (alpha == 0.0) return false;
            for (size_t i = 0; i < x.size(); ++i) x[i] = x[i] + alpha * dx[i];
        }
        return false;
    }

    // IPCSolver::initializeAdaptiveBarrier (IPCSolver.cpp:327-353) + AdaptiveBarrier::initialize
    void init_barrier() {
        if (bodies.empty()) return;
        double tm = 0.0;
        const double inf = std::numeric_limits<double>::infinity();
        V3 mn(inf, inf, inf), mx(-inf, -inf, -inf);
        for (auto& b : bodies) { tm += b.mass; mn = cmin(mn, b.p); mx = cmax(mx, b.p); }
        const double avg = tm / double(bodies.size());
        double diag = norm(mx - mn);
        if (diag < 1e-6) diag = 1.0;
        AdaptiveBarrierO ab;
        ab.initialize(avg, diag);
        kappa_min = ab.kappa_min;
        kappa = ab.kappa;
        dhat = ab.dhat;
    }

    // computeMinDistance / computeSceneMinDistance (AdaptiveBarrier.h:89-162). With
    // prune_scene_min the body pairs whose world boxes are farther apart than
    // 1e-8 + 1e-5 * (longest edge of either mesh) are skipped: a vertex-face distance
    // below 1e-9 with barycentric coordinates inside [-1e-6, 1 + 1e-6] puts the vertex
    // within 1e-9 + 2e-6 * (longest edge) of the face, so the result's comparison with
    // 1e-9 (IPCSolver.cpp:824-831) is unchanged; only the value of larger minima may
    // differ (+inf when every pair is skipped).
    // ---- ipc mode, function level: the primitive-distance barrier of rigid IPC over a
    // list of pairs (kind 0 PT: body of the vertex, body of the face, vertex, face; kind 1
    // EE: bodies and mesh edge indices of edge a and edge b) at Newton variable x. Pairs
    // whose bodies form the same unordered pair and follow each other form a segment;
    // per segment the slot-0 body is the lower index. Per segment the energy, gradient
    // and Hessian are summed in pair order over the pairs within dhat; the total energy
    // is the sum of the segments' energies in segment order.
    static constexpr int kIpcAdvanceIters = 1000;  // conservative-advancement steps per pair
    struct IpcSeg {
        int lo, hi;
        double E, g[12], H[144];
    };
    bool ipc_pair_ok(const int32_t* pr) const {
        const int kind = pr[0], b0 = pr[1], b1 = pr[2];
        if (kind != 0 && kind != 1) return false;
        if (b0 < 0 || b1 < 0 || b0 >= n() || b1 >= n() || b0 == b1) return false;
        if (bodies[b0].mesh < 0 || bodies[b1].mesh < 0) return false;
        const Mesh& M0 = meshes[bodies[b0].mesh];
        const Mesh& M1 = meshes[bodies[b1].mesh];
        if (pr[3] < 0 || pr[4] < 0) return false;
        if (kind == 0) return pr[3] < int(M0.V.size()) && pr[4] < int(M1.F.size());
        return pr[3] < int(M0.E.size()) && pr[4] < int(M1.E.size());
    }
    bool ipc_pair_vertices(const int32_t* pr, const vector<double>& x, double xw[12], int owner[4],
                           double Jth[4][9], int& lo, int& hi) const {
        if (!ipc_pair_ok(pr)) return false;
        const int kind = pr[0], b0 = pr[1], b1 = pr[2];
        const Mesh& M0 = meshes[bodies[b0].mesh];
        const Mesh& M1 = meshes[bodies[b1].mesh];
        int vb[4];
        V3 r[4];
        if (kind == 0) {
            vb[0] = b0; r[0] = M0.V[pr[3]];
            for (int k = 0; k < 3; ++k) { vb[1 + k] = b1; r[1 + k] = M1.V[M1.F[pr[4]][k]]; }
        } else {
            for (int k = 0; k < 2; ++k) {
                vb[k] = b0; r[k] = M0.V[M0.E[pr[3]][k]];
                vb[2 + k] = b1; r[2 + k] = M1.V[M1.E[pr[4]][k]];
            }
        }
        lo = std::min(b0, b1);
        hi = std::max(b0, b1);
        for (int v = 0; v < 4; ++v) {
            const Body& b = bodies[vb[v]];
            const double q[4] = {b.q.w, b.q.x, b.q.y, b.q.z};
            const double rr[3] = {r[v].x, r[v].y, r[v].z};
            nd::ipc::vertex_and_jacobian(&x[6 * size_t(vb[v])], q, rr, xw + 3 * v, Jth[v]);
            owner[v] = vb[v] == lo ? 0 : 1;
        }
        return true;
    }
    // total energy and the per-segment gradient / Hessian blocks; false on a bad pair
    bool ipc_barrier_eval(const int32_t* pairs, int np, const vector<double>& x, double dhat, double kappa, bool psd,
                          double& value, vector<IpcSeg>& segs) const {
        value = 0.0;
        segs.clear();
        for (int i = 0; i < np; ++i) {
            const int32_t* pr = pairs + 5 * i;
            double xw[12], Jth[4][9], g[12], H[144], gb[12], Hb[144], E = 0.0;
            int owner[4], lo, hi;
            if (!ipc_pair_vertices(pr, x, xw, owner, Jth, lo, hi)) return false;
            if (segs.empty() || segs.back().lo != lo || segs.back().hi != hi) {
                IpcSeg sg{};
                sg.lo = lo;
                sg.hi = hi;
                segs.push_back(sg);
            }
            if (!nd::ipc::pair_energy(pr[0], xw, dhat, kappa, &E, g, H, nullptr, nullptr)) continue;
            if (psd) nd::ipc::psd_project<12>(H);
            nd::ipc::fold_to_bodies(g, H, owner, Jth, gb, Hb);
            IpcSeg& sg = segs.back();
            sg.E += E;
            for (int k = 0; k < 12; ++k) sg.g[k] += gb[k];
            for (int k = 0; k < 144; ++k) sg.H[k] += Hb[k];
        }
        for (const IpcSeg& sg : segs) value += sg.E;
        return true;
    }
    // ipc mode maxStep over a pair list; 0, or 1 with fail_index = the first bad pair
    int ipc_max_step(const int32_t* pairs, int np, const vector<double>& x, const vector<double>& dx, double sep,
                     double safety, double& toi, double& alpha, int& fail_index) const;
    // ipc mode time step (see ndo_ipc_step); 0, or -1 for a scene with joints other than
    // single-body fixed joints
    struct IpcStepParams {
        double dhat, kappa, sep, safety, newton_tol, pcg_tol, inflation;
        int max_newton, max_ls, pcg_max_it;
    };
    int ipc_step(double h, const IpcStepParams& P, int32_t* stats);

    double scene_min_distance() const {
        if (prune_scene_min) return scene_min_distance_pruned();
        double md = std::numeric_limits<double>::infinity();
        long long vcount = 0;
        for (int i = 0; i < n(); ++i) {
            if (bodies[i].mesh < 0) continue;
            const Mesh& MA = meshes[bodies[i].mesh];
            for (size_t vi = 0; vi < MA.V.size(); ++vi) {
                if ((vcount++ % sample_stride) != 0) continue;
                const V3 p = bodies[i].p + rotate(bodies[i].q, MA.V[vi]);
                for (int j = i + 1; j < n(); ++j) {
                    if (bodies[j].mesh < 0) continue;
                    const Mesh& MB = meshes[bodies[j].mesh];
                    vector<V3> VB(MB.V.size());
                    for (size_t k = 0; k < VB.size(); ++k) VB[k] = bodies[j].p + rotate(bodies[j].q, MB.V[k]);
                    double d = std::numeric_limits<double>::infinity();
                    for (auto& f : MB.F) {
                        const V3 v0 = VB[f[0]], v1 = VB[f[1]], v2 = VB[f[2]];
                        const V3 e1 = v1 - v0, e2 = v2 - v0;
                        V3 nn = cross(e1, e2);
                        const double area = norm(nn);
                        if (area < 1e-12) continue;
                        nn = normalized(nn);
                        const double dist = std::abs(dot(p - v0, nn));
                        M3 M;
                        M(0, 0) = e1.x; M(1, 0) = e1.y; M(2, 0) = e1.z;
                        M(0, 1) = e2.x; M(1, 1) = e2.y; M(2, 1) = e2.z;
                        M(0, 2) = nn.x; M(1, 2) = nn.y; M(2, 2) = nn.z;
                        const V3 co = mul(inverse(M), p - v0);
                        const double b0 = 1.0 - co.x - co.y, b1 = co.x, b2 = co.y;
                        const double bmin = std::min({b0, b1, b2}), bmax = std::max({b0, b1, b2});
                        if (bmin >= -1e-6 && bmax <= 1.0 + 1e-6) d = std::min(d, dist);
                    }
                    md = std::min(md, d);
                }
            }
        }
        return md;
    }

    // The same test with exact pruning for the result's comparison with 1e-9
    // (IPCSolver.cpp:824-831): body pairs whose world boxes are farther apart than
    // m = 1e-8 + 1e-5 * (longest edge of either mesh), and within a pair the faces whose
    // world box is farther than m from body i's box and the vertices farther than m from
    // body j's box, are skipped. A vertex-face distance below 1e-9 with barycentric
    // coordinates inside [-1e-6, 1 + 1e-6] puts the vertex within 1e-9 + 2e-6 * (longest
    // edge) of the face, so every such pair is kept; each kept test computes the same
    // values as the loop above (per-face data computed once instead of per vertex), so
    // the minimum is the same whenever it is below m. Larger minima may differ (+inf when
    // everything is skipped).
    double scene_min_distance_pruned() const {
        double md = std::numeric_limits<double>::infinity();
        const int N = n();
        vector<AABB> wb(N);
        vector<double> emax(N, 0.0);
        for (int i = 0; i < N; ++i) {
            const Body& b = bodies[i];
            if (b.mesh < 0) continue;
            const Mesh& M = meshes[b.mesh];
            wb[i] = transform_local_aabb(M.lmin, M.lmax, b.p, b.q);
            for (auto& f : M.F)
                for (int k = 0; k < 3; ++k) emax[i] = std::max(emax[i], norm(M.V[f[(k + 1) % 3]] - M.V[f[k]]));
        }
        struct FaceData { V3 v0, nn; M3 inv; AABB box; };
        for (int i = 0; i < N; ++i) {
            if (bodies[i].mesh < 0) continue;
            const Mesh& MA = meshes[bodies[i].mesh];
            for (int j = i + 1; j < N; ++j) {
                if (bodies[j].mesh < 0) continue;
                const double m = 1e-8 + 1e-5 * std::max(emax[i], emax[j]);
                if (aabb_distance(wb[i], wb[j]) > m) continue;
                const Mesh& MB = meshes[bodies[j].mesh];
                vector<V3> VB(MB.V.size());
                for (size_t k = 0; k < VB.size(); ++k) VB[k] = bodies[j].p + rotate(bodies[j].q, MB.V[k]);
                vector<FaceData> fd;
                for (auto& f : MB.F) {
                    const V3 v0 = VB[f[0]], v1 = VB[f[1]], v2 = VB[f[2]];
                    AABB fb;
                    fb.expand(v0); fb.expand(v1); fb.expand(v2);
                    if (aabb_distance(fb, wb[i]) > m) continue;
                    const V3 e1 = v1 - v0, e2 = v2 - v0;
                    V3 nn = cross(e1, e2);
                    const double area = norm(nn);
                    if (area < 1e-12) continue;
                    nn = normalized(nn);
                    M3 M;
                    M(0, 0) = e1.x; M(1, 0) = e1.y; M(2, 0) = e1.z;
                    M(0, 1) = e2.x; M(1, 1) = e2.y; M(2, 1) = e2.z;
                    M(0, 2) = nn.x; M(1, 2) = nn.y; M(2, 2) = nn.z;
                    fd.push_back({v0, nn, inverse(M), fb});
                }
                if (fd.empty()) continue;
                for (size_t vi = 0; vi < MA.V.size(); ++vi) {
                    const V3 p = bodies[i].p + rotate(bodies[i].q, MA.V[vi]);
                    if (point_aabb_distance(p, wb[j]) > m) continue;
                    for (const FaceData& F : fd) {
                        if (point_aabb_distance(p, F.box) > m) continue;
                        const double dist = std::abs(dot(p - F.v0, F.nn));
                        const V3 co = mul(F.inv, p - F.v0);
                        const double b0 = 1.0 - co.x - co.y, b1 = co.x, b2 = co.y;
                        const double bmin = std::min({b0, b1, b2}), bmax = std::max({b0, b1, b2});
                        if (bmin >= -1e-6 && bmax <= 1.0 + 1e-6) md = std::min(md, dist);
                    }
                }
            }
        }
        return md;
    }

    double ccd_earliest_toi(double dt) const;

    // IPCSolver::computeMaxStep (IPCSolver.cpp:355-392)
    double compute_max_step(double dt) {
        if (!enable_ccd || bodies.empty()) return dt;
        double vmax = 0.0;
        for (auto& b : bodies) vmax = std::max(vmax, norm(b.v));
        const double infl = std::max(1e-6, dhat) + vmax * dt;
        auto close = detect_body_pairs(infl);
        bool any = false;
        for (auto& wp : close)
            if (pair_allowed(bodies[wp.first].id, bodies[wp.second].id)) { any = true; break; }
        if (!any) return dt;
        if (cands.fv.empty()) return dt;
        const double toi = ccd_earliest_toi(dt);
        st_toi = toi;
        if (toi < 1.0) return toi * ccd_safety * dt;
        return dt;
    }

    // IPCSolver::step (IPCSolver.cpp:808-1231)
    void step(double dt) {
        st_alm = st_newton = st_contacts = st_evals = st_kappa = 0;
        st_toi = 1.0;
        assembler_build();
        build_candidates(dhat, cands);
        st_fv = double(cands.fv.size());
        st_ee = double(cands.ee.size());
        if (enable_adaptive_barrier) {
            if (!barrier_initialized) { init_barrier(); barrier_initialized = true; }
            const double dh = dhat;
            double bmd = min_body_pair_distance(dh * 2.0);
            if (!std::isfinite(bmd)) bmd = std::numeric_limits<double>::infinity();
            if (bmd < dh * 2.0) {
                const auto t0 = std::chrono::steady_clock::now();
                const double md = scene_min_distance();
                t_scene_min += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() *
                               double(sample_stride);
                if (md < std::numeric_limits<double>::infinity()) {
                    AdaptiveBarrierO ab{kappa, dhat, kappa_min, 1e20};
                    if (ab.update(md)) { kappa = ab.kappa; st_kappa = 1; }
                }
            }
        }
        double adt = dt;
        if (enable_ccd) adt = compute_max_step(dt);
        const int N = n();
        const int nd = N * 6;
        vector<double> xc(nd, 0.0), vc(nd, 0.0), ac(nd, 0.0);
        for (int i = 0; i < N; ++i) {
            const Body& b = bodies[i];
            xc[i * 6 + 0] = b.p.x; xc[i * 6 + 1] = b.p.y; xc[i * 6 + 2] = b.p.z;
            vc[i * 6 + 0] = b.v.x; vc[i * 6 + 1] = b.v.y; vc[i * 6 + 2] = b.v.z;
            vc[i * 6 + 3] = b.w.x; vc[i * 6 + 4] = b.w.y; vc[i * 6 + 5] = b.w.z;
            ac[i * 6 + 0] = b.al.x; ac[i * 6 + 1] = b.al.y; ac[i * 6 + 2] = b.al.z;
            ac[i * 6 + 3] = b.aa.x; ac[i * 6 + 4] = b.aa.y; ac[i * 6 + 5] = b.aa.z;
        }
        x_prev = xc; v_prev = vc; a_prev = ac; idt = adt;
        x_hat.assign(nd, 0.0);
        for (int i = 0; i < nd; ++i) x_hat[i] = x_prev[i] + adt * v_prev[i];  // ImplicitEulerIntegrator.cpp:5-7
        vector<double> x_new = x_hat;
        if (adj_rec) {
            AdjStep st;
            st.bodies = bodies;
            st.x_hat = x_hat;
            st.adt = adt;
            st.dhat = dhat;
            st.cands = cands;
            adj_rec->push_back(std::move(st));
        }

        vector<Contact> active;
        const bool needs = enable_adaptive_barrier || (enable_friction && friction_coeff > 0.0);
        if (needs) active = narrow_phase(nullptr, &st_evals);
        has_barrier = false;
        has_friction = false;
        if (needs) {
            b_dhat = std::max(1e-9, dhat);
            b_kappa = std::max(0.0, kappa);
            b_contacts = active;
        }
        if (enable_adaptive_barrier && needs) has_barrier = true;
        if (enable_friction && friction_coeff > 0) {
            double emu = friction_coeff;
            if (!active.empty() && !overrides.empty()) {
                double s = 0.0;
                int cnt = 0;
                for (auto& c : active) {
                    auto it = overrides.find(pair_key(bodies[c.A].id, bodies[c.B].id));
                    s += (it != overrides.end()) ? it->second : friction_coeff;
                    ++cnt;
                }
                if (cnt > 0) emu = s / double(cnt);
            }
            f_mu = emu;
            f_eps = 1e-8;
            f_dt = std::max(0.0, adt);
            f_contacts = active;
            f_normal.clear();
            has_friction = true;
        }
        if (needs) { b_dhat = std::max(1e-9, dhat); b_kappa = std::max(0.0, kappa); }

        for (int alm = 0; alm < alm_max_iters; ++alm) {
            for (auto& j : joints) {  // IPCSolver.cpp:966-1066
                if (j.type == 1) {
                    j.gA = j.a * 6; j.gB = j.b * 6;
                    j.pA_ref = bodies[j.a].p; j.qA_ref = bodies[j.a].q;
                    j.pB_ref = bodies[j.b].p; j.qB_ref = bodies[j.b].q;
                } else {
                    j.gA = j.a * 6;
                    j.qA_ref = bodies[j.a].q;
                }
            }
            AdjSolve rec_solve;
            if (adj_rec) {
                rec_solve.joints = joints;
                rec_solve.b_contacts = b_contacts;
                rec_solve.f_contacts = f_contacts;
                rec_solve.f_normal = f_normal;
                rec_solve.b_dhat = b_dhat; rec_solve.b_kappa = b_kappa;
                rec_solve.f_mu = f_mu; rec_solve.f_eps = f_eps; rec_solve.f_dt = f_dt;
                rec_solve.has_barrier = has_barrier; rec_solve.has_friction = has_friction;
            }
            int nit = 0;
            if (adj_rec && record_newton) newton_rec = &rec_solve.iters;
            const bool conv = newton(x_new, nit);
            newton_rec = nullptr;
            st_newton += nit + 1;
            double viol = 0.0, dual = 0.0;
            for (auto& j : joints) {  // IPCSolver.cpp:1084-1108
                double C[6];
                joint_C(j, x_new, C);
                double inf = 0.0;
                for (int k = 0; k < 6; ++k) inf = std::max(inf, std::abs(C[k]));
                viol = std::max(viol, inf);
                const double dc = j.mu * inf;
                dual = std::isfinite(dc) ? std::max(dual, std::min(dc, 1e12)) : std::max(dual, 1e12);
                if (inf > alm_hardening_trigger * alm_constraint_tol) {
                    const double nm = j.mu * alm_hardening_ratio;
                    j.mu = std::isfinite(nm) ? std::min(nm, 1e12) : 1e12;
                }
            }
            if (adj_rec) {
                rec_solve.x_star = x_new;
                rec_solve.converged = conv;
                rec_solve.newton_iters = nit;
                for (auto& j : joints) rec_solve.mu_post.push_back(j.mu);
                adj_rec->back().solves.push_back(std::move(rec_solve));
            }
            for (auto& j : joints) {  // ConstraintForm::updateLambdas -> Joint::updateLambda
                double C[6];
                joint_C(j, x_new, C);
                for (int k = 0; k < 6; ++k) j.lambda[k] += j.mu * C[k];
            }
            if (needs) {  // IPCSolver.cpp:1112-1130
                const auto ic = narrow_phase(&x_new, &st_evals);
                b_contacts = ic;
                if (enable_friction && has_friction) {
                    f_contacts = ic;
                    f_normal = bar_normal_forces(b_contacts, x_new, b_dhat, b_kappa);
                }
                st_contacts = double(ic.size());
            }
            double cres = 0.0;
            if (needs) {  // IPCSolver.cpp:1132-1154
                vector<double> cgr(nd, 0.0), t;
                bar_grad(b_contacts, x_new, b_dhat, b_kappa, t);
                for (int i = 0; i < nd; ++i) cgr[i] += t[i];
                if (enable_friction && has_friction) {
                    fric_grad(f_contacts, f_normal, x_new, f_mu, f_eps, f_dt, t);
                    for (int i = 0; i < nd; ++i) cgr[i] += t[i];
                }
                for (double v : cgr) cres = std::max(cres, std::abs(v));
                if (!std::isfinite(cres)) cres = 1e12;
            }
            st_alm = alm + 1;
            if (conv && viol <= alm_constraint_tol && dual <= alm_dual_tol && cres <= alm_contact_tol) break;
        }
        // ImplicitEuler::update_quantities + write-back (IPCSolver.cpp:1168-1195)
        for (int i = 0; i < N; ++i) {
            Body& b = bodies[i];
            double vn[6], an[6];
            for (int k = 0; k < 6; ++k) {
                vn[k] = (x_new[i * 6 + k] - x_prev[i * 6 + k]) / adt;
                an[k] = (vn[k] - v_prev[i * 6 + k]) / adt;
            }
            b.al = V3(an[0], an[1], an[2]);
            b.aa = V3(an[3], an[4], an[5]);
            b.v = V3(vn[0], vn[1], vn[2]);
            b.w = V3(vn[3], vn[4], vn[5]);
            b.p = seg3(x_new, i * 6);
            b.q = write_back_q(b.q, seg3(x_new, i * 6 + 3));
        }
        has_barrier = has_friction = false;
    }
    // IPCSolver.cpp:1187-1191
    static Q write_back_q(const Q& q, const V3& th) {
        const double ang = norm(th);
        if (ang > 1e-10) {
            const V3 axis = th / ang;
            return qnormalized(mul(from_angle_axis(ang, axis), q));
        }
        return q;
    }

    // ---------------- CPU adjoint
    // The adjoint of recorded steps, with every derivative taken by central differences
    // of this file's own functions (objective gradient, joint constraint, quaternion
    // write-back): an FP64 check of the device adjoint that shares none of its
    // derivative code. Per step, in reverse: the write-back's cotangent gives x-bar of
    // the last solve; each solve a (last to first) adds the multiplier update's
    // mu_a (dC/dx)^T lambda-bar, solves A^T eta = x-bar with A = dg/dx at x*_a (dense,
    // by differences, LU with partial pivoting), then sends -eta^T dg/dz to every input z
    // of g_a: x_hat (step-start p, v, w), masses, step-start poses (friction frame, joint
    // reference rotations), multipliers and the friction coefficient. A solve whose
    // Newton loop accepted no step (NewtonSolver returned at its first iteration, with or
    // without convergence) is the identity on its starting point, so its x-bar passes to
    // the previous solve's x* (or to x_hat). Contact points, normals and lagged normal
    // forces follow the state the narrow phase ran at (x*_{a-1}, or the step-start poses
    // for the first solve); the contact sets, the barrier stiffness and the hardening are
    // held fixed, as on the device.
    void adj_load(const AdjStep& st, const AdjSolve& so) {
        bodies = st.bodies;
        x_hat = st.x_hat;
        idt = st.adt;
        joints = so.joints;
        b_contacts = so.b_contacts;
        f_contacts = so.f_contacts;
        f_normal = so.f_normal;
        b_dhat = so.b_dhat; b_kappa = so.b_kappa;
        f_mu = so.f_mu; f_eps = so.f_eps; f_dt = so.f_dt;
        has_barrier = so.has_barrier; has_friction = so.has_friction;
    }
    // step-start pose of body i, which the friction frame and the joint references read
    void set_start_pose(int i, const V3& p, const Q& q) {
        bodies[i].p = p;
        bodies[i].q = q;
        for (auto& j : joints) {
            if (j.a == i) { j.pA_ref = p; j.qA_ref = q; }
            if (j.b == i) { j.pB_ref = p; j.qB_ref = q; }
        }
    }
    double eta_dot_grad(const vector<double>& x, const vector<double>& eta) const {
        vector<double> g;
        obj_grad(x, g);
        double s = 0.0;
        for (size_t i = 0; i < g.size(); ++i) s += eta[i] * g[i];
        return s;
    }
    // x = M^-1 b, M dense row-major n x n (LU with partial pivoting); false if singular
    static bool lu_solve(vector<double> M, vector<double>& b, int n) {
        vector<int> piv(n);
        for (int k = 0; k < n; ++k) {
            int p = k;
            for (int i = k + 1; i < n; ++i)
                if (std::abs(M[size_t(i) * n + k]) > std::abs(M[size_t(p) * n + k])) p = i;
            if (!(std::abs(M[size_t(p) * n + k]) > 0.0)) return false;
            if (p != k) {
                for (int c = 0; c < n; ++c) std::swap(M[size_t(k) * n + c], M[size_t(p) * n + c]);
                std::swap(b[k], b[p]);
            }
            const double d = M[size_t(k) * n + k];
            for (int i = k + 1; i < n; ++i) {
                const double f = M[size_t(i) * n + k] / d;
                if (f == 0.0) continue;
                M[size_t(i) * n + k] = f;
                for (int c = k + 1; c < n; ++c) M[size_t(i) * n + c] -= f * M[size_t(k) * n + c];
                b[i] -= f * b[k];
            }
        }
        for (int k = n - 1; k >= 0; --k) {
            double s = b[k];
            for (int c = k + 1; c < n; ++c) s -= M[size_t(k) * n + c] * b[c];
            b[k] = s / M[size_t(k) * n + k];
        }
        return true;
    }
    // central difference of f at z with step h*max(1, |z|); f reads z by reference
    template <typename F>
    static double cdiff(double& z, double rel, F&& f) {
        const double z0 = z, h = rel * std::max(1.0, std::abs(z0));
        z = z0 + h;
        const double fp = f();
        z = z0 - h;
        const double fm = f();
        z = z0;
        return (fp - fm) / (2.0 * h);
    }
    // Cotangents in: P (3N), Qb (4N), V (3N), W (3N) of the end state. Out: the same of
    // the start state, gm (N) and gmu. Returns false if a transposed solve is singular.
    bool adjoint_sweep(vector<AdjStep>& rec, vector<double>& P, vector<double>& Qb, vector<double>& V,
                       vector<double>& W, vector<double>& gm, double& gmu) {
        const int N = n(), nd = 6 * N;
        const double rel = 1e-7;
        const int J = int(joints.size());
        vector<double> lbar(6 * size_t(J), 0.0);
        gm.assign(N, 0.0);
        gmu = 0.0;
        for (int k = int(rec.size()) - 1; k >= 0; --k) {
            AdjStep& st = rec[k];
            const double h = st.adt;
            const int last = int(st.solves.size()) - 1;
            vector<double> Pn(3 * N, 0.0), Qn(4 * N, 0.0), Vn(3 * N, 0.0), Wn(3 * N, 0.0), xbar(nd, 0.0);
            vector<double> carry(nd, 0.0);  // cotangent of the previous solve's x* (its warm start)
            const vector<double>& xf = st.solves[last].x_star;
            for (int i = 0; i < N; ++i) {
                for (int c = 0; c < 3; ++c) {
                    xbar[6 * i + c] = P[3 * i + c] + V[3 * i + c] / h;
                    Pn[3 * i + c] -= V[3 * i + c] / h;
                    xbar[6 * i + 3 + c] = W[3 * i + c] / h;
                }
                // q' = write_back_q(q, theta)
                double in[7] = {xf[6 * i + 3], xf[6 * i + 4], xf[6 * i + 5], st.bodies[i].q.w, st.bodies[i].q.x,
                                st.bodies[i].q.y, st.bodies[i].q.z};
                for (int t = 0; t < 7; ++t) {
                    double dq[4];
                    for (int r = 0; r < 4; ++r) {
                        dq[r] = cdiff(in[t], rel, [&] {
                            const Q o = write_back_q(Q(in[3], in[4], in[5], in[6]), V3(in[0], in[1], in[2]));
                            const double oo[4] = {o.w, o.x, o.y, o.z};
                            return oo[r];
                        });
                    }
                    double acc = 0.0;
                    for (int r = 0; r < 4; ++r) acc += Qb[4 * i + r] * dq[r];
                    if (t < 3) xbar[6 * i + 3 + t] += acc;
                    else Qn[4 * i + t - 3] += acc;
                }
            }
            for (int a = last; a >= 0; --a) {
                AdjSolve& so = st.solves[a];
                adj_load(st, so);

                vector<double> xs = so.x_star;
                if (a != last) xbar = carry;
                carry.assign(nd, 0.0);
                // multiplier update lambda_a = lambda_{a-1} + mu_post C(x*_a)
                for (int j = 0; j < J; ++j) {
                    bool any = false;
                    for (int r = 0; r < 6; ++r) any = any || lbar[6 * j + r] != 0.0;
                    if (!any) continue;
                    Joint& jt = joints[j];
                    const double mp = so.mu_post[j];
                    auto lc = [&] {
                        double C[6];
                        joint_C(jt, xs, C);
                        double t = 0.0;
                        for (int r = 0; r < 6; ++r) t += lbar[6 * j + r] * C[r];
                        return t;
                    };
                    for (int g : {jt.gA, jt.gB}) {
                        if (g < 0) continue;
                        for (int c = 0; c < 6; ++c) xbar[g + c] += mp * cdiff(xs[g + c], rel, lc);
                    }
                    for (int side = 0; side < 2; ++side) {
                        const int bi = side == 0 ? jt.a : jt.b;
                        if (bi < 0) continue;
                        Q& qr = side == 0 ? jt.qA_ref : jt.qB_ref;
                        double* qc[4] = {&qr.w, &qr.x, &qr.y, &qr.z};
                        for (int r = 0; r < 4; ++r) Qn[4 * bi + r] += mp * cdiff(*qc[r], rel, lc);
                    }
                }
                // a solve that accepted no Newton step returned its starting point: the
                // previous solve's x*, or x_hat for the first
                if (so.newton_iters == 0) {
                    if (a == 0) {
                        for (int i = 0; i < nd; ++i) {
                            const int b = i / 6, c = i % 6;
                            if (c < 3) {
                                Pn[3 * b + c] += xbar[i];
                                Vn[3 * b + c] += h * xbar[i];
                            } else {
                                Wn[3 * b + c - 3] += h * xbar[i];
                            }
                        }
                    } else {
                        carry = xbar;
                    }
                    continue;
                }
                // A = dg/dx at x*, by columns; solve A^T eta = xbar
                vector<double> At(size_t(nd) * nd);  // row c of At = column c of A
                for (int c = 0; c < nd; ++c) {
                    const double x0 = xs[c], hh = rel * std::max(1.0, std::abs(x0));
                    vector<double> gp, gmi;
                    xs[c] = x0 + hh;
                    obj_grad(xs, gp);
                    xs[c] = x0 - hh;
                    obj_grad(xs, gmi);
                    xs[c] = x0;
                    for (int r = 0; r < nd; ++r) At[size_t(c) * nd + r] = (gp[r] - gmi[r]) / (2.0 * hh);
                }
                vector<double> eta = xbar;
                if (!lu_solve(At, eta, nd)) return false;
                auto s_eta = [&] { return eta_dot_grad(xs, eta); };
                for (int i = 0; i < nd; ++i) {
                    const double xh = -cdiff(x_hat[i], rel, s_eta);
                    const int b = i / 6, c = i % 6;
                    if (c < 3) {
                        Pn[3 * b + c] += xh;
                        Vn[3 * b + c] += h * xh;
                    } else {
                        Wn[3 * b + c - 3] += h * xh;
                    }
                }
                for (int b = 0; b < N; ++b) gm[b] -= cdiff(bodies[b].mass, rel, s_eta);
                // step-start poses: bodies in a friction contact or a joint
                vector<char> touched(N, 0);
                if (has_friction)
                    for (auto& c : f_contacts) touched[c.A] = touched[c.B] = 1;
                for (auto& jt : joints) {
                    if (jt.a >= 0) touched[jt.a] = 1;
                    if (jt.b >= 0) touched[jt.b] = 1;
                }
                for (int b = 0; b < N; ++b) {
                    if (!touched[b]) continue;
                    double pq[7] = {bodies[b].p.x, bodies[b].p.y, bodies[b].p.z, bodies[b].q.w, bodies[b].q.x,
                                    bodies[b].q.y, bodies[b].q.z};
                    for (int t = 0; t < 7; ++t) {
                        const double d = cdiff(pq[t], rel, [&] {
                            set_start_pose(b, V3(pq[0], pq[1], pq[2]), Q(pq[3], pq[4], pq[5], pq[6]));
                            return s_eta();
                        });
                        set_start_pose(b, V3(pq[0], pq[1], pq[2]), Q(pq[3], pq[4], pq[5], pq[6]));
                        if (t < 3) Pn[3 * b + t] -= d;
                        else Qn[4 * b + t - 3] -= d;
                    }
                }
                for (int j = 0; j < J; ++j)
                    for (int r = 0; r < 6; ++r) lbar[6 * j + r] -= cdiff(joints[j].lambda[r], rel, s_eta);
                if (has_friction) gmu -= cdiff(f_mu, rel, s_eta);
                // The solve's contact points and normals, and (a >= 1) its lagged normal
                // forces, are functions of the state the narrow phase ran at: x*_{a-1} for
                // a >= 1 (detectContactsAtState + computeNormalForces, IPCSolver.cpp:
                // 1112-1130), the step-start poses for a == 0 (detectContacts, :394-580).
                // Each recorded contact keeps its primitives (the contact set is discrete
                // and held fixed) and its geometry is recomputed at the perturbed state.
                if ((has_barrier || has_friction) && !b_contacts.empty()) {
                    const vector<Contact> bc0 = b_contacts, fc0 = f_contacts;
                    const vector<double> fn0 = f_normal;
                    // the clamp side each contact was recorded on
                    vector<signed char> cx(bc0.size()), cn(bc0.size());
                    for (size_t i = 0; i < bc0.size(); ++i) {
                        const Contact& c = bc0[i];
                        cx[i] = dot(seg3(xs, c.B * 6) - seg3(xs, c.A * 6), c.normal) > 1e-9 ? 0 : 1;
                        cn[i] = 1;
                        if (a >= 1) {
                            const vector<double>& xq = st.solves[a - 1].x_star;
                            cn[i] = dot(seg3(xq, c.B * 6) - seg3(xq, c.A * 6), c.normal) > 1e-9 ? 0 : 1;
                        }
                    }
                    auto use = [&](const vector<double>* xs_prev) {
                        vector<Contact> ic = bc0;
                        for (size_t i = 0; i < ic.size(); ++i) {
                            Contact& c = ic[i];
                            contact_geom(c, xs_prev, c.point, c.normal);
                            c.clamp_x = cx[i];
                            c.clamp_n = cn[i];
                        }
                        b_contacts = ic;
                        f_contacts = ic;
                        if (xs_prev && has_friction)
                            f_normal = has_barrier ? bar_normal_forces(ic, *xs_prev, b_dhat, b_kappa)
                                                   : vector<double>(ic.size(), 0.0);
                    };
                    vector<char> inc(N, 0);
                    for (auto& c : bc0) inc[c.A] = inc[c.B] = 1;
                    if (a >= 1) {
                        vector<double> xp = st.solves[a - 1].x_star;
                        auto s_geom = [&] {
                            use(&xp);
                            return s_eta();
                        };
                        for (int b = 0; b < N; ++b)
                            if (inc[b])
                                for (int c = 0; c < 6; ++c)
                                    carry[6 * b + c] -= cdiff(xp[6 * b + c], rel, s_geom);
                    } else {
                        for (int b = 0; b < N; ++b) {
                            if (!inc[b]) continue;
                            const Body b0 = bodies[b];
                            double pq[7] = {b0.p.x, b0.p.y, b0.p.z, b0.q.w, b0.q.x, b0.q.y, b0.q.z};
                            for (int t = 0; t < 7; ++t) {
                                const double d = cdiff(pq[t], rel, [&] {
                                    bodies[b].p = V3(pq[0], pq[1], pq[2]);
                                    bodies[b].q = Q(pq[3], pq[4], pq[5], pq[6]);
                                    use(nullptr);
                                    bodies[b] = b0;
                                    return s_eta();
                                });
                                if (t < 3) Pn[3 * b + t] -= d;
                                else Qn[4 * b + t - 3] -= d;
                            }
                        }
                    }
                    b_contacts = bc0;
                    f_contacts = fc0;
                    f_normal = fn0;
                }
            }
            P = Pn; Qb = Qn; V = Vn; W = Wn;
        }
        return true;
    }
};

// ---------------------------------------------------------------------------------
// CCD (RigidTrajectoryAABB.h, IntervalRootFinder.h, TimeOfImpact.cpp, CCD.h)
struct Pose {
    V3 p0, p1;
    Q q0, q1;
};
struct I3 {
    I x, y, z;
};

// RigidTrajectoryAABB.h:37-73
static I3 vertex_traj(const Pose& ps, const V3& lv, const I& t) {
    const Q qlo = slerp(ps.q0, t.lo, ps.q1), qhi = slerp(ps.q0, t.hi, ps.q1);
    const V3 a = rotate(qlo, lv), b = rotate(qhi, lv);
    const I3 rot{I(std::min(a.x, b.x), std::max(a.x, b.x)), I(std::min(a.y, b.y), std::max(a.y, b.y)),
                 I(std::min(a.z, b.z), std::max(a.z, b.z))};
    const I3 pos{I(std::min(ps.p0.x, ps.p1.x), std::max(ps.p0.x, ps.p1.x)),
                 I(std::min(ps.p0.y, ps.p1.y), std::max(ps.p0.y, ps.p1.y)),
                 I(std::min(ps.p0.z, ps.p1.z), std::max(ps.p0.z, ps.p1.z))};
    return {pos.x + rot.x, pos.y + rot.y, pos.z + rot.z};
}
// RigidTrajectoryAABB.h:75-94
static I3 edge_traj(const Pose& ps, const V3& a, const V3& b, const I& t, const I& al) {
    const I3 v0 = vertex_traj(ps, a, t), v1 = vertex_traj(ps, b, t);
    const I A(al.lo, al.hi);
    return {v0.x * (1.0 - A) + v1.x * A, v0.y * (1.0 - A) + v1.y * A, v0.z * (1.0 - A) + v1.z * A};
}
// RigidTrajectoryAABB.h:96-120
static I3 face_traj(const Pose& ps, const V3& a, const V3& b, const V3& c, const I& t, const I& u, const I& v) {
    const I3 v0 = vertex_traj(ps, a, t), v1 = vertex_traj(ps, b, t), v2 = vertex_traj(ps, c, t);
    const I one(1.0);
    const I ui(u.lo, u.hi), vi(v.lo, v.hi);
    const I wi = one - ui - vi;
    return {wi * v0.x + ui * v1.x + vi * v2.x, wi * v0.y + ui * v1.y + vi * v2.y, wi * v0.z + ui * v1.z + vi * v2.z};
}

// ipc mode (query kinds 3 and 4, no reference counterpart). R(t) r over the box: the
// slerp-endpoint hull widened by the arc's largest distance from its chord,
// |r| (1 - cos(phi / 2)) with cos(phi / 2) = |<q(t.lo), q(t.hi)>|, plus 8e-15 |r| for
// the rounding of the endpoint rotations.
static I3 rot_hull(const Pose& ps, const V3& r, const I& t) {
    const Q qlo = slerp(ps.q0, t.lo, ps.q1), qhi = slerp(ps.q0, t.hi, ps.q1);
    const V3 a = rotate(qlo, r), b = rotate(qhi, r);
    const double cd = std::min(1.0, std::abs(((qlo.w * qhi.w + qlo.x * qhi.x) + qlo.y * qhi.y) + qlo.z * qhi.z));
    const double rl = std::sqrt((r.x * r.x + r.y * r.y) + r.z * r.z);
    const double e = rl * (1.0 - cd) + 8e-15 * rl;
    return {I(sub_dn(std::min(a.x, b.x), e), add_up(std::max(a.x, b.x), e)),
            I(sub_dn(std::min(a.y, b.y), e), add_up(std::max(a.y, b.y), e)),
            I(sub_dn(std::min(a.z, b.z), e), add_up(std::max(a.z, b.z), e))};
}
// p_A(t) - p_B(t) over the box, both translations linear in t
static I rel_trans1(double a0, double a1, double b0, double b1, const I& t) {
    const I pa0(a0), pb0(b0);
    return (pa0 - pb0) + t * ((I(a1) - pa0) - (I(b1) - pb0));
}
static I3 rel_trans(const Pose& A, const Pose& B, const I& t) {
    return {rel_trans1(A.p0.x, A.p1.x, B.p0.x, B.p1.x, t), rel_trans1(A.p0.y, A.p1.y, B.p0.y, B.p1.y, t),
            rel_trans1(A.p0.z, A.p1.z, B.p0.z, B.p1.z, t)};
}

struct Box {
    I d[3];
};

// IntervalRootFinder.h:20-114 (capped DFS, max 1000 pops). earliest_first (the ipc
// mode's kinds 3 and 4, no reference counterpart): the low half is popped first, and a
// search stopped by the cap reports an impact at the least t.lo of the best leaf and of
// every box still on the stack (a lower bound of any impact it did not rule out).
static bool root_finder(int D, const std::function<I3(const Box&)>& f, const std::function<bool(const Box&)>& valid,
                        const Box& x0, const double* tol, Box& out, int* pops, int max_it = 1000,
                        bool earliest_first = false) {
    Box earliest;
    for (int i = 0; i < D; ++i) earliest.d[i] = I(std::numeric_limits<double>::infinity());
    bool found = false;
    std::stack<Box> st;
    st.push(x0);
    Box xt;
    for (int i = 0; i < D; ++i) xt.d[i] = I(0, tol[i]);
    (void)f(xt);  // pre-check: evaluated, result unused (IntervalRootFinder.h:41-48)
    int it = 0;
    for (; !st.empty() && it < max_it; ++it) {
        Box x = st.top();
        st.pop();
        if (x.d[0].lo >= earliest.d[0].lo) continue;
        if (!valid(x)) continue;
        const I3 y = f(x);
        if (!(zero_in(y.x) && zero_in(y.y) && zero_in(y.z))) continue;
        bool within = true;
        for (int i = 0; i < D; ++i)
            if (width(x.d[i]) > tol[i]) { within = false; break; }
        if (within) {
            if (x.d[0].lo < earliest.d[0].lo) { earliest = x; found = true; }
            continue;
        }
        int sd = 0;
        double mw = width(x.d[0]);
        for (int i = 1; i < D; ++i) {
            const double w = width(x.d[i]);
            if (w > mw) { mw = w; sd = i; }
        }
        const double mid = median(x.d[sd]);
        Box lo = x, hi = x;
        lo.d[sd] = I(x.d[sd].lo, mid);
        hi.d[sd] = I(mid, x.d[sd].hi);
        if (earliest_first) {
            st.push(hi);
            st.push(lo);
        } else {
            st.push(lo);
            st.push(hi);
        }
    }
    if (pops) *pops = it;
    if (earliest_first && it >= max_it && !st.empty()) {
        double m = earliest.d[0].lo;
        for (; !st.empty(); st.pop()) m = std::min(m, st.top().d[0].lo);
        earliest.d[0] = I(m, m);
        found = true;
    }
    if (found) out = earliest;
    return found;
}

static Pose make_pose(const double* ps) {
    Pose p;
    p.p0 = V3(ps[0], ps[1], ps[2]);
    p.q0 = Q(ps[3], ps[4], ps[5], ps[6]);
    p.p1 = V3(ps[7], ps[8], ps[9]);
    p.q1 = Q(ps[10], ps[11], ps[12], ps[13]);
    return p;
}

// computeEdgeVertexTOIImpl / computeEdgeEdgeTOIImpl / computeFaceVertexTOIImpl
// (TimeOfImpact.cpp:8-150), kinds 0-2. Body A carries the vertex (or edge A), body B the
// edge/face. Kinds 3 (face-vertex) and 4 (edge-edge) are the ipc mode's: the
// separation written in body-relative form (relative translation linear in t, each
// rotated body-frame vector and edge vector as its rot_hull), the earliest-first
// search, and for 3 a box is kept when it meets the triangle u, v >= 0, u + v <= 1 (u
// and v stay inside [0, 1], so: upper(1 - u.lo - v.lo) >= 0) where the reference keeps
// only boxes whose weights are all certainly inside [0, 1].
static int ccd_toi(int kind, const Pose& A, const Pose& B, const V3* a, const V3* b, double earliest, double tol,
                   double* toi, int* pops) {
    g_interval_nan = false;
    Box x0, out;
    bool hit = false;
    if (kind == 0) {
        const double tl[2] = {tol, 1.0};
        x0.d[0] = I(0, earliest);
        x0.d[1] = I(0, 1);
        auto f = [&](const Box& x) {
            const I3 v = vertex_traj(A, a[0], x.d[0]);
            const I3 e = edge_traj(B, b[0], b[1], x.d[0], x.d[1]);
            return I3{v.x - e.x, v.y - e.y, v.z - e.z};
        };
        hit = root_finder(2, f, [](const Box&) { return true; }, x0, tl, out, pops);
    } else if (kind == 1) {
        const double tl[3] = {tol, 1.0, 1.0};
        x0.d[0] = I(0, earliest);
        x0.d[1] = I(0, 1);
        x0.d[2] = I(0, 1);
        auto f = [&](const Box& x) {
            const I3 ea = edge_traj(A, a[0], a[1], x.d[0], x.d[1]);
            const I3 eb = edge_traj(B, b[0], b[1], x.d[0], x.d[2]);
            return I3{ea.x - eb.x, ea.y - eb.y, ea.z - eb.z};
        };
        hit = root_finder(3, f, [](const Box&) { return true; }, x0, tl, out, pops);
    } else if (kind == 4) {
        // (p_A - p_B) + R_A a0 + alpha R_A (a1 - a0) - (R_B b0 + beta R_B (b1 - b0))
        const double tl[3] = {tol, 1.0, 1.0};
        x0.d[0] = I(0, earliest);
        x0.d[1] = I(0, 1);
        x0.d[2] = I(0, 1);
        const V3 ea = a[1] - a[0], eb = b[1] - b[0];
        auto f = [&](const Box& x) {
            const I& t = x.d[0];
            const I3 tr = rel_trans(A, B, t), ra = rot_hull(A, a[0], t), rea = rot_hull(A, ea, t);
            const I3 rb = rot_hull(B, b[0], t), reb = rot_hull(B, eb, t);
            const I& al = x.d[1];
            const I& be = x.d[2];
            return I3{((tr.x + ra.x) + al * rea.x) - (rb.x + be * reb.x),
                      ((tr.y + ra.y) + al * rea.y) - (rb.y + be * reb.y),
                      ((tr.z + ra.z) + al * rea.z) - (rb.z + be * reb.z)};
        };
        hit = root_finder(3, f, [](const Box&) { return true; }, x0, tl, out, pops, 1000, true);
    } else if (kind == 3) {
        // (p_A - p_B) + R_A a - R_B b0 - (u R_B (b1 - b0) + v R_B (b2 - b0))
        const double tl[3] = {tol, 1.0, 1.0};
        x0.d[0] = I(0, earliest);
        x0.d[1] = I(0, 1);
        x0.d[2] = I(0, 1);
        const V3 e1 = b[1] - b[0], e2 = b[2] - b[0];
        auto f = [&](const Box& x) {
            const I& t = x.d[0];
            const I3 tr = rel_trans(A, B, t), ra = rot_hull(A, a[0], t), rb = rot_hull(B, b[0], t);
            const I3 r1 = rot_hull(B, e1, t), r2 = rot_hull(B, e2, t);
            const I& u = x.d[1];
            const I& v = x.d[2];
            return I3{((tr.x + ra.x) - rb.x) - (u * r1.x + v * r2.x), ((tr.y + ra.y) - rb.y) - (u * r1.y + v * r2.y),
                      ((tr.z + ra.z) - rb.z) - (u * r1.z + v * r2.z)};
        };
        auto valid = [](const Box& x) {
            const I u = x.d[1], v = x.d[2];
            const I w = I(1.0) - u - v;
            return u.hi >= 0 && v.hi >= 0 && w.hi >= 0;
        };
        hit = root_finder(3, f, valid, x0, tl, out, pops, 1000, true);
    } else {
        const double tl[3] = {tol, 1.0, 1.0};
        x0.d[0] = I(0, earliest);
        x0.d[1] = I(0, 1);
        x0.d[2] = I(0, 1);
        auto f = [&](const Box& x) {
            const I3 v = vertex_traj(A, a[0], x.d[0]);
            const I3 fp = face_traj(B, b[0], b[1], b[2], x.d[0], x.d[1], x.d[2]);
            return I3{v.x - fp.x, v.y - fp.y, v.z - fp.z};
        };
        auto valid = [](const Box& x) {
            const I u = x.d[1], v = x.d[2];
            const I w = I(1.0) - u - v;
            return u.lo >= 0 && u.hi <= 1 && v.lo >= 0 && v.hi <= 1 && w.lo >= 0 && w.hi <= 1;
        };
        hit = root_finder(3, f, valid, x0, tl, out, pops);
    }
    if (g_interval_nan) return -1;
    *toi = hit ? out.d[0].lo : std::numeric_limits<double>::infinity();
    return hit ? 1 : 0;
}

// ipc mode, OptimizationProblem::maxStep (NewtonSolver.h:42, NewtonSolver.cpp:53-54):
// per pair, conservative advancement (nd::ipc::pair_advance) along the line-search path
// x + t dx (positions p + t dp, orientations apply_theta(q, theta + t dtheta)), keeping
// sep of the pair's current distance; mu = sum over the two bodies of |dp| + |dtheta|
// r_max (r_max: the largest body-frame radius of the pair's vertices on that body, which
// bounds their speed along the path), margin = 2e-10 (r_max A + r_max B) for the
// rotation skipped below |theta| = 1e-10. toi = the least over the pairs; alpha = toi *
// safety when toi < 1 (the form of IPCSolver::computeMaxStep, IPCSolver.cpp:385-391),
// else 1.
int Solver::ipc_max_step(const int32_t* pairs, int np, const vector<double>& x, const vector<double>& dx, double sep,
                         double safety, double& toi, double& alpha, int& fail_index) const {
    for (int i = 0; i < np; ++i)
        if (!ipc_pair_ok(pairs + 5 * i)) {
            fail_index = i;
            return 1;
        }
    toi = std::numeric_limits<double>::infinity();
    for (int i = 0; i < np; ++i) {
        const int32_t* pr = pairs + 5 * i;
        const int kind = pr[0];
        int vb[4];
        V3 r[4];
        const Mesh& MA = meshes[bodies[pr[1]].mesh];
        const Mesh& MB = meshes[bodies[pr[2]].mesh];
        if (kind == 0) {
            vb[0] = pr[1];
            r[0] = MA.V[pr[3]];
            for (int k = 0; k < 3; ++k) { vb[1 + k] = pr[2]; r[1 + k] = MB.V[MB.F[pr[4]][k]]; }
        } else {
            for (int k = 0; k < 2; ++k) {
                vb[k] = pr[1];
                r[k] = MA.V[MA.E[pr[3]][k]];
                vb[2 + k] = pr[2];
                r[2 + k] = MB.V[MB.E[pr[4]][k]];
            }
        }
        double rmax[2] = {0.0, 0.0};
        for (int v = 0; v < 4; ++v) {
            const int side = vb[v] == pr[1] ? 0 : 1;
            rmax[side] = std::max(rmax[side], std::sqrt((r[v].x * r[v].x + r[v].y * r[v].y) + r[v].z * r[v].z));
        }
        double mu = 0.0;
        for (int side = 0; side < 2; ++side) {
            const size_t o = 6 * size_t(pr[1 + side]);
            const double lp = std::sqrt((dx[o] * dx[o] + dx[o + 1] * dx[o + 1]) + dx[o + 2] * dx[o + 2]);
            const double lt = std::sqrt((dx[o + 3] * dx[o + 3] + dx[o + 4] * dx[o + 4]) + dx[o + 5] * dx[o + 5]);
            mu = mu + (lp + lt * rmax[side]);
        }
        const double margin = 2e-10 * (rmax[0] + rmax[1]);
        auto world = [&](double t, double* xw) {
            for (int v = 0; v < 4; ++v) {
                const size_t o = 6 * size_t(vb[v]);
                const V3 p(x[o] + t * dx[o], x[o + 1] + t * dx[o + 1], x[o + 2] + t * dx[o + 2]);
                const V3 th(x[o + 3] + t * dx[o + 3], x[o + 4] + t * dx[o + 4], x[o + 5] + t * dx[o + 5]);
                const V3 w = p + rotate(apply_theta(bodies[vb[v]].q, th), r[v]);
                xw[3 * v] = w.x;
                xw[3 * v + 1] = w.y;
                xw[3 * v + 2] = w.z;
            }
        };
        const double t = nd::ipc::pair_advance(kind, world, mu, sep, margin, kIpcAdvanceIters,
                                               [](double) { return false; });
        toi = std::min(toi, t);
    }
    alpha = toi < 1.0 ? toi * safety : 1.0;
    return 0;
}

// Block-Jacobi PCG over 6x6 blocks in the device solver's order of operations: every dot
// product as 256 strided partial sums (lane t takes i = t, t + 256, ...) combined by a
// halving tree; each block row's product sums its diagonal block, then the off-diagonal
// blocks in list order (a block (i, j) also serves row j transposed); block-Jacobi
// preconditioner by the 6x6 Cholesky inverse.
static bool inv6_o(const double* A, double* out) {
    double L[36];
    for (int k = 0; k < 36; ++k) L[k] = 0.0;
    for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
        if (!(s > 0.0)) return false;
        const double dj = std::sqrt(s);
        L[j * 6 + j] = dj;
        for (int i = j + 1; i < 6; ++i) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = t / dj;
        }
    }
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

static int pcg_ordered(int nb, const vector<double>& diag, const vector<int>& offij, const vector<double>& offb,
                       const vector<double>& b, double tol, int max_it, vector<double>& x) {
    const int nd = 6 * nb, noff = int(offij.size() / 2), T = 256;
    auto dot = [&](const vector<double>& a, const vector<double>& c) {
        double part[256];
        for (int t = 0; t < T; ++t) {
            double s = 0.0;
            for (int i = t; i < nd; i += T) s += a[i] * c[i];
            part[t] = s;
        }
        for (int h = T / 2; h > 0; h >>= 1)
            for (int t = 0; t < h; ++t) part[t] = part[t] + part[t + h];
        return part[0];
    };
    vector<double> Dinv(36 * size_t(nb), 0.0), r(nd), z(nd), p(nd), Ap(nd);
    for (int i = 0; i < nb; ++i) inv6_o(&diag[36 * size_t(i)], &Dinv[36 * size_t(i)]);
    auto mv = [&](const vector<double>& v, vector<double>& y) {
        for (int i = 0; i < nb; ++i)
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int c = 0; c < 6; ++c) s += diag[36 * size_t(i) + a * 6 + c] * v[6 * i + c];
                for (int o = 0; o < noff; ++o) {
                    const int bi = offij[2 * o], bj = offij[2 * o + 1];
                    const double* B = &offb[36 * size_t(o)];
                    if (bi == i) for (int c = 0; c < 6; ++c) s += B[a * 6 + c] * v[6 * bj + c];
                    if (bj == i) for (int c = 0; c < 6; ++c) s += B[c * 6 + a] * v[6 * bi + c];
                }
                y[6 * i + a] = s;
            }
    };
    auto prec = [&](const vector<double>& v, vector<double>& y) {
        for (int i = 0; i < nb; ++i)
            for (int a = 0; a < 6; ++a) {
                double s = 0.0;
                for (int c = 0; c < 6; ++c) s += Dinv[36 * size_t(i) + a * 6 + c] * v[6 * i + c];
                y[6 * i + a] = s;
            }
    };
    x.assign(nd, 0.0);
    r = b;
    prec(r, z);
    p = z;
    double rz = dot(r, z);
    const double bn = std::sqrt(dot(b, b));
    double rn = bn;
    int it = 0;
    for (; it < max_it && rn > tol * bn; ++it) {
        mv(p, Ap);
        const double pAp = dot(p, Ap);
        const double alpha = rz / pAp;
        for (int i = 0; i < nd; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        rn = std::sqrt(dot(r, r));
        prec(r, z);
        const double rz2 = dot(r, z);
        const double beta = rz2 / rz;
        rz = rz2;
        for (int i = 0; i < nd; ++i) p[i] = z[i] + beta * p[i];
    }
    return it;
}

// ipc mode time step: IPCSolver::step's implicit-Euler Newton solve (IPCSolver.cpp:137-262,
// NewtonSolver.cpp:10-95) with the ipc barrier in place of the centre-distance barrier and
// maxStep wired into the line search, for scenes whose only joints pin single bodies
// (those bodies are static: their DOFs are held). Objective, as the reference's:
// sum over moving bodies of 1/2 m |x_p - x^_p|^2 + 1/2 dtheta^T I dtheta (InertiaForm,
// body-frame I), + h^2 (-sum m g.x_p (GravityForm) + E_ipc(x)), x^ = [p + h v, h w].
// Pairs: the top-K lists of the close pairs at inflation P.inflation (ipc_pairs_from_lists),
// fixed for the step. Newton: H dx = -g by block-Jacobi PCG (pcg_ordered; diagonal blocks:
// inertia then h^2 x each segment's block in segment order; one off-diagonal block per
// segment, zero when a body is static), stop when |dx|_inf <= newton_tol or the direction
// does not descend; alpha starts at min(1, maxStep) (ipc_max_step, sep, safety) and halves
// until Armijo (c = 1e-4) holds. Sums run in body, segment and pair order. Write-back as
// IPCSolver::step's: p = x_p, v = (x_p - p) / h, w = x_theta / h, q = apply_theta(q, x_theta).
int Solver::ipc_step(double h, const IpcStepParams& P, int32_t* stats) {
    const int N = n(), nd = 6 * N;
    vector<char> stat(N, 0);
    for (auto& j : joints) {
        if (j.type != 0 || j.b >= 0 || j.a < 0) return -1;
        stat[j.a] = 1;
    }
    assembler_build();
    vector<std::array<int32_t, 5>> pr;
    ipc_pairs_from_lists(P.inflation, stat, pr);
    const int np = int(pr.size());
    vector<int32_t> flat(5 * size_t(np));
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < 5; ++k) flat[5 * size_t(i) + k] = pr[i][k];
    const double h2 = h * h;
    vector<double> x(nd, 0.0), xh(nd, 0.0);
    for (int b = 0; b < N; ++b) {
        const Body& B = bodies[b];
        x[6 * b] = B.p.x; x[6 * b + 1] = B.p.y; x[6 * b + 2] = B.p.z;
        xh[6 * b] = B.p.x + h * B.v.x; xh[6 * b + 1] = B.p.y + h * B.v.y; xh[6 * b + 2] = B.p.z + h * B.v.z;
        xh[6 * b + 3] = h * B.w.x; xh[6 * b + 4] = h * B.w.y; xh[6 * b + 5] = h * B.w.z;
    }
    auto energy = [&](const vector<double>& xx, double* e_ipc_out, vector<IpcSeg>* segs_out) {
        double si = 0.0, sg = 0.0;
        for (int b = 0; b < N; ++b) {
            if (stat[b]) continue;
            double e, g6[6], H36[36];
            nd::ipc::body_terms(bodies[b].mass, &bodies[b].I.m[0][0], &xx[6 * b], &xh[6 * b], e, g6, H36, &sg);
            si += e;
        }
        double ev = 0.0;
        vector<IpcSeg> segs;
        ipc_barrier_eval(flat.data(), np, xx, P.dhat, P.kappa, true, ev, segs);
        if (e_ipc_out) *e_ipc_out = ev;
        if (segs_out) *segs_out = std::move(segs);
        return (si + h2 * sg) + h2 * ev;
    };
    int newton = 0, trials = 0, pcg_total = 0;
    for (int it = 0; it < P.max_newton; ++it) {
        vector<IpcSeg> segs;
        const double E0 = energy(x, nullptr, &segs);
        // gradient and blocks
        vector<double> g(nd, 0.0), diag(36 * size_t(N), 0.0);
        for (int b = 0; b < N; ++b) {
            double* D = &diag[36 * size_t(b)];
            if (stat[b]) {
                for (int k = 0; k < 6; ++k) D[7 * k] = 1.0;
                continue;
            }
            double e, sgd = 0.0;
            nd::ipc::body_terms(bodies[b].mass, &bodies[b].I.m[0][0], &x[6 * b], &xh[6 * b], e, &g[6 * b], D, &sgd);
            for (int k = 0; k < 3; ++k) g[6 * b + k] += h2 * (-bodies[b].mass * nd::ipc::gravity(k));
            for (const IpcSeg& sgm : segs) {
                const int side = sgm.lo == b ? 0 : (sgm.hi == b ? 1 : -1);
                if (side < 0) continue;
                for (int r = 0; r < 6; ++r) {
                    g[6 * b + r] += h2 * sgm.g[6 * side + r];
                    for (int c = 0; c < 6; ++c) D[6 * r + c] += h2 * sgm.H[12 * (6 * side + r) + 6 * side + c];
                }
            }
        }
        vector<int> offij;
        vector<double> offb;
        for (const IpcSeg& sgm : segs) {
            offij.push_back(sgm.lo);
            offij.push_back(sgm.hi);
            const bool both = !stat[sgm.lo] && !stat[sgm.hi];
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) offb.push_back(both ? h2 * sgm.H[12 * r + 6 + c] : 0.0);
        }
        vector<double> rhs(nd), dx;
        for (int i = 0; i < nd; ++i) rhs[i] = -g[i];
        pcg_total += pcg_ordered(N, diag, offij, offb, rhs, P.pcg_tol, P.pcg_max_it, dx);
        double dmax = 0.0;
        for (int i = 0; i < nd; ++i) dmax = std::max(dmax, std::abs(dx[i]));
        if (!(dmax > P.newton_tol)) break;
        double slope = 0.0;
        for (int i = 0; i < nd; ++i) slope += g[i] * dx[i];
        if (!(slope < 0.0)) break;
        double alpha = 1.0, toi = 0.0;
        int fi = -1;
        ipc_max_step(flat.data(), np, x, dx, P.sep, P.safety, toi, alpha, fi);
        alpha = std::min(1.0, alpha);
        bool ok = false;
        vector<double> xt(nd);
        for (int k = 0; k < P.max_ls; ++k) {
            ++trials;
            for (int i = 0; i < nd; ++i) xt[i] = x[i] + alpha * dx[i];
            const double Et = energy(xt, nullptr, nullptr);
            if (Et <= E0 + 1e-4 * alpha * slope) {
                ok = true;
                break;
            }
            alpha *= 0.5;
        }
        if (!ok) break;
        x = xt;
        ++newton;
    }
    for (int b = 0; b < N; ++b) {
        if (stat[b]) continue;
        Body& B = bodies[b];
        const V3 pn = seg3(x, 6 * b), th = seg3(x, 6 * b + 3);
        B.v = (pn - B.p) / h;
        B.w = th / h;
        B.p = pn;
        B.q = apply_theta(B.q, th);
    }
    if (stats) {
        stats[0] = newton;
        stats[1] = trials;
        stats[2] = np;
        stats[3] = pcg_total;
    }
    return 0;
}

// CCDSystem::computeEarliestTOI (CCD.h:22-82)
double Solver::ccd_earliest_toi(double dt) const {
    if (bodies.empty() || cands.fv.empty()) return 1.0;
    double earliest = 1.0;
    for (auto& c : cands.fv) {
        const int fi = world_index_of_id(c[0]), vi = world_index_of_id(c[1]);
        if (fi < 0 || vi < 0) continue;
        const Body& fb = bodies[fi];
        const Body& vb = bodies[vi];
        if (fb.mesh < 0 || vb.mesh < 0) continue;
        Pose A, B;
        A.p0 = vb.p; A.q0 = vb.q; A.p1 = vb.p + vb.v * dt; A.q1 = vb.q;
        B.p0 = fb.p; B.q0 = fb.q; B.p1 = fb.p + fb.v * dt; B.q1 = fb.q;
        const Mesh& MA = meshes[vb.mesh];
        const Mesh& MB = meshes[fb.mesh];
        const V3 a[1] = {MA.V[c[3]]};
        const auto& f = MB.F[c[2]];
        const V3 b[3] = {MB.V[f[0]], MB.V[f[1]], MB.V[f[2]]};
        double toi = std::numeric_limits<double>::infinity();
        int pops = 0;
        const int r = ccd_toi(2, A, B, a, b, earliest, 1e-8, &toi, &pops);
        if (r == 1 && toi < earliest) earliest = toi;
    }
    return earliest;
}

}  // namespace ndo

// =================================================================================
// C ABI
using namespace ndo;

namespace {
struct Handle {
    Solver s;
    std::vector<Solver::AdjStep> rec;  // ndo_record_steps
    Solver probe;                      // a recorded solve's objective (ndo_record_objective)
};
Solver& S(void* h) { return static_cast<Handle*>(h)->s; }

void contacts_out(const std::vector<Contact>& cs, int32_t cap, int32_t* n, int32_t* ints, double* dbl) {
    *n = int32_t(cs.size());
    for (size_t i = 0; i < cs.size() && int32_t(i) < cap; ++i) {
        const Contact& c = cs[i];
        if (ints) { ints[5 * i] = c.A; ints[5 * i + 1] = c.B; ints[5 * i + 2] = c.type; ints[5 * i + 3] = c.pa; ints[5 * i + 4] = c.pb; }
        if (dbl) {
            dbl[7 * i] = c.point.x; dbl[7 * i + 1] = c.point.y; dbl[7 * i + 2] = c.point.z;
            dbl[7 * i + 3] = c.normal.x; dbl[7 * i + 4] = c.normal.y; dbl[7 * i + 5] = c.normal.z;
            dbl[7 * i + 6] = c.dist;
        }
    }
}
std::vector<Contact> contacts_in(int32_t nc, const int32_t* ints, const double* dbl) {
    std::vector<Contact> cs(nc);
    for (int i = 0; i < nc; ++i) {
        Contact& c = cs[i];
        c.A = ints[5 * i]; c.B = ints[5 * i + 1]; c.type = ints[5 * i + 2]; c.pa = ints[5 * i + 3]; c.pb = ints[5 * i + 4];
        c.point = V3(dbl[7 * i], dbl[7 * i + 1], dbl[7 * i + 2]);
        c.normal = V3(dbl[7 * i + 3], dbl[7 * i + 4], dbl[7 * i + 5]);
        c.dist = dbl[7 * i + 6];
    }
    return cs;
}
void dense_out(const SymAccum& H, double* out) {
    const int n = H.n;
    std::fill(out, out + size_t(n) * n, 0.0);
    for (int r = 0; r < n; ++r)
        for (auto& kv : H.rows[r]) out[size_t(r) * n + kv.first] = kv.second;
}
}  // namespace

extern "C" {

void* ndo_create(const ndo_scene_desc* d, int32_t scene) {
    if (!d || scene < 0 || scene >= d->n_scenes) return nullptr;
    auto* h = new Handle();
    Solver& s = h->s;
    const int b0 = d->scene_body_off[scene], b1 = d->scene_body_off[scene + 1];
    for (int b = b0; b < b1; ++b) {
        Body body;
        body.id = d->body_id[b];
        body.mass = d->mass[b];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) body.I(i, j) = d->inertia[b * 9 + i * 3 + j];
        if (d->n_verts[b] > 0) {
            Mesh m;
            const int vo = d->vert_off[b], nv = d->n_verts[b], fo = d->face_off[b], nf = d->n_faces[b];
            for (int v = 0; v < nv; ++v) m.V.push_back(V3(d->verts[(vo + v) * 3], d->verts[(vo + v) * 3 + 1], d->verts[(vo + v) * 3 + 2]));
            for (int f = 0; f < nf; ++f) m.F.push_back({d->faces[(fo + f) * 3], d->faces[(fo + f) * 3 + 1], d->faces[(fo + f) * 3 + 2]});
            m.E = build_unique_edges(m.F);
            m.lmin = m.V[0];
            m.lmax = m.V[0];
            for (auto& v : m.V) { m.lmin = cmin(m.lmin, v); m.lmax = cmax(m.lmax, v); }  // MeshShape.cpp:6-11
            body.mesh = int(s.meshes.size());
            s.meshes.push_back(std::move(m));
        }
        s.bodies.push_back(body);
    }
    if (d->scene_joint_off) {
        for (int j = d->scene_joint_off[scene]; j < d->scene_joint_off[scene + 1]; ++j) {
            Joint jt;
            jt.type = d->joint_type[j];
            jt.a = d->joint_body[2 * j];
            jt.b = d->joint_body[2 * j + 1];
            const double* jd = d->joint_data + 16 * j;
            if (jt.type == 0) {
                jt.tp = V3(jd[0], jd[1], jd[2]);
                jt.tq = Q(jd[3], jd[4], jd[5], jd[6]);
            } else {
                jt.anchorA = V3(jd[0], jd[1], jd[2]);
                jt.anchorB = V3(jd[3], jd[4], jd[5]);
                jt.axisA = V3(jd[6], jd[7], jd[8]);
                jt.axisB = V3(jd[9], jd[10], jd[11]);
            }
            jt.mu = d->joint_mu[j];
            s.joints.push_back(jt);
        }
    }
    if (d->scene_allowed_off) {
        for (int k = d->scene_allowed_off[scene]; k < d->scene_allowed_off[scene + 1]; ++k) {
            const uint64_t key = pair_key(d->allowed_pairs[2 * k], d->allowed_pairs[2 * k + 1]);
            s.allowed.insert(key);
            if (d->allowed_mu[k] >= 0.0) s.overrides[key] = d->allowed_mu[k];
        }
    }
    return h;
}

void ndo_destroy(void* h) { delete static_cast<Handle*>(h); }

int32_t ndo_configure(void* h, const ndo_params* p) {
    Solver& s = S(h);
    s.dhat = std::max(1e-9, p->dhat);                       // setContactDistanceThreshold
    if (p->kappa > 0.0) s.kappa = p->kappa;                 // setContactStiffness
    s.friction_coeff = p->friction_mu;
    s.enable_friction = p->enable_friction != 0;
    s.enable_ccd = p->enable_ccd != 0;
    s.enable_adaptive_barrier = p->enable_adaptive_barrier != 0;
    s.alm_max_iters = std::max(1, int(p->alm_max_iters));
    s.alm_constraint_tol = std::max(1e-12, p->alm_constraint_tol);
    s.alm_dual_tol = std::max(1e-12, p->alm_dual_tol);
    s.alm_contact_tol = std::max(1e-12, p->alm_contact_tol);
    s.alm_hardening_trigger = std::max(1.0, p->alm_hardening_trigger);
    s.alm_hardening_ratio = std::max(1.0, p->alm_hardening_ratio);
    s.ccd_safety = p->ccd_safety_factor;
    return 0;
}

int32_t ndo_n_bodies(void* h) { return S(h).n(); }

int32_t ndo_set_state(void* h, const double* p, const double* q, const double* v, const double* w, const double* al,
                      const double* aa) {
    Solver& s = S(h);
    for (int i = 0; i < s.n(); ++i) {
        Body& b = s.bodies[i];
        if (p) b.p = V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        if (q) b.q = Q(q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]);
        if (v) b.v = V3(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        if (w) b.w = V3(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
        if (al) b.al = V3(al[3 * i], al[3 * i + 1], al[3 * i + 2]);
        if (aa) b.aa = V3(aa[3 * i], aa[3 * i + 1], aa[3 * i + 2]);
    }
    s.assembler_build();
    return 0;
}

int32_t ndo_get_state(void* h, double* p, double* q, double* v, double* w, double* al, double* aa) {
    Solver& s = S(h);
    for (int i = 0; i < s.n(); ++i) {
        const Body& b = s.bodies[i];
        if (p) { p[3 * i] = b.p.x; p[3 * i + 1] = b.p.y; p[3 * i + 2] = b.p.z; }
        if (q) { q[4 * i] = b.q.w; q[4 * i + 1] = b.q.x; q[4 * i + 2] = b.q.y; q[4 * i + 3] = b.q.z; }
        if (v) { v[3 * i] = b.v.x; v[3 * i + 1] = b.v.y; v[3 * i + 2] = b.v.z; }
        if (w) { w[3 * i] = b.w.x; w[3 * i + 1] = b.w.y; w[3 * i + 2] = b.w.z; }
        if (al) { al[3 * i] = b.al.x; al[3 * i + 1] = b.al.y; al[3 * i + 2] = b.al.z; }
        if (aa) { aa[3 * i] = b.aa.x; aa[3 * i + 1] = b.aa.y; aa[3 * i + 2] = b.aa.z; }
    }
    return 0;
}

int32_t ndo_get_barrier(void* h, double* kappa, double* dhat, int32_t* init) {
    Solver& s = S(h);
    *kappa = s.kappa; *dhat = s.dhat; *init = s.barrier_initialized ? 1 : 0;
    return 0;
}
int32_t ndo_set_barrier(void* h, double kappa, double dhat, int32_t init) {
    Solver& s = S(h);
    s.kappa = kappa; s.dhat = dhat; s.barrier_initialized = init != 0;
    return 0;
}
int32_t ndo_get_joint_state(void* h, double* lambda6, double* mu) {
    Solver& s = S(h);
    for (size_t j = 0; j < s.joints.size(); ++j) {
        for (int k = 0; k < 6; ++k) lambda6[6 * j + k] = s.joints[j].lambda[k];
        mu[j] = s.joints[j].mu;
    }
    return 0;
}

int32_t ndo_set_sampling(void* h, int32_t stride) {
    S(h).sample_stride = std::max(1, int(stride));
    return 0;
}

int32_t ndo_timing(void* h, double* out2) {
    out2[0] = S(h).t_step;
    out2[1] = S(h).t_scene_min;
    return 0;
}

int32_t ndo_step(void* h, double dt) {
    const auto t0 = std::chrono::steady_clock::now();
    S(h).step(dt);
    S(h).t_step += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    S(h).assembler_build();
    return 0;
}

int32_t ndo_set_newton(void* h, double tol, int32_t max_iters) {
    Solver& s = S(h);
    s.newton_tol = tol;
    s.newton_max_iters = std::max(1, int(max_iters));
    return 0;
}

int32_t ndo_set_mass(void* h, const double* mass) {
    Solver& s = S(h);
    for (int i = 0; i < s.n(); ++i) s.bodies[i].mass = mass[i];
    return 0;
}

int32_t ndo_adjoint(void* h, double dt, int32_t n_steps, const double* dl_dp, const double* dl_dq,
                    const double* dl_dv, const double* dl_dw, double* g_p0, double* g_q0, double* g_v0,
                    double* g_w0, double* g_mass, double* g_mu) {
    Solver& s = S(h);
    const int N = s.n();
    std::vector<Solver::AdjStep> rec;
    s.adj_rec = &rec;
    for (int k = 0; k < n_steps; ++k) ndo_step(h, dt);
    s.adj_rec = nullptr;
    // the sweep overwrites the scene's forms; keep the end state to restore it
    const std::vector<Body> end_bodies = s.bodies;
    const std::vector<Joint> end_joints = s.joints;
    std::vector<double> P(3 * N, 0.0), Qb(4 * N, 0.0), V(3 * N, 0.0), W(3 * N, 0.0), gm;
    for (int i = 0; i < 3 * N; ++i) {
        if (dl_dp) P[i] = dl_dp[i];
        if (dl_dv) V[i] = dl_dv[i];
        if (dl_dw) W[i] = dl_dw[i];
    }
    for (int i = 0; i < 4 * N; ++i)
        if (dl_dq) Qb[i] = dl_dq[i];
    double gmu = 0.0;
    const bool ok = s.adjoint_sweep(rec, P, Qb, V, W, gm, gmu);
    s.bodies = end_bodies;
    s.joints = end_joints;
    s.has_barrier = s.has_friction = false;
    s.assembler_build();
    if (!ok) return -1;
    for (int i = 0; i < 3 * N; ++i) {
        if (g_p0) g_p0[i] = P[i];
        if (g_v0) g_v0[i] = V[i];
        if (g_w0) g_w0[i] = W[i];
    }
    for (int i = 0; i < 4 * N; ++i)
        if (g_q0) g_q0[i] = Qb[i];
    for (int i = 0; i < N; ++i)
        if (g_mass) g_mass[i] = gm[i];
    if (g_mu) *g_mu = gmu;
    return 0;
}

int32_t ndo_last_stats(void* h, double* o) {
    Solver& s = S(h);
    o[0] = s.st_alm; o[1] = s.st_newton; o[2] = s.st_contacts; o[3] = s.st_fv; o[4] = s.st_ee;
    o[5] = s.st_evals; o[6] = s.st_kappa; o[7] = s.st_toi;
    return 0;
}

int32_t ndo_detect_body_pairs(void* h, double r, int32_t* out, int32_t cap, int32_t* n) {
    Solver& s = S(h);
    s.assembler_build();
    auto pairs = s.detect_body_pairs(r);
    *n = int32_t(pairs.size());
    for (size_t i = 0; i < pairs.size() && int32_t(i) < cap; ++i) { out[2 * i] = pairs[i].first; out[2 * i + 1] = pairs[i].second; }
    return 0;
}

int32_t ndo_min_body_pair_distance(void* h, double r, double* out) {
    Solver& s = S(h);
    s.assembler_build();
    *out = s.min_body_pair_distance(r);
    return 0;
}

int32_t ndo_build_candidates(void* h, double dhat, int32_t* counts, int32_t* vv, int32_t* ev, int32_t* ee, int32_t* fv,
                             int32_t cap) {
    Solver& s = S(h);
    s.assembler_build();
    s.build_candidates(dhat, s.cands);
    counts[0] = int32_t(s.cands.vv.size());
    counts[1] = int32_t(s.cands.ev.size());
    counts[2] = int32_t(s.cands.ee.size());
    counts[3] = int32_t(s.cands.fv.size());
    for (size_t i = 0; i < s.cands.vv.size() && vv && int32_t(i) < cap; ++i) for (int k = 0; k < 4; ++k) vv[4 * i + k] = s.cands.vv[i][k];
    for (size_t i = 0; i < s.cands.ev.size() && ev && int32_t(i) < cap; ++i) for (int k = 0; k < 5; ++k) ev[5 * i + k] = s.cands.ev[i][k];
    for (size_t i = 0; i < s.cands.ee.size() && ee && int32_t(i) < cap; ++i) for (int k = 0; k < 6; ++k) ee[6 * i + k] = s.cands.ee[i][k];
    for (size_t i = 0; i < s.cands.fv.size() && fv && int32_t(i) < cap; ++i) for (int k = 0; k < 4; ++k) fv[4 * i + k] = s.cands.fv[i][k];
    return 0;
}

int32_t ndo_select_top_k(const double* scores, int32_t n, int32_t k, int32_t* out) {
    std::vector<double> s(scores, scores + n);
    auto r = select_top_k(s, k);
    for (size_t i = 0; i < r.size(); ++i) out[i] = r[i];
    return int32_t(r.size());
}

int32_t ndo_narrow_phase(void* h, const double* x, int32_t cap, int32_t* n, int32_t* ints, double* dbl) {
    Solver& s = S(h);
    std::vector<double> xv;
    if (x) xv.assign(x, x + size_t(s.n()) * 6);
    auto cs = s.narrow_phase(x ? &xv : nullptr);
    contacts_out(cs, cap, n, ints, dbl);
    return 0;
}

int32_t ndo_scene_min_distance(void* h, double* out) {
    *out = S(h).scene_min_distance();
    return 0;
}

int32_t ndo_barrier_eval(void* h, int32_t nc, const int32_t* ints, const double* dbl, const double* x, double dhat,
                         double kappa, double* value, double* grad, double* hess, double* nf) {
    Solver& s = S(h);
    const auto cs = contacts_in(nc, ints, dbl);
    std::vector<double> xv(x, x + size_t(s.n()) * 6);
    const double dh = std::max(1e-9, dhat), k = std::max(0.0, kappa);
    if (value) *value = Solver::bar_value(cs, xv, dh, k);
    if (grad) { std::vector<double> g; Solver::bar_grad(cs, xv, dh, k, g); std::copy(g.begin(), g.end(), grad); }
    if (hess) { SymAccum H; H.reset(int(xv.size())); Solver::bar_hess(cs, xv, dh, k, 1.0, H); dense_out(H, hess); }
    if (nf) { auto v = Solver::bar_normal_forces(cs, xv, dh, k); std::copy(v.begin(), v.end(), nf); }
    return 0;
}

int32_t ndo_friction_eval(void* h, int32_t nc, const int32_t* ints, const double* dbl, const double* nfp, const double* x,
                          double mu, double eps, double dt, double* value, double* grad, double* hess) {
    Solver& s = S(h);
    const auto cs = contacts_in(nc, ints, dbl);
    std::vector<double> nf(nfp, nfp + nc);
    std::vector<double> xv(x, x + size_t(s.n()) * 6);
    if (value) *value = s.fric_value(cs, nf, xv, mu, eps, dt);
    if (grad) { std::vector<double> g; s.fric_grad(cs, nf, xv, mu, eps, dt, g); std::copy(g.begin(), g.end(), grad); }
    if (hess) { SymAccum H; H.reset(int(xv.size())); s.fric_hess(cs, nf, xv, mu, eps, dt, 1.0, H); dense_out(H, hess); }
    return 0;
}

int32_t ndo_friction_potential(const double* u, double N, double mu, double eps, double* value, double* g3, double* h9) {
    const V3 uu(u[0], u[1], u[2]);
    if (value) *value = friction_value(uu, N, mu, eps);
    if (g3) { const V3 g = friction_grad(uu, N, mu, eps); g3[0] = g.x; g3[1] = g.y; g3[2] = g.z; }
    if (h9) { const M3 H = friction_hess(uu, N, mu, eps); for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) h9[i * 3 + j] = H(i, j); }
    return 0;
}

int32_t ndo_barrier_potential(double d, double dhat, double kappa, double* o) {
    o[0] = barrier_value(d, dhat, kappa);
    o[1] = barrier_gradient(d, dhat, kappa);
    o[2] = barrier_hessian(d, dhat, kappa);
    return 0;
}

int32_t ndo_ccd_toi(int32_t kind, const double* pa, const double* pb, const double* a_pts, const double* b_pts,
                    double earliest, double tol, double* toi, int32_t* pops) {
    const Pose A = make_pose(pa), B = make_pose(pb);
    V3 a[2], b[3];
    if (kind < 0 || kind > 4) return -3;
    const int na = (kind == 1 || kind == 4) ? 2 : 1, nb = (kind == 2 || kind == 3) ? 3 : 2;
    for (int i = 0; i < na; ++i) a[i] = V3(a_pts[3 * i], a_pts[3 * i + 1], a_pts[3 * i + 2]);
    for (int i = 0; i < nb; ++i) b[i] = V3(b_pts[3 * i], b_pts[3 * i + 1], b_pts[3 * i + 2]);
    int p = 0;
    const int r = ccd_toi(kind, A, B, a, b, earliest, tol, toi, &p);
    if (pops) *pops = p;
    return r;
}

int32_t ndo_ccd_earliest_toi(void* h, double dt, double* out) {
    *out = S(h).ccd_earliest_toi(dt);
    return 0;
}

int32_t ndo_root_finder_kat(int32_t which, double* lo, double* hi) {
    // tests/physics/test_ccd.cpp:79-125: f(x) = x - 0.5; x*x - 0.25; x*x + 1 on [0, 1], tol 1e-8.
    Box x0, out;
    x0.d[0] = I(0, 1);
    const double tol[1] = {1e-8};
    auto f = [&](const Box& b) {
        const I x = b.d[0];
        I y;
        if (which == 0) y = x - I(0.5);
        else if (which == 1) y = x * x - I(0.25);
        else y = x * x + I(1.0);
        return I3{y, I(0.0), I(0.0)};
    };
    int pops = 0;
    const bool found = root_finder(1, f, [](const Box&) { return true; }, x0, tol, out, &pops);
    if (found) { *lo = out.d[0].lo; *hi = out.d[0].hi; }
    return found ? 1 : 0;
}

void ndo_round_ops(const double* a, const double* b, int32_t n, double* out6) {
    for (int i = 0; i < n; ++i) {
        out6[6 * i + 0] = add_dn(a[i], b[i]);
        out6[6 * i + 1] = add_up(a[i], b[i]);
        out6[6 * i + 2] = sub_dn(a[i], b[i]);
        out6[6 * i + 3] = sub_up(a[i], b[i]);
        out6[6 * i + 4] = mul_dn(a[i], b[i]);
        out6[6 * i + 5] = mul_up(a[i], b[i]);
    }
}

// Elementary functions of the path, element-wise: fn 0 log(x), 1 sin(x), 2 cos(x),
// 3 atan2(x, y), 4 acos(x).
int32_t ndo_elemfn(int32_t fn, const double* x, const double* y, int32_t n, double* out) {
    for (int i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = nd::ef::log(x[i]); break;
            case 1: out[i] = nd::ef::sin(x[i]); break;
            case 2: out[i] = nd::ef::cos(x[i]); break;
            case 3: out[i] = nd::ef::atan2(x[i], y[i]); break;
            case 4: out[i] = nd::ef::acos(x[i]); break;
            default: return -1;
        }
    }
    return 0;
}

// ---- function-level pins of the narrow phase's primitives (IPCSolver.cpp:32-134)
int32_t ndo_closest_points_segments(const double* s12, double* out6) {
    V3 c1, c2;
    closest_points_on_segments(V3(s12[0], s12[1], s12[2]), V3(s12[3], s12[4], s12[5]), V3(s12[6], s12[7], s12[8]),
                               V3(s12[9], s12[10], s12[11]), c1, c2);
    out6[0] = c1.x; out6[1] = c1.y; out6[2] = c1.z;
    out6[3] = c2.x; out6[4] = c2.y; out6[5] = c2.z;
    return 0;
}

int32_t ndo_point_in_triangle(const double* p, const double* tri9) {
    return point_in_triangle(V3(p[0], p[1], p[2]), V3(tri9[0], tri9[1], tri9[2]), V3(tri9[3], tri9[4], tri9[5]),
                             V3(tri9[6], tri9[7], tri9[8])) ? 1 : 0;
}

int32_t ndo_safe_normalized(const double* v, const double* fb, double* out3) {
    const V3 r = safe_normalized(V3(v[0], v[1], v[2]), V3(fb[0], fb[1], fb[2]));
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
    return 0;
}

// ---- BVH over given boxes (BVH.cpp), broad phase with inflation (BroadPhase.cpp)
static vector<AABB> boxes_in(const double* b6, int32_t n) {
    vector<AABB> v;
    for (int i = 0; i < n; ++i)
        v.emplace_back(V3(b6[6 * i], b6[6 * i + 1], b6[6 * i + 2]), V3(b6[6 * i + 3], b6[6 * i + 4], b6[6 * i + 5]));
    return v;
}

int32_t ndo_bvh_query(const double* boxes6, int32_t n, const double* q6, int32_t* out, int32_t cap) {
    BVH bvh;
    bvh.init(boxes_in(boxes6, n));
    vector<unsigned> hits;
    bvh.query(AABB(V3(q6[0], q6[1], q6[2]), V3(q6[3], q6[4], q6[5])), hits);
    for (int i = 0; i < int(hits.size()) && i < cap; ++i) out[i] = int32_t(hits[i]);
    return int32_t(hits.size());
}

int32_t ndo_broad_phase_boxes(const double* boxes6, int32_t n, double inflation, int32_t* out_pairs, int32_t cap) {
    auto b = boxes_in(boxes6, n);
    for (auto& x : b) x = x.inflated(inflation);
    const auto pr = broad_phase_detect(b);
    for (int i = 0; i < int(pr.size()) && i < cap; ++i) {
        out_pairs[2 * i] = pr[i].first;
        out_pairs[2 * i + 1] = pr[i].second;
    }
    return int32_t(pr.size());
}

// ---- interval operations of the CCD (Interval.h): op 0 a+b, 1 a-b, 2 a*b, 3 width(a),
// 4 median(a), 5 zero_in(a)
int32_t ndo_interval_op(int32_t op, const double* a2, const double* b2, double* out2) {
    const ndo::I a(a2[0], a2[1]), b(b2[0], b2[1]);
    ndo::I r;
    switch (op) {
        case 0: r = a + b; break;
        case 1: r = a - b; break;
        case 2: r = a * b; break;
        case 3: out2[0] = out2[1] = ndo::width(a); return 0;
        case 4: out2[0] = out2[1] = ndo::median(a); return 0;
        case 5: out2[0] = out2[1] = ndo::zero_in(a) ? 1.0 : 0.0; return 0;
        default: return -1;
    }
    out2[0] = r.lo;
    out2[1] = r.hi;
    return 0;
}

// ---- AdaptiveBarrier: initialize(avg_mass, diag), then updateStiffness(min_dist[k]);
// kappas[0] after initialize, kappas[k + 1] after update k.
int32_t ndo_adaptive_barrier(double avg_mass, double diag, double max_stiffness, int32_t n, const double* min_dist,
                             double* kappas, double* kappa_min) {
    AdaptiveBarrierO ab;
    if (max_stiffness > 0) ab.max_stiffness = max_stiffness;
    ab.initialize(avg_mass, diag);
    kappas[0] = ab.kappa;
    for (int k = 0; k < n; ++k) {
        ab.update(min_dist[k]);
        kappas[k + 1] = ab.kappa;
    }
    *kappa_min = ab.kappa_min;
    return 0;
}

int32_t ndo_set_narrow_raw(void* h, int32_t on) {
    S(h).narrow_raw = on != 0;
    return 0;
}

int32_t ndo_set_scene_min_pruned(void* h, int32_t on) {
    S(h).prune_scene_min = on != 0;
    return 0;
}

// ---- recorded steps: every Newton solve (and its iterations) of n_steps steps
int32_t ndo_record_steps(void* h, double dt, int32_t n_steps) {
    Handle& H = *static_cast<Handle*>(h);
    H.rec.clear();
    H.s.adj_rec = &H.rec;
    H.s.record_newton = true;
    for (int k = 0; k < n_steps; ++k) H.s.step(dt);
    H.s.adj_rec = nullptr;
    H.s.record_newton = false;
    return 0;
}

// counts: [steps]; with step >= 0: [solves of the step]; with solve >= 0 also
// [converged, newton iterations reported, recorded iterations, trials of iteration it]
int32_t ndo_record_info(void* h, int32_t step, int32_t solve, int32_t it, int32_t* out4) {
    Handle& H = *static_cast<Handle*>(h);
    out4[0] = int32_t(H.rec.size());
    if (step < 0) return 0;
    if (step >= int(H.rec.size())) return -1;
    const auto& st = H.rec[step];
    out4[0] = int32_t(st.solves.size());
    if (solve < 0) return 0;
    if (solve >= int(st.solves.size())) return -1;
    const auto& so = st.solves[solve];
    out4[0] = so.converged ? 1 : 0;
    out4[1] = so.newton_iters;
    out4[2] = int32_t(so.iters.size());
    out4[3] = (it >= 0 && it < int(so.iters.size())) ? int32_t(so.iters[it].trials.size()) : 0;
    return 0;
}

// x_hat of the step and x* of the solve (6n each); per iteration it (>= 0): x and dx (6n
// each, into x_hat / x_star), scal[4] = f(x), g.dx, contact value, cg.dx and trials[3k]
// = alpha, f, contact value of every trial. Joint state before the solve: lambda6 (6J)
// and mu (J); after: mu_post (J).
int32_t ndo_record_solve(void* h, int32_t step, int32_t solve, int32_t it, double* x_hat, double* x_star,
                         double* scal, double* trials, double* lambda6, double* mu, double* mu_post) {
    Handle& H = *static_cast<Handle*>(h);
    if (step < 0 || step >= int(H.rec.size())) return -1;
    const auto& st = H.rec[step];
    if (solve < 0 || solve >= int(st.solves.size())) return -1;
    const auto& so = st.solves[solve];
    if (it < 0) {
        if (x_hat) std::copy(st.x_hat.begin(), st.x_hat.end(), x_hat);
        if (x_star) std::copy(so.x_star.begin(), so.x_star.end(), x_star);
    } else {
        if (it >= int(so.iters.size())) return -1;
        const auto& nr = so.iters[it];
        if (x_hat) std::copy(nr.x.begin(), nr.x.end(), x_hat);
        if (x_star) std::copy(nr.dx.begin(), nr.dx.end(), x_star);
        if (scal) { scal[0] = nr.fx; scal[1] = nr.descent; scal[2] = nr.fc; scal[3] = nr.cdesc; }
        if (trials)
            for (size_t k = 0; k < nr.trials.size(); ++k) {
                trials[3 * k] = nr.trials[k].alpha;
                trials[3 * k + 1] = nr.trials[k].fn;
                trials[3 * k + 2] = nr.trials[k].fcn;
            }
    }
    for (size_t j = 0; j < so.joints.size(); ++j) {
        if (lambda6) for (int k = 0; k < 6; ++k) lambda6[6 * j + k] = so.joints[j].lambda[k];
        if (mu) mu[j] = so.joints[j].mu;
        if (mu_post) mu_post[j] = so.mu_post[j];
    }
    return 0;
}

// The objective of a recorded solve at x (value, gradient, dense Hessian, contact value
// and contact gradient; any may be NULL) and the joints' constraints C (6J).
int32_t ndo_record_objective(void* h, int32_t step, int32_t solve, const double* x, double* value, double* grad,
                             double* hess, double* cvalue, double* cgrad, double* jointC) {
    Handle& H = *static_cast<Handle*>(h);
    if (step < 0 || step >= int(H.rec.size())) return -1;
    if (solve < 0 || solve >= int(H.rec[step].solves.size())) return -1;
    H.probe = H.s;
    H.probe.adj_rec = nullptr;
    H.probe.adj_load(H.rec[step], H.rec[step].solves[solve]);
    const Solver& P = H.probe;
    std::vector<double> xv(x, x + size_t(P.n()) * 6);
    if (value) *value = P.obj_value(xv);
    if (grad) { std::vector<double> g; P.obj_grad(xv, g); std::copy(g.begin(), g.end(), grad); }
    if (hess) { SymAccum A; P.obj_hess(xv, A); dense_out(A, hess); }
    if (cvalue) *cvalue = P.contact_value(xv);
    if (cgrad) { std::vector<double> g; P.contact_grad(xv, g); std::copy(g.begin(), g.end(), cgrad); }
    if (jointC)
        for (size_t j = 0; j < P.joints.size(); ++j) P.joint_C(P.joints[j], xv, jointC + 6 * j);
    return 0;
}

// ---- ipc mode (function level)
int32_t ndo_ipc_pair(int32_t kind, const double* x12, double dhat, double kappa, int32_t psd, double* E, double* g12,
                     double* H144, int32_t* type, double* d) {
    int t = 0;
    double dd = 0.0;
    nd::ipc::pair_energy(kind, x12, dhat, kappa, E, g12, H144, &t, &dd);
    if (psd) nd::ipc::psd_project<12>(H144);
    if (type) *type = t;
    if (d) *d = dd;
    return 0;
}

int32_t ndo_ipc_distance(int32_t kind, const double* x12, int32_t* type, double* d) {
    int t = 0;
    *d = nd::ipc::pair_distance(kind, x12, &t);
    if (type) *type = t;
    return 0;
}

double ndo_psd_project12(double* H144) { return nd::ipc::psd_project<12>(H144); }

int32_t ndo_vertex_jacobian(const double* xb6, const double* q4, const double* r3, double* xw3, double* J9) {
    nd::ipc::vertex_and_jacobian(xb6, q4, r3, xw3, J9);
    return 0;
}

int32_t ndo_ipc_max_step(void* h, int32_t np, const int32_t* pairs5, const double* x, const double* dx, double sep,
                         double safety, double* alpha, double* toi, int32_t* fail_index) {
    Solver& s = S(h);
    const size_t nd = size_t(s.n()) * 6;
    std::vector<double> xv(x, x + nd), dv(dx, dx + nd);
    int fi = -1;
    const int r = s.ipc_max_step(pairs5, np, xv, dv, sep, safety, *toi, *alpha, fi);
    if (fail_index) *fail_index = fi;
    return -r;
}

int32_t ndo_pcg_ordered(int32_t nb, const double* diag, int32_t noff, const int32_t* offij, const double* offb,
                        const double* b, double tol, int32_t max_it, double* x) {
    std::vector<double> dg(diag, diag + 36 * size_t(nb)), ob(offb, offb + 36 * size_t(noff)), bb(b, b + 6 * size_t(nb)), xx;
    std::vector<int> oi(offij, offij + 2 * size_t(noff));
    const int it = pcg_ordered(nb, dg, oi, ob, bb, tol, max_it, xx);
    std::copy(xx.begin(), xx.end(), x);
    return it;
}

int32_t ndo_ipc_step(void* h, double dt, const double* prm7, const int32_t* prm3, int32_t* stats4) {
    Solver::IpcStepParams P{prm7[0], prm7[1], prm7[2], prm7[3], prm7[4], prm7[5], prm7[6], prm3[0], prm3[1], prm3[2]};
    return S(h).ipc_step(dt, P, stats4);
}

int32_t ndo_ipc_barrier_eval(void* h, int32_t np, const int32_t* pairs5, const double* x, double dhat, double kappa,
                             int32_t psd, double* value, int32_t cap, int32_t* n_seg, int32_t* seg_bodies,
                             double* seg_g, double* seg_H) {
    Solver& s = S(h);
    std::vector<double> xv(x, x + size_t(s.n()) * 6);
    std::vector<Solver::IpcSeg> segs;
    if (!s.ipc_barrier_eval(pairs5, np, xv, dhat, kappa, psd != 0, *value, segs)) return -1;
    *n_seg = int32_t(segs.size());
    for (size_t k = 0; k < segs.size() && int32_t(k) < cap; ++k) {
        seg_bodies[2 * k] = segs[k].lo;
        seg_bodies[2 * k + 1] = segs[k].hi;
        std::copy(segs[k].g, segs[k].g + 12, seg_g + 12 * k);
        std::copy(segs[k].H, segs[k].H + 144, seg_H + 144 * k);
    }
    return 0;
}

int32_t ndo_objective(void* h, const double* x, double* value, double* grad, double* hess) {
    Solver& s = S(h);
    std::vector<double> xv(x, x + size_t(s.n()) * 6);
    if (value) *value = s.obj_value(xv);
    if (grad) { std::vector<double> g; s.obj_grad(xv, g); std::copy(g.begin(), g.end(), grad); }
    if (hess) { SymAccum H; s.obj_hess(xv, H); dense_out(H, hess); }
    return 0;
}

}  // extern "C"
