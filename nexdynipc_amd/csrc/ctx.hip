// ndipc C ABI: context, scene upload, step orchestration and the function-level entry
// points. The host side only stages data and enqueues kernels on the context stream;
// a step (IPCSolver::step) is a fixed kernel sequence whose control decisions
// (Newton convergence, line search, ALM break) are made on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "../../include/ndipc.h"
#include "ndipc_internal.h"

using nd::Dev;
using nd::SceneParams;
using nd::Work;

struct TimingRec {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    hipEvent_t open = nullptr;
};

// device buffer kept across calls, grown on demand (the ipc-mode step's work space)
struct GrowBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct ndipc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream for the Newton line-search helpers, forked from and joined to
    // `stream` around each Newton launch
    hipStream_t hstream = nullptr;
    hipEvent_t hfork = nullptr, hjoin = nullptr;
    // third stream for the step prefix's stiffness control (scene-min flag, kappa update),
    // forked after the broad phase and joined before the contact setup
    hipStream_t pstream = nullptr;
    hipEvent_t pfork = nullptr, pjoin = nullptr;
    std::string err;
    bool uploaded = false;
    Dev d{};
    Work w{};
    int S = 0, N = 0, J = 0;
    std::vector<int> body_off, nv, nf, ne, e_off, v_off, f_off, body_id, joint_off;
    std::vector<int> edges_host;
    long long pair_cap = 0;
    std::vector<SceneParams> sp;
    std::vector<void*> allocs;
    bool alloc_failed = false;  // a dalloc / dupload of the current upload failed
    double* d_dt = nullptr;
    double* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    bool timing = false;
    std::map<std::string, TimingRec> tim;
    double* mass_dev = nullptr;  // writable view of d.mass
    // adjoint: checkpoints every `adj_every` steps of the last recorded forward run
    nd::Adj adj{};
    std::vector<void*> adj_allocs;
    int adj_every = 0, adj_steps = 0, adj_nck = 0;
    std::vector<double> adj_dt;
    double* ck_state = nullptr;  // [nck][19N] p q v w al aa
    double* ck_joint = nullptr;  // [nck][7J] lambda(6J) mu(J)
    SceneParams* ck_sp = nullptr;  // [nck][S]
    // one whole step captured as a HIP graph (kernel nodes only), replayed per step
    bool use_graph = true;
    hipGraphExec_t step_exec = nullptr;
    int graph_alm = -1;  // ALM iterations the captured step enqueues
    // adaptive ALM rounds (default): the step's prefix and one ALM round captured as
    // graphs; the host issues rounds until every scene's ALM loop has ended
    bool adaptive = true;
    int round_lookahead = 1;          // rounds issued beyond the last one reported
    hipGraphExec_t pre_exec = nullptr, round_exec = nullptr;
    // asynchronous stepping of a batch (adaptive mode, several scenes, several steps): each
    // scene starts its next step as soon as its ALM loop ends; an iteration runs the
    // write-back of the scenes that ended, the prefix for the scenes starting a step and
    // one ALM round for every scene inside one
    bool async_steps = true;
    unsigned int* hflag = nullptr;    // pinned host memory: [round number, all scenes done]
    unsigned int rounds_issued = 0;   // rounds with a flag issued since upload
    GrowBuf ipc_buf[8];               // ndipc_ipc_step's work space
};

namespace {
template <typename T>
hipError_t grow(GrowBuf& b, size_t bytes, T** out) {
    if (bytes > b.bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        const hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) return e;
        b.bytes = bytes;
    }
    *out = static_cast<T*>(b.p);
    return hipSuccess;
}
}  // namespace

namespace {
ndipc_ctx* g_timing_ctx = nullptr;  // context whose stream is being timed (one at a time)

int fail(ndipc_ctx* c, const std::string& m) {
    if (c) c->err = m;
    return -1;
}
#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(ctx, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Work-groups of the top-K and scene-min kernels when a mesh exceeds the LDS-resident
// size (one HBM slab each; 1024 single-wave top-K groups, 256 scene-min groups of 256).
constexpr int kTopkSlabBlocks = 1024;
constexpr int kSmfSlabBlocks = 256;

// Device scratch of one call: freed when the call returns, on every path.
struct ScratchBufs {
    std::vector<void*> p;
    ScratchBufs() = default;
    ScratchBufs(const ScratchBufs&) = delete;
    ScratchBufs& operator=(const ScratchBufs&) = delete;
    ~ScratchBufs() {
        for (void* q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T** out, size_t bytes) {
        void* q = nullptr;
        const hipError_t e = hipMalloc(&q, bytes);
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T*>(q);
        return e;
    }
};

template <typename T>
T* dalloc(ndipc_ctx* ctx, size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
        ctx->alloc_failed = true;
        return nullptr;
    }
    ctx->allocs.push_back(p);
    if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) {
        ctx->alloc_failed = true;
        return nullptr;
    }
    return static_cast<T*>(p);
}
template <typename T>
T* dupload(ndipc_ctx* ctx, const T* h, size_t n) {
    T* p = dalloc<T>(ctx, n);
    if (p && h && n && hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        ctx->alloc_failed = true;
        return nullptr;
    }
    return p;
}

void free_adjoint(ndipc_ctx* ctx) {
    for (void* p : ctx->adj_allocs) (void)hipFree(p);
    ctx->adj_allocs.clear();
    ctx->adj = nd::Adj{};
    ctx->adj_every = ctx->adj_steps = ctx->adj_nck = 0;
    ctx->ck_state = ctx->ck_joint = nullptr;
    ctx->ck_sp = nullptr;
}

void drop_graph(ndipc_ctx* ctx) {
    if (ctx->step_exec) (void)hipGraphExecDestroy(ctx->step_exec);
    if (ctx->pre_exec) (void)hipGraphExecDestroy(ctx->pre_exec);
    if (ctx->round_exec) (void)hipGraphExecDestroy(ctx->round_exec);
    ctx->step_exec = ctx->pre_exec = ctx->round_exec = nullptr;
    ctx->graph_alm = -1;
}

void free_all(ndipc_ctx* ctx) {
    drop_graph(ctx);
    free_adjoint(ctx);
    for (void* p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
    ctx->alloc_failed = false;
    ctx->uploaded = false;
}

int sync_sp_to_dev(ndipc_ctx* ctx) {
    HIPCHK(hipMemcpyAsync(ctx->d.sp, ctx->sp.data(), sizeof(SceneParams) * ctx->S, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}
int sync_sp_from_dev(ndipc_ctx* ctx) {
    HIPCHK(hipMemcpyAsync(ctx->sp.data(), ctx->d.sp, sizeof(SceneParams) * ctx->S, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ContactAssembler.cpp:90-120 — unique edges in face order, stored (min, max).
std::vector<int> unique_edges(const int* faces, int nf) {
    std::unordered_set<uint64_t> seen;
    std::vector<int> E;
    for (int f = 0; f < nf; ++f) {
        const int v[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
        for (int k = 0; k < 3; ++k) {
            const int a = v[k], b = v[(k + 1) % 3];
            const int lo = std::min(a, b), hi = std::max(a, b);
            const uint64_t key = (uint64_t(uint32_t(lo)) << 32) | uint32_t(hi);
            if (seen.insert(key).second) {
                E.push_back(lo);
                E.push_back(hi);
            }
        }
    }
    return E;
}

void default_params(SceneParams& p) {
    p.dhat = 1e-3;
    p.kappa = 0.0;
    p.mu = 0.0;
    p.ccd_safety = 0.8;
    p.alm_ctol = 1e-4;
    p.alm_dtol = 1e-3;
    p.alm_ctcol = 1e-3;
    p.alm_trig = 2.0;
    p.alm_ratio = 2.0;
    p.enable_friction = 0;
    p.enable_ccd = 1;
    p.enable_adaptive = 1;
    p.alm_max_iters = 10;
    p.barrier_initialized = 0;
    p.newton_tol = 1e-6;
    p.newton_max_iters = 20;
}

int check_ctx(ndipc_ctx* ctx) {
    if (!ctx) return -1;
    if (!ctx->uploaded) return fail(ctx, "no scenes uploaded");
    HIPCHK(hipSetDevice(ctx->device));
    return 0;
}

int max_grid_pairs(ndipc_ctx* ctx) {
    return int(std::min<long long>(ctx->pair_cap, 1 << 30));
}
}  // namespace

// Kernel timing with HIP events on the launching stream. Events come from a pool that
// is created once per context, so recording adds no allocation to the timed region.
namespace {
std::vector<hipEvent_t> g_event_pool;
size_t g_event_next = 0;
hipEvent_t pooled_event() {
    if (g_event_next == g_event_pool.size()) {
        for (int i = 0; i < 4096; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) break;
            g_event_pool.push_back(e);
        }
    }
    return g_event_next < g_event_pool.size() ? g_event_pool[g_event_next++] : nullptr;
}
}  // namespace

namespace nd {
void timing_begin(const char* name, hipStream_t st) {
    ndipc_ctx* c = g_timing_ctx;
    if (!c || !c->timing || (st != c->stream && st != c->pstream)) return;
    hipEvent_t e = pooled_event();
    if (!e) return;
    (void)hipEventRecord(e, st);
    c->tim[name].open = e;
}
void timing_end(const char* name, hipStream_t st) {
    ndipc_ctx* c = g_timing_ctx;
    if (!c || !c->timing || (st != c->stream && st != c->pstream)) return;
    TimingRec& r = c->tim[name];
    if (!r.open) return;
    hipEvent_t e = pooled_event();
    if (!e) return;
    (void)hipEventRecord(e, st);
    r.ev.emplace_back(r.open, e);
    r.open = nullptr;
}
}  // namespace nd

// Work vectors of ndipc_pcg_solve_noalloc: one fixed array per GPU, held until the
// process ends (per block r, z, p, A p: 24 doubles; the inverted diagonal block: 36).
constexpr int kNoAllocBlockDoubles = 60;
static __device__ __align__(256) double scratch[size_t(NDIPC_PCG_NOALLOC_MAX_BLOCKS) * kNoAllocBlockDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each GPU

// Once per device, with the device current: the device address of `scratch`. Host code
// uses only this address (the array's name is a host-side stand-in).
int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return 0;
}
}  // namespace

extern "C" {

int32_t ndipc_create(int32_t device, ndipc_ctx** out) {
    if (!out) return -1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device) return -2;
    if (hipSetDevice(device) != hipSuccess) return -3;
    nd::geom_configure();
    if (nd::geom_configure_smf() != 0) return -5;
    if (scratch_setup(device) != 0) return -6;
    auto* c = new ndipc_ctx();
    c->device = device;
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return -4;
    }
    auto release = [&] {
        if (c->pjoin) (void)hipEventDestroy(c->pjoin);
        if (c->pfork) (void)hipEventDestroy(c->pfork);
        if (c->pstream) (void)hipStreamDestroy(c->pstream);
        if (c->hjoin) (void)hipEventDestroy(c->hjoin);
        if (c->hfork) (void)hipEventDestroy(c->hfork);
        if (c->hstream) (void)hipStreamDestroy(c->hstream);
        (void)hipStreamDestroy(c->stream);
        delete c;
    };
    if (hipStreamCreate(&c->hstream) != hipSuccess ||
        hipEventCreateWithFlags(&c->hfork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->hjoin, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreate(&c->pstream) != hipSuccess ||
        hipEventCreateWithFlags(&c->pfork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->pjoin, hipEventDisableTiming) != hipSuccess) {
        release();
        return -4;
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&c->hflag), 2 * sizeof(unsigned int), hipHostMallocCoherent) !=
        hipSuccess) {
        release();
        return -4;
    }
    c->hflag[0] = c->hflag[1] = 0;
    if (const char* env = std::getenv("NDIPC_FIXED_ROUNDS")) c->adaptive = env[0] != '1';
    if (const char* env = std::getenv("NDIPC_ROUND_LOOKAHEAD")) c->round_lookahead = std::max(0, std::atoi(env));
    if (const char* env = std::getenv("NDIPC_ASYNC_STEPS")) c->async_steps = env[0] != '0';
    *out = c;
    return 0;
}

int32_t ndipc_destroy(ndipc_ctx* ctx) {
    if (!ctx) return 0;
    // tear down in any case; -1 reports that the device or the stream had failed
    int rc = 0;
    if (hipSetDevice(ctx->device) != hipSuccess) rc = -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = -1;
    if (g_timing_ctx == ctx) g_timing_ctx = nullptr;
    free_all(ctx);
    for (GrowBuf& b : ctx->ipc_buf)
        if (b.p && hipFree(b.p) != hipSuccess) rc = -1;
    if (hipStreamSynchronize(ctx->hstream) != hipSuccess) rc = -1;
    if (hipStreamSynchronize(ctx->pstream) != hipSuccess) rc = -1;
    if (hipEventDestroy(ctx->hfork) != hipSuccess) rc = -1;
    if (hipEventDestroy(ctx->hjoin) != hipSuccess) rc = -1;
    if (hipStreamDestroy(ctx->hstream) != hipSuccess) rc = -1;
    if (hipEventDestroy(ctx->pfork) != hipSuccess) rc = -1;
    if (hipEventDestroy(ctx->pjoin) != hipSuccess) rc = -1;
    if (hipStreamDestroy(ctx->pstream) != hipSuccess) rc = -1;
    if (hipStreamDestroy(ctx->stream) != hipSuccess) rc = -1;
    if (hipHostFree(ctx->hflag) != hipSuccess) rc = -1;
    delete ctx;
    return rc;
}

const char* ndipc_last_error(const ndipc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* ndipc_stream(ndipc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int32_t ndipc_upload_scenes(ndipc_ctx* ctx, const ndipc_scene_desc* dsc) {
    if (!ctx || !dsc) return -1;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    free_all(ctx);
    ctx->hflag[0] = ctx->hflag[1] = 0;  // the device's round counter restarts with the upload
    ctx->rounds_issued = 0;
    const int S = dsc->n_scenes, N = dsc->n_bodies;
    if (S <= 0 || N < 0) return fail(ctx, "empty scene batch");
    if (dsc->scene_body_off[0] != 0 || dsc->scene_body_off[S] != N) return fail(ctx, "scene_body_off inconsistent");
    ctx->S = S;
    ctx->N = N;
    ctx->body_off.assign(dsc->scene_body_off, dsc->scene_body_off + S + 1);
    ctx->nv.assign(dsc->n_verts, dsc->n_verts + N);
    ctx->nf.assign(dsc->n_faces, dsc->n_faces + N);
    ctx->v_off.assign(dsc->vert_off, dsc->vert_off + N);
    ctx->f_off.assign(dsc->face_off, dsc->face_off + N);
    ctx->body_id.assign(dsc->body_id, dsc->body_id + N);
    int total_v = 0, total_f = 0;
    for (int b = 0; b < N; ++b) {
        total_v = std::max(total_v, ctx->v_off[b] + ctx->nv[b]);
        total_f = std::max(total_f, ctx->f_off[b] + ctx->nf[b]);
    }
    // per-body scene, local AABB, max edge, unique edges (per distinct mesh)
    std::vector<int> body_scene(N);
    std::vector<double> lbox(6 * size_t(N), 0.0), max_edge(N, 0.0), vbound(N, 0.0);
    ctx->e_off.assign(N, 0);
    ctx->ne.assign(N, 0);
    ctx->edges_host.clear();
    std::map<std::pair<int, int>, std::pair<int, int>> edge_cache;
    long long pair_cap = 0;
    for (int s = 0; s < S; ++s) {
        const int b0 = ctx->body_off[s], b1 = ctx->body_off[s + 1];
        if (b1 - b0 > nd::kMaxSceneBodies) return fail(ctx, "scene has more than 4096 bodies");
        long long m = 0;
        for (int b = b0; b < b1; ++b) {
            body_scene[b] = s;
            if (ctx->nv[b] > 0) ++m;
        }
        pair_cap += m * (m - 1) / 2;
    }
    for (int b = 0; b < N; ++b) {
        const int nvb = ctx->nv[b], nfb = ctx->nf[b];
        if (nvb == 0) continue;
        if (nvb > nd::kMaxPrims || nfb > nd::kMaxPrims) return fail(ctx, "mesh exceeds 65535 vertices or faces");
        const double* V = dsc->verts + 3 * size_t(ctx->v_off[b]);
        double mn[3] = {V[0], V[1], V[2]}, mx[3] = {V[0], V[1], V[2]};
        for (int v = 0; v < nvb; ++v)
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], V[3 * v + a]);
                mx[a] = std::max(mx[a], V[3 * v + a]);
            }
        for (int a = 0; a < 3; ++a) { lbox[6 * b + a] = mn[a]; lbox[6 * b + 3 + a] = mx[a]; }
        double vb = 0.0;
        for (int v = 0; v < 3 * nvb; ++v) vb = std::isfinite(V[v]) ? std::max(vb, std::fabs(V[v])) : INFINITY;
        vbound[b] = vb;
        const int* F = dsc->faces + 3 * size_t(ctx->f_off[b]);
        double me = 0.0;
        for (int f = 0; f < nfb; ++f)
            for (int k = 0; k < 3; ++k) {
                const int a = F[3 * f + k], c = F[3 * f + (k + 1) % 3];
                if (a < 0 || a >= nvb || c < 0 || c >= nvb) return fail(ctx, "face index out of range");
                const double dx = V[3 * a] - V[3 * c], dy = V[3 * a + 1] - V[3 * c + 1], dz = V[3 * a + 2] - V[3 * c + 2];
                me = std::max(me, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        max_edge[b] = me;
        const auto key = std::make_pair(ctx->f_off[b], nfb);
        auto it = edge_cache.find(key);
        if (it == edge_cache.end()) {
            const std::vector<int> E = unique_edges(F, nfb);
            if (int(E.size() / 2) > nd::kMaxPrims) return fail(ctx, "mesh exceeds 65535 edges");
            const int off = int(ctx->edges_host.size() / 2);
            ctx->edges_host.insert(ctx->edges_host.end(), E.begin(), E.end());
            it = edge_cache.emplace(key, std::make_pair(off, int(E.size() / 2))).first;
        }
        ctx->e_off[b] = it->second.first;
        ctx->ne[b] = it->second.second;
    }
    // joints
    const int J = dsc->n_joints;
    ctx->J = J;
    std::vector<int> ja(std::max(J, 1), -1), jb(std::max(J, 1), -1), jt(std::max(J, 1), 0);
    std::vector<int> joint_off(S + 1, 0);
    for (int s = 0; s <= S; ++s) joint_off[s] = dsc->scene_joint_off ? dsc->scene_joint_off[s] : 0;
    for (int s = 0; s < S; ++s) {
        for (int j = joint_off[s]; j < joint_off[s + 1]; ++j) {
            jt[j] = dsc->joint_type[j];
            const int a = dsc->joint_body[2 * j], b = dsc->joint_body[2 * j + 1];
            const int n = ctx->body_off[s + 1] - ctx->body_off[s];
            if (a < 0 || a >= n) return fail(ctx, "joint body out of range");
            ja[j] = ctx->body_off[s] + a;
            if (jt[j] == 1) {
                if (b < 0 || b >= n || b == a) return fail(ctx, "hinge joint needs two distinct bodies of its scene");
                jb[j] = ctx->body_off[s] + b;
            } else if (jt[j] != 0) {
                return fail(ctx, "unsupported joint type");
            }
        }
    }
    ctx->joint_off = joint_off;
    // allowed pairs: per-scene sorted keys
    std::vector<int> allowed_off(S + 1, 0);
    std::vector<unsigned long long> akeys;
    std::vector<double> amu;
    for (int s = 0; s < S; ++s) {
        std::vector<std::pair<unsigned long long, double>> kv;
        if (dsc->scene_allowed_off) {
            for (int k = dsc->scene_allowed_off[s]; k < dsc->scene_allowed_off[s + 1]; ++k) {
                const int a = dsc->allowed_pairs[2 * k], b = dsc->allowed_pairs[2 * k + 1];
                const int lo = std::min(a, b), hi = std::max(a, b);
                kv.emplace_back((unsigned long long)(uint32_t)lo << 32 | (uint32_t)hi, dsc->allowed_mu[k]);
            }
        }
        std::sort(kv.begin(), kv.end(), [](auto& x, auto& y) { return x.first < y.first; });
        // IPCSolver::addAllowedContactPair: later calls overwrite the override
        std::vector<std::pair<unsigned long long, double>> u;
        for (auto& e : kv) {
            if (!u.empty() && u.back().first == e.first) {
                if (e.second >= 0.0) u.back().second = e.second;
            } else {
                u.push_back(e);
            }
        }
        for (auto& e : u) { akeys.push_back(e.first); amu.push_back(e.second); }
        allowed_off[s + 1] = int(akeys.size());
    }
    // device scene data
    Dev& d = ctx->d;
    d = Dev{};
    d.S = S;
    d.N = N;
    d.max_nb = 0;
    for (int s = 0; s < S; ++s) d.max_nb = std::max(d.max_nb, ctx->body_off[s + 1] - ctx->body_off[s]);
    d.max_nv = d.max_nf = 1;
    for (int b = 0; b < N; ++b) {
        d.max_nv = std::max(d.max_nv, ctx->nv[b]);
        d.max_nf = std::max(d.max_nf, ctx->nf[b]);
    }
    d.body_off = dupload(ctx, ctx->body_off.data(), S + 1);
    d.body_scene = dupload(ctx, body_scene.data(), N);
    d.body_id = dupload(ctx, ctx->body_id.data(), N);
    d.v_off = dupload(ctx, ctx->v_off.data(), N);
    d.nv = dupload(ctx, ctx->nv.data(), N);
    d.f_off = dupload(ctx, ctx->f_off.data(), N);
    d.nf = dupload(ctx, ctx->nf.data(), N);
    d.e_off = dupload(ctx, ctx->e_off.data(), N);
    d.ne = dupload(ctx, ctx->ne.data(), N);
    d.verts = dupload(ctx, dsc->verts, 3 * size_t(total_v));
    d.faces = dupload(ctx, dsc->faces, 3 * size_t(total_f));
    d.edges = dupload(ctx, ctx->edges_host.data(), ctx->edges_host.size());
    d.lbox = dupload(ctx, lbox.data(), lbox.size());
    d.max_edge = dupload(ctx, max_edge.data(), N);
    d.vbound = dupload(ctx, vbound.data(), N);
    ctx->mass_dev = dupload(ctx, dsc->mass, N);
    d.mass = ctx->mass_dev;
    d.inertia = dupload(ctx, dsc->inertia, 9 * size_t(N));
    d.joint_off = dupload(ctx, joint_off.data(), S + 1);
    d.jtype = dupload(ctx, jt.data(), jt.size());
    d.ja = dupload(ctx, ja.data(), ja.size());
    d.jb = dupload(ctx, jb.data(), jb.size());
    {
        // body -> joint incidence in joint order, for the per-body sums over joints
        std::vector<int> off(size_t(N) + 1, 0), ent;
        for (int j = 0; j < J; ++j) {
            if (ja[j] >= 0) ++off[ja[j] + 1];
            if (jb[j] >= 0) ++off[jb[j] + 1];
        }
        for (int b = 0; b < N; ++b) off[b + 1] += off[b];
        ent.assign(std::max<size_t>(off[N], 1), 0);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int j = 0; j < J; ++j) {
            if (ja[j] >= 0) ent[fill[ja[j]]++] = (j << 1) | 0;
            if (jb[j] >= 0) ent[fill[jb[j]]++] = (j << 1) | 1;
        }
        d.bj_off = dupload(ctx, off.data(), off.size());
        d.bj_ent = dupload(ctx, ent.data(), ent.size());
    }
    std::vector<double> jdata(16 * size_t(std::max(J, 1)), 0.0), jmu(std::max(J, 1), 1e8);
    if (J > 0) {
        std::memcpy(jdata.data(), dsc->joint_data, sizeof(double) * 16 * J);
        std::memcpy(jmu.data(), dsc->joint_mu, sizeof(double) * J);
    }
    d.jdata = dupload(ctx, jdata.data(), jdata.size());
    d.jmu = dupload(ctx, jmu.data(), jmu.size());
    d.jlambda = dalloc<double>(ctx, 6 * size_t(std::max(J, 1)));
    d.jC = dalloc<double>(ctx, 6 * size_t(std::max(J, 1)));
    d.allowed_off = dupload(ctx, allowed_off.data(), S + 1);
    d.allowed_key = dupload(ctx, akeys.data(), akeys.size());
    d.allowed_mu = dupload(ctx, amu.data(), amu.size());
    d.p = dalloc<double>(ctx, 3 * size_t(N));
    d.q = dalloc<double>(ctx, 4 * size_t(N));
    d.v = dalloc<double>(ctx, 3 * size_t(N));
    d.w = dalloc<double>(ctx, 3 * size_t(N));
    d.al = dalloc<double>(ctx, 3 * size_t(N));
    d.aa = dalloc<double>(ctx, 3 * size_t(N));
    d.sp = dalloc<SceneParams>(ctx, S);
    // per-scene prune margin for the scene-min-distance flag
    ctx->sp.assign(S, SceneParams{});
    for (int s = 0; s < S; ++s) {
        default_params(ctx->sp[s]);
        double me = 0.0;
        for (int b = ctx->body_off[s]; b < ctx->body_off[s + 1]; ++b) me = std::max(me, max_edge[b]);
        ctx->sp[s].prune_margin = 1e-8 + 1e-5 * me;
    }
    // work buffers
    Work& w = ctx->w;
    w = Work{};
    const size_t PC = size_t(std::max<long long>(pair_cap, 1));
    ctx->pair_cap = pair_cap;
    w.wbox = dalloc<double>(ctx, 6 * size_t(N));
    w.pose_p = dalloc<double>(ctx, 3 * size_t(N));
    w.pose_q = dalloc<double>(ctx, 4 * size_t(N));
    w.leaf_body = dalloc<int>(ctx, N);
    w.node_child = dalloc<int>(ctx, 2 * size_t(N));
    w.node_range = dalloc<int>(ctx, 2 * size_t(N));
    w.node_box = dalloc<double>(ctx, 6 * size_t(N));
    w.nleaves = dalloc<int>(ctx, S);
    w.bvh_age = dalloc<int>(ctx, S);  // zero: the first call builds
    {
        // line-search helper work-groups for small batches (NDIPC_NEWTON_HELPERS=0: off)
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) ncu = 0;
        const char* env = std::getenv("NDIPC_NEWTON_HELPERS");
        const bool off = env && env[0] == '0';
        w.newton_helpers = (!off && (long long)S * (1 + nd::kHelpers) <= ncu) ? 1 : 0;
        w.hbox = dalloc<nd::HelperBox>(ctx, S);
        w.hsingle = dalloc<int>(ctx, N);
        const nd::HelperCfg cfg{nd::kHelperWaitTicks, nd::kHelperEpochTicks, nd::kHelperIdleTicks,
                                nd::kHelperStraggleTicks};
        w.hcfg = dupload(ctx, &cfg, 1);
        w.round_seq = dalloc<unsigned int>(ctx, 1);
        w.post_count = dalloc<unsigned int>(ctx, 1);
    }
    w.pair_cnt = dalloc<int>(ctx, S);
    w.pair_off = dalloc<int>(ctx, S + 1);
    w.pairs = dalloc<int>(ctx, 2 * PC);
    w.pair_scene = dalloc<int>(ctx, PC);
    w.ccd_list = dalloc<int>(ctx, PC);
    w.ccd_nlist = dalloc<int>(ctx, 1);
    {
        const char* env = std::getenv("NDIPC_CCD_EVAL_ALL");
        w.ccd_bound = (env && env[0] == '1') ? -1.0 : 1e50;
    }
    w.n_pairs_total = dalloc<int>(ctx, 1);
    w.prox_cnt = dalloc<int>(ctx, S);
    w.prox_off = dalloc<int>(ctx, S + 1);
    w.prox_pairs = dalloc<int>(ctx, 2 * PC);
    w.prox_scene = dalloc<int>(ctx, PC);
    w.n_prox_total = dalloc<int>(ctx, 1);
    w.lists = dalloc<int>(ctx, PC * 6 * nd::kTopK);
    w.lcnt = dalloc<int>(ctx, PC * 6);
    w.hits = dalloc<unsigned long long>(ctx, PC * nd::kHitWords);
    w.hit_fv = dalloc<int>(ctx, PC);
    w.hit_ee = dalloc<int>(ctx, PC);
    w.hit_off_fv = dalloc<int>(ctx, PC);
    w.hit_off_ee = dalloc<int>(ctx, PC);
    w.raw_cnt = dalloc<int>(ctx, S);
    w.cint = dalloc<int>(ctx, size_t(S) * nd::kMaxRawContacts * 5);
    w.cdbl = dalloc<double>(ctx, size_t(S) * nd::kMaxRawContacts * 7);
    w.cext = dalloc<int>(ctx, size_t(S) * nd::kMaxRawContacts * 2);
    w.sort_idx = dalloc<int>(ctx, size_t(S) * nd::kSceneScratch);
    w.sort_key = dalloc<double>(ctx, size_t(S) * nd::kSceneScratch);
    const size_t V6 = 6 * size_t(N);
    w.x = dalloc<double>(ctx, V6);
    w.xhat = dalloc<double>(ctx, V6);
    w.xprev = dalloc<double>(ctx, V6);
    w.vprev = dalloc<double>(ctx, V6);
    w.g = dalloc<double>(ctx, V6);
    w.dx = dalloc<double>(ctx, V6);
    w.xt = dalloc<double>(ctx, V6);
    w.cg = dalloc<double>(ctx, V6);
    w.r = dalloc<double>(ctx, V6);
    w.z = dalloc<double>(ctx, V6);
    w.pv = dalloc<double>(ctx, V6);
    w.Ap = dalloc<double>(ctx, V6);
    w.tmp = dalloc<double>(ctx, V6);
    w.D = dalloc<double>(ctx, 36 * size_t(N));
    w.Dinv = dalloc<double>(ctx, 36 * size_t(N));
    w.nf = dalloc<double>(ctx, size_t(S) * nd::kMaxContacts);
    w.cblk = dalloc<double>(ctx, size_t(S) * nd::kMaxContacts * 144);
    w.cgrad = dalloc<double>(ctx, size_t(S) * nd::kMaxContacts * 18);
    w.cactive = dalloc<int>(ctx, size_t(S) * nd::kMaxContacts);
    w.csr_off = dalloc<int>(ctx, size_t(N) + S + 1);
    w.csr_ent = dalloc<int>(ctx, size_t(S) * 2 * nd::kMaxContacts);
    w.jblk = dalloc<double>(ctx, 108 * size_t(std::max(J, 1)));
    w.jgrad = dalloc<double>(ctx, 12 * size_t(std::max(J, 1)));
    w.dense = dalloc<double>(ctx, size_t(S) * nd::kDenseRows * nd::kDenseRows);
    w.flag = dalloc<int>(ctx, S);
    w.bcnt = dalloc<int>(ctx, 2 * size_t(N));
    w.boff = dalloc<int>(ctx, 2 * size_t(N));
    w.bp_cache = dalloc<int>(ctx, size_t(N) * nd::kBpCache);
    // meshes above the LDS-resident size: per work-group HBM slabs of the top-K and
    // scene-min kernels (their grids are kept within the slab counts at launch)
    {
        int max_ne = 1;
        for (int b = 0; b < N; ++b) max_ne = std::max(max_ne, ctx->ne[b]);
        const int M = std::max(max_ne, std::max(d.max_nv, d.max_nf));
        if (M > nd::kTkLdsPrims) {
            w.tk_slab_m = M;
            w.tk_slab_blocks = kTopkSlabBlocks;
            w.tk_slab_stride = nd::tk_slab_bytes(M);
            w.tk_slab = dalloc<unsigned char>(ctx, w.tk_slab_stride * kTopkSlabBlocks);
        }
        if (d.max_nv > nd::kTkLdsPrims || d.max_nf > nd::kTkLdsPrims) {
            w.smf_slab_blocks = kSmfSlabBlocks;
            w.smf_slab_stride = nd::smf_slab_bytes(d.max_nv, d.max_nf);
            w.smf_slab = dalloc<unsigned char>(ctx, w.smf_slab_stride * kSmfSlabBlocks);
        }
    }
    ctx->d_dt = dalloc<double>(ctx, S);
    ctx->scratch_bytes = 0;
    ctx->d_scratch = nullptr;
    if (ctx->alloc_failed) return fail(ctx, "device allocation or upload failed");
    ctx->uploaded = true;
    return sync_sp_to_dev(ctx);
}

int32_t ndipc_configure(ndipc_ctx* ctx, const ndipc_params* p) {
    if (check_ctx(ctx)) return -1;
    if (!p) return fail(ctx, "null params");
    for (int s = 0; s < ctx->S; ++s) {
        SceneParams& q = ctx->sp[s];
        q.dhat = std::max(1e-9, p->dhat);
        if (p->kappa > 0.0) q.kappa = p->kappa;
        q.mu = p->friction_mu;
        q.enable_friction = p->enable_friction != 0;
        q.enable_ccd = p->enable_ccd != 0;
        q.enable_adaptive = p->enable_adaptive_barrier != 0;
        q.alm_max_iters = std::max(1, int(p->alm_max_iters));
        q.alm_ctol = std::max(1e-12, p->alm_constraint_tol);
        q.alm_dtol = std::max(1e-12, p->alm_dual_tol);
        q.alm_ctcol = std::max(1e-12, p->alm_contact_tol);
        q.alm_trig = std::max(1.0, p->alm_hardening_trigger);
        q.alm_ratio = std::max(1.0, p->alm_hardening_ratio);
        q.ccd_safety = p->ccd_safety_factor;
        q.newton_tol = p->newton_tol > 0.0 ? p->newton_tol : 1e-6;
        q.newton_max_iters = p->newton_max_iters > 0 ? int(p->newton_max_iters) : 20;
        if (q.prune_margin > 4.0 * 1e-3 && q.enable_adaptive)
            return fail(ctx, "longest face edge too long for the exact scene-min-distance pruning");
    }
    return sync_sp_to_dev(ctx);
}

int32_t ndipc_set_scene_mu(ndipc_ctx* ctx, const double* mu) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) ctx->sp[s].mu = mu[s];
    return sync_sp_to_dev(ctx);
}

int32_t ndipc_set_joint_state(ndipc_ctx* ctx, const double* lambda6, const double* mu) {
    if (check_ctx(ctx)) return -1;
    if (ctx->J == 0) return 0;
    if (lambda6) HIPCHK(hipMemcpy(ctx->d.jlambda, lambda6, 8 * 6 * size_t(ctx->J), hipMemcpyHostToDevice));
    if (mu) HIPCHK(hipMemcpy(ctx->d.jmu, mu, 8 * size_t(ctx->J), hipMemcpyHostToDevice));
    return 0;
}

int32_t ndipc_set_mass(ndipc_ctx* ctx, const double* mass) {
    if (check_ctx(ctx)) return -1;
    if (!mass) return fail(ctx, "null mass");
    for (int b = 0; b < ctx->N; ++b)
        if (!(mass[b] > 0.0)) return fail(ctx, "masses must be positive");
    HIPCHK(hipMemcpy(ctx->mass_dev, mass, 8 * size_t(ctx->N), hipMemcpyHostToDevice));
    return 0;
}

int32_t ndipc_set_state(ndipc_ctx* ctx, const double* p, const double* q, const double* v, const double* w,
                        const double* al, const double* aa) {
    if (check_ctx(ctx)) return -1;
    const size_t N = ctx->N;
    if (p) HIPCHK(hipMemcpy(ctx->d.p, p, 3 * N * 8, hipMemcpyHostToDevice));
    if (q) HIPCHK(hipMemcpy(ctx->d.q, q, 4 * N * 8, hipMemcpyHostToDevice));
    if (v) HIPCHK(hipMemcpy(ctx->d.v, v, 3 * N * 8, hipMemcpyHostToDevice));
    if (w) HIPCHK(hipMemcpy(ctx->d.w, w, 3 * N * 8, hipMemcpyHostToDevice));
    if (al) HIPCHK(hipMemcpy(ctx->d.al, al, 3 * N * 8, hipMemcpyHostToDevice));
    if (aa) HIPCHK(hipMemcpy(ctx->d.aa, aa, 3 * N * 8, hipMemcpyHostToDevice));
    return 0;
}

int32_t ndipc_get_state(ndipc_ctx* ctx, double* p, double* q, double* v, double* w, double* al, double* aa) {
    if (check_ctx(ctx)) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const size_t N = ctx->N;
    if (p) HIPCHK(hipMemcpy(p, ctx->d.p, 3 * N * 8, hipMemcpyDeviceToHost));
    if (q) HIPCHK(hipMemcpy(q, ctx->d.q, 4 * N * 8, hipMemcpyDeviceToHost));
    if (v) HIPCHK(hipMemcpy(v, ctx->d.v, 3 * N * 8, hipMemcpyDeviceToHost));
    if (w) HIPCHK(hipMemcpy(w, ctx->d.w, 3 * N * 8, hipMemcpyDeviceToHost));
    if (al) HIPCHK(hipMemcpy(al, ctx->d.al, 3 * N * 8, hipMemcpyDeviceToHost));
    if (aa) HIPCHK(hipMemcpy(aa, ctx->d.aa, 3 * N * 8, hipMemcpyDeviceToHost));
    return 0;
}

int32_t ndipc_get_barrier(ndipc_ctx* ctx, double* kappa, double* dhat, int32_t* init) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) {
        if (kappa) kappa[s] = ctx->sp[s].kappa;
        if (dhat) dhat[s] = ctx->sp[s].dhat;
        if (init) init[s] = ctx->sp[s].barrier_initialized;
    }
    return 0;
}

int32_t ndipc_set_barrier(ndipc_ctx* ctx, const double* kappa, const double* dhat, const int32_t* init) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) {
        if (kappa) ctx->sp[s].kappa = kappa[s];
        if (dhat) ctx->sp[s].dhat = dhat[s];
        if (init) ctx->sp[s].barrier_initialized = init[s];
    }
    return sync_sp_to_dev(ctx);
}

int32_t ndipc_get_joint_state(ndipc_ctx* ctx, double* lambda6, double* mu) {
    if (check_ctx(ctx)) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->J == 0) return 0;
    if (lambda6) HIPCHK(hipMemcpy(lambda6, ctx->d.jlambda, 6 * size_t(ctx->J) * 8, hipMemcpyDeviceToHost));
    if (mu) HIPCHK(hipMemcpy(mu, ctx->d.jmu, size_t(ctx->J) * 8, hipMemcpyDeviceToHost));
    return 0;
}

// The kernel sequence of one IPCSolver::step for every scene, in three parts: the
// prefix (candidates, stiffness control, CCD gate, first narrow phase, contact setup),
// one ALM round (Newton, poses, narrow phase at the new iterate, ALM post-pass; scenes
// whose ALM loop has ended skip through device flags) and the suffix (write-back).
// With `rec`, the data the adjoint needs is recorded into slot k of the segment: the
// step-start state, the objective of every Newton solve, the converged variable and the
// solve's iteration count. Kernels only (no copies, fills or host work), except the
// recording copies.
static int enqueue_prefix(ndipc_ctx* ctx, const nd::Adj* rec, int k, bool side = true) {
    const Dev& d = ctx->d;
    const Work& w = ctx->w;
    hipStream_t st = ctx->stream;
    const int mp = max_grid_pairs(ctx);
    const size_t N = size_t(ctx->N);
    if (rec) {
        double* dst = rec->s_state + size_t(k) * 13 * N;
        HIPCHK(hipMemcpyAsync(dst, d.p, 8 * 3 * N, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(dst + 3 * N, d.q, 8 * 4 * N, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(dst + 7 * N, d.v, 8 * 3 * N, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(dst + 10 * N, d.w, 8 * 3 * N, hipMemcpyDeviceToDevice, st));
    }
    nd::launch_step_begin(d, w, ctx->d_dt, st);
    nd::launch_world_boxes(d, w, st);
    nd::launch_lbvh_build(d, w, st);
    nd::launch_barrier_init_ccd(d, w, st);
    nd::launch_broad_phase_step(d, w, st);
    // The stiffness control reads only the step-start poses and the broad phase's pairs,
    // and writes only the scene flag and kappa, which nothing before the contact setup
    // reads: it runs on the third stream beside the candidate build, the CCD gate and the
    // first narrow phase (latency-bound kernels that leave most CUs idle).
    hipStream_t sst = side ? ctx->pstream : st;
    if (side) {
        HIPCHK(hipEventRecord(ctx->pfork, st));
        HIPCHK(hipStreamWaitEvent(ctx->pstream, ctx->pfork, 0));
    }
    nd::timing_begin("scene_min_flag", sst);
    nd::launch_scene_min_flag(d, w, mp, sst);
    nd::timing_end("scene_min_flag", sst);
    nd::launch_kappa(d, w, sst);
    if (side) HIPCHK(hipEventRecord(ctx->pjoin, ctx->pstream));
    nd::timing_begin("topk", st);
    nd::launch_topk(d, w, mp, st);
    nd::timing_end("topk", st);
    nd::launch_ccd_earliest(d, w, mp, st);
    nd::launch_ccd_dt(d, w, st);
    nd::launch_pack(d, w, st);
    nd::launch_pose_from_state(d, w, nullptr, st);
    nd::launch_narrow_phase(d, w, mp, false, st);
    if (side) HIPCHK(hipStreamWaitEvent(st, ctx->pjoin, 0));
    nd::launch_contacts_setup(d, w, st);
    return 0;
}

// One ALM round; with `flag`, the ALM post-pass reports to the host whether every scene's
// ALM loop has ended.
static int enqueue_round(ndipc_ctx* ctx, const nd::Adj* rec, int k, int a, bool flag) {
    const Dev& d = ctx->d;
    const Work& w = ctx->w;
    hipStream_t st = ctx->stream;
    if (rec) nd::launch_adj_snapshot(d, w, *rec, k * rec->A + a, st);
    nd::launch_newton(d, w, st);
    if (rec) nd::launch_adj_snapshot_post(d, w, *rec, k * rec->A + a, st);
    // (k_newton leaves the poses of its iterate for the narrow phase; the ALM post-pass
    // gathers the capped contacts first and reports the round)
    nd::launch_narrow_phase(d, w, max_grid_pairs(ctx), true, st, false);
    nd::launch_alm_post(d, w, st, true, flag ? ctx->hflag : nullptr);
    return 0;
}

// The line-search helpers of the step's Newton launches, on the second stream.
static int fork_helpers(ndipc_ctx* ctx) {
    HIPCHK(hipEventRecord(ctx->hfork, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->hstream, ctx->hfork, 0));
    nd::launch_newton_helpers(ctx->d, ctx->w, ctx->hstream);
    HIPCHK(hipEventRecord(ctx->hjoin, ctx->hstream));
    return 0;
}

static int join_helpers(ndipc_ctx* ctx) {
    nd::launch_helpers_close(ctx->d, ctx->w, ctx->stream);
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->hjoin, 0));
    return 0;
}

// The whole step with a fixed number of ALM rounds (every round enqueued; scenes whose
// loop ended skip through device flags): the adjoint's recorded path and the
// fixed-round mode.
static int enqueue_step(ndipc_ctx* ctx, const nd::Adj* rec, int k, int alm_max) {
    const bool helpers = ctx->w.newton_helpers && alm_max > 0;
    if (enqueue_prefix(ctx, rec, k)) return -1;
    if (helpers && fork_helpers(ctx)) return -1;
    for (int a = 0; a < alm_max; ++a)
        if (enqueue_round(ctx, rec, k, a, false)) return -1;
    if (helpers && join_helpers(ctx)) return -1;
    nd::launch_finalize(ctx->d, ctx->w, ctx->stream);
    return 0;
}

// Captures `fn` (kernels on the context stream only) into an executable graph.
static int capture(ndipc_ctx* ctx, hipGraphExec_t* out, const std::function<int()>& fn) {
    hipStream_t st = ctx->stream;
    ndipc_ctx* saved = g_timing_ctx;
    g_timing_ctx = nullptr;  // no event records inside the graph
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = fn();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    g_timing_ctx = saved;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(ctx, std::string("step capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        *out = nullptr;
        return fail(ctx, std::string("step graph: ") + hipGetErrorString(ei));
    }
    return 0;
}

// Captures the graphs of the step. The step's kernels take only device pointers and
// sizes fixed at upload, and read dt and every per-step value from device memory, so
// one capture serves every later step until the next upload (or, for the fixed-round
// whole-step graph, a change of the ALM iteration count).
static int capture_step(ndipc_ctx* ctx, int alm_max) {
    drop_graph(ctx);
    if (ctx->adaptive) {
        if (capture(ctx, &ctx->pre_exec, [&] { return enqueue_prefix(ctx, nullptr, 0); })) return -1;
        if (capture(ctx, &ctx->round_exec, [&] { return enqueue_round(ctx, nullptr, 0, 0, true); })) return -1;
    } else {
        // the helpers' fork and join are part of the whole-step graph
        if (capture(ctx, &ctx->step_exec, [&] { return enqueue_step(ctx, nullptr, 0, alm_max); })) return -1;
    }
    ctx->graph_alm = alm_max;
    return 0;
}

// Waits (host) until round `seq` has reported; *all_done: every scene's ALM loop ended.
// Only the latest round's report is kept, and the device may be rounds ahead of the
// host; every round issued since the step began belongs to it, and once a round reports
// every loop ended so do all later rounds of the step (a loop that ended stays ended), so
// the latest report answers for round `seq` too (if the latest says not ended, neither
// had round seq).
static int wait_round(ndipc_ctx* ctx, unsigned int seq, int* all_done) {
    volatile unsigned int* f = ctx->hflag;
    for (long long spin = 0;; ++spin) {
        const unsigned int v = __atomic_load_n(&f[0], __ATOMIC_ACQUIRE);
        if (int(v - seq) >= 0) {
            *all_done = int(__atomic_load_n(&f[1], __ATOMIC_ACQUIRE));
            return 0;
        }
        if ((spin & 0xffff) == 0xffff) {  // a stream that has ended or failed will not report
            const hipError_t q = hipStreamQuery(ctx->stream);
            if (q == hipSuccess && int(__atomic_load_n(&f[0], __ATOMIC_ACQUIRE) - seq) < 0)
                return fail(ctx, "ALM round flag: the stream ended without reporting the round");
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(ctx, std::string("ALM round flag: ") + hipGetErrorString(q));
        }
        __builtin_ia32_pause();
    }
}

// One step with as many ALM rounds as the slowest scene needs (at most alm_max): the
// host issues rounds ahead of the device by `round_lookahead` and stops once a round
// reports that every scene's ALM loop has ended; the rounds a scene runs and their
// results are those of the fixed-round step (a round after the end is a no-op).
static int run_step_adaptive(ndipc_ctx* ctx, int alm_max, bool graph) {
    const bool helpers = ctx->w.newton_helpers && alm_max > 0;
    hipStream_t st = ctx->stream;
    if (graph) HIPCHK(hipGraphLaunch(ctx->pre_exec, st));
    else if (enqueue_prefix(ctx, nullptr, 0)) return -1;
    if (helpers && fork_helpers(ctx)) return -1;
    const unsigned int base = ctx->rounds_issued;
    int issued = 0;
    auto issue = [&]() -> int {
        if (graph) HIPCHK(hipGraphLaunch(ctx->round_exec, st));
        else if (enqueue_round(ctx, nullptr, 0, 0, true)) return -1;
        ++issued;
        ++ctx->rounds_issued;
        return 0;
    };
    while (issued < std::min(alm_max, 1 + ctx->round_lookahead))
        if (issue()) return -1;
    for (int r = 0; r < alm_max; ++r) {
        int done = 0;
        if (wait_round(ctx, base + unsigned(r) + 1u, &done)) return -1;
        if (done) break;
        if (issued < alm_max && issue()) return -1;
    }
    if (helpers && join_helpers(ctx)) return -1;
    nd::launch_finalize(ctx->d, ctx->w, st);
    return 0;
}

// One iteration of asynchronous stepping: the write-back and state change of the scenes
// whose ALM loop ended in the previous iteration, the step prefix (its per-step kernels act
// only on the scenes starting a step; the candidate build, a function of each scene's
// step-start state, runs for all) and one ALM round with its report. A scene's kernels
// and their inputs are those of its lockstep step, so its results are the same bits.
// Everything runs on the context stream, enqueued kernel by kernel: an iteration is
// about a millisecond of device work, so launching its ~30 kernels costs the device
// nothing, while replaying the iteration as a captured graph gave, in about one run of
// the 256-scene batch in 80, a contact record mixed from two contacts (cause not found).
static int enqueue_async_kernels(ndipc_ctx* ctx) {
    nd::launch_async_advance(ctx->d, ctx->w, ctx->stream);
    if (enqueue_prefix(ctx, nullptr, 0, false)) return -1;
    return enqueue_round(ctx, nullptr, 0, 0, true);
}

static int enqueue_async_iter(ndipc_ctx* ctx) {
    ctx->w.as_mode = 1;  // kernel arguments are copied at launch
    const int rc = enqueue_async_kernels(ctx);
    ctx->w.as_mode = 0;
    return rc;
}

// n_steps steps of every scene, each scene advancing on its own: iterations are issued
// ahead of the device by `round_lookahead` until one reports that every scene is done or
// in the last round of its last step; then the final write-back. A step takes at most
// alm_max rounds, so n_steps * alm_max iterations always suffice.
static int run_steps_async(ndipc_ctx* ctx, int n_steps, int alm_max) {
    hipStream_t st = ctx->stream;
    nd::launch_async_begin(ctx->d, ctx->w, n_steps, st);
    const long long cap = (long long)n_steps * alm_max;
    const unsigned int base = ctx->rounds_issued;
    long long issued = 0;
    auto issue = [&]() -> int {
        if (enqueue_async_iter(ctx)) return -1;
        ++issued;
        ++ctx->rounds_issued;
        return 0;
    };
    while (issued < std::min(cap, 1LL + ctx->round_lookahead))
        if (issue()) return -1;
    for (long long r = 0; r < cap; ++r) {
        int done = 0;
        if (wait_round(ctx, base + unsigned(r) + 1u, &done)) return -1;
        if (done) break;
        if (issued < cap && issue()) return -1;
    }
    ctx->w.as_mode = 1;
    nd::launch_async_advance(ctx->d, ctx->w, st);
    ctx->w.as_mode = 0;
    return 0;
}

// The step sequence of IPCSolver::step for every scene, n_steps times. When recording
// for the adjoint, or in fixed-round mode, every step enqueues all alm_max rounds (the
// whole step replayed from one graph, or kernel by kernel with timing or graphs off);
// otherwise each step issues only the rounds the scenes need (adaptive mode), from the
// prefix and round graphs or kernel by kernel.
static int run_steps(ndipc_ctx* ctx, int n_steps, const nd::Adj* rec) {
    int alm_max = 1;
    for (auto& q : ctx->sp) alm_max = std::max(alm_max, q.alm_max_iters);
    const bool graph = !rec && ctx->use_graph && !ctx->timing;
    // asynchronous stepping pays for the prefix every round, which only several scenes
    // taking several steps recover; the line-search helpers keep per-step protocols
    if (!rec && ctx->adaptive && ctx->async_steps && ctx->S > 1 && n_steps > 1 && !ctx->w.newton_helpers) {
        g_timing_ctx = ctx;
        if (run_steps_async(ctx, n_steps, alm_max)) return -1;
        HIPCHK(hipGetLastError());
        return 0;
    }
    const bool have = ctx->adaptive ? ctx->pre_exec && ctx->round_exec : ctx->step_exec && ctx->graph_alm == alm_max;
    if (graph && !have)
        if (capture_step(ctx, alm_max)) return -1;
    g_timing_ctx = ctx;
    for (int k = 0; k < n_steps; ++k) {
        if (!rec && ctx->adaptive) {
            if (run_step_adaptive(ctx, alm_max, graph)) return -1;
        } else if (graph) {
            HIPCHK(hipGraphLaunch(ctx->step_exec, ctx->stream));
        } else if (enqueue_step(ctx, rec, k, alm_max)) {
            return -1;
        }
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static int upload_dt(ndipc_ctx* ctx, const double* dt) {
    if (!dt) return fail(ctx, "null dt");
    for (int s = 0; s < ctx->S; ++s)
        if (!(dt[s] > 0.0)) return fail(ctx, "dt must be positive");
    HIPCHK(hipMemcpyAsync(ctx->d_dt, dt, sizeof(double) * ctx->S, hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

int32_t ndipc_step(ndipc_ctx* ctx, const double* dt, int32_t n_steps) {
    if (check_ctx(ctx)) return -1;
    if (upload_dt(ctx, dt)) return -1;
    return run_steps(ctx, n_steps, nullptr);
}

int32_t ndipc_pair_count(ndipc_ctx* ctx, int32_t* out) {
    if (check_ctx(ctx)) return -1;
    HIPCHK(hipMemcpyAsync(out, ctx->w.n_pairs_total, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int32_t ndipc_eval_counter(ndipc_ctx* ctx, double* out) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) out[s] = ctx->sp[s].evals_total;
    return 0;
}

int32_t ndipc_last_stats(ndipc_ctx* ctx, double* out8) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) {
        const SceneParams& q = ctx->sp[s];
        double* o = out8 + 8 * s;
        o[0] = q.alm_iters;
        o[1] = q.newton_iters;
        o[2] = q.n_contacts;
        o[3] = q.n_fv;
        o[4] = q.n_ee;
        o[5] = q.evals;
        o[6] = q.kappa_flag;
        o[7] = q.toi;
    }
    return 0;
}

int32_t ndipc_last_residuals(ndipc_ctx* ctx, double* out3) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) {
        const SceneParams& q = ctx->sp[s];
        out3[3 * s] = q.viol;
        out3[3 * s + 1] = q.dual;
        out3[3 * s + 2] = q.cres;
    }
    return 0;
}

int32_t ndipc_solver_counters(ndipc_ctx* ctx, double* out8) {
    if (check_ctx(ctx)) return -1;
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) {
        const SceneParams& q = ctx->sp[s];
        double* o = out8 + 8 * s;
        o[6] = q.newton_total;
        o[7] = q.trials_total;
        o[0] = q.newton_iters;
        o[1] = q.ls_trials;
        o[2] = q.solves_direct;
        o[3] = q.solves_pcg;
        o[4] = q.pcg_iters;
        o[5] = q.comp_max;
    }
    return 0;
}

int32_t ndipc_helper_config(ndipc_ctx* ctx, int64_t wait_ticks, int64_t epoch_ticks, int64_t idle_ticks,
                            int64_t straggle_ticks) {
    if (check_ctx(ctx)) return -1;
    if (!ctx->w.hcfg) return fail(ctx, "helper_config: no scenes uploaded");
    const nd::HelperCfg cfg{
        wait_ticks >= 0 ? (unsigned long long)wait_ticks : nd::kHelperWaitTicks,
        epoch_ticks >= 0 ? (unsigned long long)epoch_ticks : nd::kHelperEpochTicks,
        idle_ticks >= 0 ? (unsigned long long)idle_ticks : nd::kHelperIdleTicks,
        straggle_ticks >= 0 ? (unsigned long long)straggle_ticks : nd::kHelperStraggleTicks};
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(ctx->w.hcfg, &cfg, sizeof(cfg), hipMemcpyHostToDevice));
    return 0;
}

int32_t ndipc_helper_counters(ndipc_ctx* ctx, int32_t* out4) {
    if (check_ctx(ctx)) return -1;
    if (!ctx->w.hbox) return fail(ctx, "helper_counters: no scenes uploaded");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<nd::HelperBox> hb(ctx->S);
    HIPCHK(hipMemcpy(hb.data(), ctx->w.hbox, sizeof(nd::HelperBox) * ctx->S, hipMemcpyDeviceToHost));
    for (int s = 0; s < ctx->S; ++s) {
        out4[4 * s + 0] = ctx->w.newton_helpers;
        out4[4 * s + 1] = hb[s].n_helped;
        out4[4 * s + 2] = hb[s].n_giveup;
        out4[4 * s + 3] = hb[s].n_straggle;
    }
    return 0;
}

// ---------------------------------------------------------------- function level
static int prep_geometry(ndipc_ctx* ctx) {
    nd::launch_world_boxes(ctx->d, ctx->w, ctx->stream);
    nd::launch_lbvh_build(ctx->d, ctx->w, ctx->stream);
    return 0;
}

int32_t ndipc_detect_body_pairs(ndipc_ctx* ctx, int32_t scene, double r, int32_t* out, int32_t cap, int32_t* n) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    prep_geometry(ctx);
    nd::launch_broad_phase(ctx->d, ctx->w, 3, std::max(0.0, r), false, &scene, ctx->stream);
    int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, ctx->w.n_pairs_total, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<int> pr(2 * size_t(std::max(cnt, 1)));
    if (cnt) HIPCHK(hipMemcpy(pr.data(), ctx->w.pairs, 8 * size_t(cnt), hipMemcpyDeviceToHost));
    const int b0 = ctx->body_off[scene];
    *n = cnt;
    for (int i = 0; i < cnt && i < cap; ++i) {
        out[2 * i] = pr[2 * i] - b0;
        out[2 * i + 1] = pr[2 * i + 1] - b0;
    }
    return 0;
}

int32_t ndipc_min_body_pair_distance(ndipc_ctx* ctx, int32_t scene, double r, double* out) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    prep_geometry(ctx);
    nd::launch_broad_phase(ctx->d, ctx->w, 1, std::max(0.0, r), false, &scene, ctx->stream);
    if (sync_sp_from_dev(ctx)) return -1;
    *out = ctx->sp[scene].min_pair_dist;
    return 0;
}

int32_t ndipc_build_candidates(ndipc_ctx* ctx, int32_t scene, double dhat, int32_t* counts, int32_t* vv, int32_t* ev,
                               int32_t* ee, int32_t* fv, int32_t cap) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    prep_geometry(ctx);
    nd::launch_broad_phase(ctx->d, ctx->w, 0, std::max(0.0, dhat), false, &scene, ctx->stream);
    nd::launch_topk(ctx->d, ctx->w, max_grid_pairs(ctx), ctx->stream);
    int np = 0;
    HIPCHK(hipMemcpyAsync(&np, ctx->w.n_pairs_total, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    std::vector<int> pr(2 * size_t(std::max(np, 1))), L(size_t(std::max(np, 1)) * 6 * nd::kTopK),
        K(size_t(std::max(np, 1)) * 6);
    if (np) {
        HIPCHK(hipMemcpy(pr.data(), ctx->w.pairs, 8 * size_t(np), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(L.data(), ctx->w.lists, 4 * L.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(K.data(), ctx->w.lcnt, 4 * K.size(), hipMemcpyDeviceToHost));
    }
    long long c[4] = {0, 0, 0, 0};
    auto put = [&](int which, int* dst, int width, std::initializer_list<int> vals) {
        if (dst && c[which] < cap) {
            int k = 0;
            for (int v : vals) dst[width * c[which] + k++] = v;
        }
        c[which]++;
    };
    for (int p = 0; p < np; ++p) {
        const int A = pr[2 * p], B = pr[2 * p + 1];
        const int idA = ctx->body_id[A], idB = ctx->body_id[B];
        const int* l = &L[size_t(p) * 6 * nd::kTopK];
        const int* k = &K[size_t(p) * 6];
        auto eAv = [&](int i, int s) { return ctx->edges_host[2 * (ctx->e_off[A] + l[2 * nd::kTopK + i]) + s]; };
        auto eBv = [&](int i, int s) { return ctx->edges_host[2 * (ctx->e_off[B] + l[3 * nd::kTopK + i]) + s]; };
        for (int i = 0; i < k[0]; ++i)
            for (int j = 0; j < k[1]; ++j) put(0, vv, 4, {idA, idB, l[i], l[nd::kTopK + j]});
        for (int i = 0; i < k[2]; ++i)
            for (int j = 0; j < k[1]; ++j) put(1, ev, 5, {idA, idB, eAv(i, 0), eAv(i, 1), l[nd::kTopK + j]});
        for (int i = 0; i < k[3]; ++i)
            for (int j = 0; j < k[0]; ++j) put(1, ev, 5, {idB, idA, eBv(i, 0), eBv(i, 1), l[j]});
        for (int i = 0; i < k[2]; ++i)
            for (int j = 0; j < k[3]; ++j) put(2, ee, 6, {idA, idB, eAv(i, 0), eAv(i, 1), eBv(j, 0), eBv(j, 1)});
        for (int i = 0; i < k[4]; ++i)
            for (int j = 0; j < k[1]; ++j) put(3, fv, 4, {idA, idB, l[4 * nd::kTopK + i], l[nd::kTopK + j]});
        for (int i = 0; i < k[5]; ++i)
            for (int j = 0; j < k[0]; ++j) put(3, fv, 4, {idB, idA, l[5 * nd::kTopK + i], l[j]});
    }
    for (int i = 0; i < 4; ++i) counts[i] = int32_t(c[i]);
    return 0;
}

int32_t ndipc_select_top_k(ndipc_ctx* ctx, const double* scores, int32_t n, int32_t k, int32_t* out) {
    if (!ctx) return -1;
    HIPCHK(hipSetDevice(ctx->device));
    if (n < 0 || n > nd::kMaxPrims) return fail(ctx, "n must be in [0, 65535]");
    if (std::min(n, k) > nd::kTkLdsPrims) return fail(ctx, "min(n, k) must be at most 2048");
    double* ds = nullptr;
    int* di = nullptr;
    unsigned char* slab = nullptr;
    ScratchBufs sb;  // freed on every return
    HIPCHK(sb.alloc(&ds, sizeof(double) * std::max(n, 1)));
    HIPCHK(sb.alloc(&di, sizeof(int) * (std::max(n, 1) + 1)));
    if (n > nd::kTkLdsPrims) HIPCHK(sb.alloc(&slab, nd::tk_slab_bytes(n)));
    HIPCHK(hipMemcpy(ds, scores, sizeof(double) * n, hipMemcpyHostToDevice));
    nd::launch_select_topk_scores(ds, n, k, di + 1, di, slab, ctx->stream);
    int m = 0;
    HIPCHK(hipMemcpyAsync(&m, di, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (m > 0) HIPCHK(hipMemcpy(out, di + 1, 4 * size_t(m), hipMemcpyDeviceToHost));
    return m;
}

int32_t ndipc_narrow_phase(ndipc_ctx* ctx, int32_t scene, const double* x, int32_t cap, int32_t* n, int32_t* ints,
                           double* dbl) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    const int b0 = ctx->body_off[scene], nb = ctx->body_off[scene + 1] - b0;
    if (x) HIPCHK(hipMemcpy(ctx->w.xt + 6 * size_t(b0), x, 48 * size_t(nb), hipMemcpyHostToDevice));
    // ungated pose: every body of the scene
    if (sync_sp_from_dev(ctx)) return -1;
    std::vector<int> saved(ctx->S);
    for (int s = 0; s < ctx->S; ++s) { saved[s] = ctx->sp[s].alm_done; ctx->sp[s].alm_done = 0; ctx->sp[s].evals = 0; }
    if (sync_sp_to_dev(ctx)) return -1;
    nd::launch_pose_from_state(ctx->d, ctx->w, x ? ctx->w.xt : nullptr, ctx->stream);
    nd::launch_narrow_phase(ctx->d, ctx->w, max_grid_pairs(ctx), false, ctx->stream);
    if (sync_sp_from_dev(ctx)) return -1;
    const int nc = ctx->sp[scene].n_contacts;
    *n = nc;
    const size_t cb = size_t(scene) * nd::kMaxRawContacts;
    const int m = std::min(nc, cap);
    if (m > 0) {
        std::vector<int> ci(5 * size_t(m));
        HIPCHK(hipMemcpy(ci.data(), ctx->w.cint + 5 * cb, 20 * size_t(m), hipMemcpyDeviceToHost));
        if (dbl) HIPCHK(hipMemcpy(dbl, ctx->w.cdbl + 7 * cb, 56 * size_t(m), hipMemcpyDeviceToHost));
        if (ints)
            for (int i = 0; i < m; ++i) {
                ints[5 * i] = ci[5 * i] - b0;
                ints[5 * i + 1] = ci[5 * i + 1] - b0;
                for (int k = 2; k < 5; ++k) ints[5 * i + k] = ci[5 * i + k];
            }
    }
    for (int s = 0; s < ctx->S; ++s) ctx->sp[s].alm_done = saved[s];
    return sync_sp_to_dev(ctx);
}

int32_t ndipc_scene_min_distance_flag(ndipc_ctx* ctx, int32_t scene, int32_t* out_flag) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    prep_geometry(ctx);
    if (sync_sp_from_dev(ctx)) return -1;
    const double r = ctx->sp[scene].prune_margin;
    nd::launch_broad_phase(ctx->d, ctx->w, 1, r, false, &scene, ctx->stream);
    // force the flag kernel's gate open for this scene (min gap below 2 dhat)
    if (sync_sp_from_dev(ctx)) return -1;
    const double saved = ctx->sp[scene].min_pair_dist;
    ctx->sp[scene].min_pair_dist = -1.0;
    if (sync_sp_to_dev(ctx)) return -1;
    HIPCHK(hipMemsetAsync(ctx->w.flag, 0, sizeof(int) * ctx->S, ctx->stream));
    nd::launch_scene_min_flag(ctx->d, ctx->w, max_grid_pairs(ctx), ctx->stream, 1);
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, ctx->w.flag + scene, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out_flag = f;
    ctx->sp[scene].min_pair_dist = saved;
    return sync_sp_to_dev(ctx);
}

static int upload_contacts(ndipc_ctx* ctx, int scene, int nc, const int32_t* ints, const double* dbl,
                           const double* nf) {
    if (nc < 0 || nc > nd::kMaxContacts) return fail(ctx, "contact count must be in [0, 2000]");
    const int b0 = ctx->body_off[scene], nb = ctx->body_off[scene + 1] - b0;
    std::vector<int> ci(5 * size_t(std::max(nc, 1)));
    for (int i = 0; i < nc; ++i) {
        if (ints[5 * i] < 0 || ints[5 * i] >= nb || ints[5 * i + 1] < 0 || ints[5 * i + 1] >= nb)
            return fail(ctx, "contact body out of range");
        ci[5 * i] = ints[5 * i] + b0;
        ci[5 * i + 1] = ints[5 * i + 1] + b0;
        for (int k = 2; k < 5; ++k) ci[5 * i + k] = ints[5 * i + k];
    }
    const size_t cb = size_t(scene) * nd::kMaxRawContacts;
    if (nc) {
        HIPCHK(hipMemcpy(ctx->w.cint + 5 * cb, ci.data(), 20 * size_t(nc), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->w.cdbl + 7 * cb, dbl, 56 * size_t(nc), hipMemcpyHostToDevice));
        if (nf) HIPCHK(hipMemcpy(ctx->w.nf + size_t(scene) * nd::kMaxContacts, nf, 8 * size_t(nc),
                                 hipMemcpyHostToDevice));
    }
    if (sync_sp_from_dev(ctx)) return -1;
    ctx->sp[scene].n_contacts = nc;
    return sync_sp_to_dev(ctx);
}

static int forms_eval(ndipc_ctx* ctx, int scene, int which, int nc, const int32_t* ints, const double* dbl,
                      const double* nf, const double* x, double dhat, double kappa, double mu, double eps, double dt,
                      double* value, double* grad, double* hess, double* nfo) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    if (upload_contacts(ctx, scene, nc, ints, dbl, nf)) return -1;
    const int b0 = ctx->body_off[scene], nb = ctx->body_off[scene + 1] - b0;
    const size_t nd6 = 6 * size_t(nb);
    const size_t need = 8 * (nd6 + nd6 + 1 + (hess ? nd6 * nd6 : 0) + nd::kMaxContacts);
    double* buf = nullptr;
    ScratchBufs sb;  // freed on every return
    HIPCHK(sb.alloc(&buf, need));
    double* dx = buf;
    double* dg = buf + nd6;
    double* dv = buf + 2 * nd6;
    double* dnf = dv + 1;
    double* dh = hess ? dnf + nd::kMaxContacts : nullptr;
    HIPCHK(hipMemcpy(dx, x, 8 * nd6, hipMemcpyHostToDevice));
    nd::launch_forms_eval(ctx->d, ctx->w, scene, which, dhat, kappa, mu, eps, dt, dx, dv, dg, dh, nfo ? dnf : nullptr,
                          ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (value) HIPCHK(hipMemcpy(value, dv, 8, hipMemcpyDeviceToHost));
    if (grad) HIPCHK(hipMemcpy(grad, dg, 8 * nd6, hipMemcpyDeviceToHost));
    if (hess) HIPCHK(hipMemcpy(hess, dh, 8 * nd6 * nd6, hipMemcpyDeviceToHost));
    if (nfo && nc) HIPCHK(hipMemcpy(nfo, dnf, 8 * size_t(nc), hipMemcpyDeviceToHost));
    return 0;
}

int32_t ndipc_barrier_eval(ndipc_ctx* ctx, int32_t scene, int32_t nc, const int32_t* ints, const double* dbl,
                           const double* x, double dhat, double kappa, double* value, double* grad, double* hess,
                           double* normal_forces) {
    return forms_eval(ctx, scene, 0, nc, ints, dbl, nullptr, x, dhat, kappa, 0.0, 1e-8, 0.0, value, grad, hess,
                      normal_forces);
}

int32_t ndipc_friction_eval(ndipc_ctx* ctx, int32_t scene, int32_t nc, const int32_t* ints, const double* dbl,
                            const double* nf, const double* x, double mu, double eps, double dt, double* value,
                            double* grad, double* hess) {
    if (!nf && nc > 0) return fail(ctx, "normal forces required");
    return forms_eval(ctx, scene, 1, nc, ints, dbl, nf, x, 1e-3, 0.0, mu, eps, dt, value, grad, hess, nullptr);
}

int32_t ndipc_ipc_barrier_eval(ndipc_ctx* ctx, int32_t scene, int32_t np, const int32_t* pairs5, const double* x,
                               double dhat, double kappa, int32_t psd, double* value, int32_t cap, int32_t* n_seg,
                               int32_t* seg_bodies, double* seg_g, double* seg_H, int32_t* n_active) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    if (np < 0 || np > NDIPC_IPC_MAX_PAIRS) return fail(ctx, "pair count out of range");
    if (!(dhat > 0.0)) return fail(ctx, "dhat must be positive");
    if (!value || !n_seg || (cap > 0 && (!seg_bodies || !seg_g || !seg_H))) return fail(ctx, "null output");
    if (np == 0) {
        *value = 0.0;
        *n_seg = 0;
        if (n_active) *n_active = 0;
        return 0;
    }
    if (!pairs5 || !x) return fail(ctx, "null input");
    const int b0 = ctx->body_off[scene], nb = ctx->body_off[scene + 1] - b0;
    const size_t n = size_t(np);
    ScratchBufs sb;  // freed on every return
    nd::IpcEvalBufs B;
    int *ip = nullptr, *iw = nullptr;
    double *dp = nullptr, *dx = nullptr;
    unsigned long long* key = nullptr;
    const size_t nchunk = (n + 1023) / 1024;
    HIPCHK(sb.alloc(&ip, 4 * (5 * n + n + 1 + n + 1 + 1 + 1 + 2 * n + 2 * nchunk + n)));
    HIPCHK(sb.alloc(&key, 8 * n));
    HIPCHK(sb.alloc(&dp, 8 * (n + 12 * n + 144 * n + n + 12 * n + 144 * n + 1)));
    HIPCHK(sb.alloc(&dx, 8 * 6 * size_t(nb)));
    iw = ip;
    B.pairs = iw;
    iw += 5 * n;
    B.act = iw;
    iw += n;
    B.bad = iw;
    iw += 1;
    B.seg_start = iw;
    iw += n + 1;
    B.n_seg = iw;
    iw += 1;
    B.n_act = iw;
    iw += 1;
    B.seg_bodies = iw;
    iw += 2 * n;
    B.chunk_cnt = iw;
    iw += nchunk;
    B.chunk_act = iw;
    iw += nchunk;
    B.alist = iw;
    B.key = key;
    B.x = dx;
    double* w = dp;
    B.pe = w;
    w += n;
    B.pg = w;
    w += 12 * n;
    B.pH = w;
    w += 144 * n;
    B.seg_E = w;
    w += n;
    B.seg_g = w;
    w += 12 * n;
    B.seg_H = w;
    w += 144 * n;
    B.value = w;
    HIPCHK(hipMemcpy(ip, pairs5, 4 * 5 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(B.bad, &np, 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx, x, 8 * 6 * size_t(nb), hipMemcpyHostToDevice));
    nd::launch_ipc_eval(ctx->d, B, b0, nb, np, dhat, kappa, psd, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int bad = np, ns = 0;
    HIPCHK(hipMemcpy(&bad, B.bad, 4, hipMemcpyDeviceToHost));
    if (bad < np) {
        char msg[96];
        std::snprintf(msg, sizeof msg, "ipc pair %d: bad kind, body or primitive index", bad);
        return fail(ctx, msg);
    }
    HIPCHK(hipMemcpy(&ns, B.n_seg, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(value, B.value, 8, hipMemcpyDeviceToHost));
    if (n_active) HIPCHK(hipMemcpy(n_active, B.n_act, 4, hipMemcpyDeviceToHost));
    *n_seg = ns;
    const size_t k = size_t(std::min(ns, std::max(cap, 0)));
    if (k) {
        HIPCHK(hipMemcpy(seg_bodies, B.seg_bodies, 4 * 2 * k, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(seg_g, B.seg_g, 8 * 12 * k, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(seg_H, B.seg_H, 8 * 144 * k, hipMemcpyDeviceToHost));
    }
    return 0;
}

int32_t ndipc_ipc_step(ndipc_ctx* ctx, int32_t scene, double dt, const ndipc_ipc_params* prm, int32_t* stats4) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    if (!prm) return fail(ctx, "null params");
    const ndipc_ipc_params P = *prm;
    if (!(dt > 0.0) || !(P.dhat > 0.0) || !(P.sep >= 0.0 && P.sep < 1.0) || P.max_newton < 0 || P.max_ls < 1)
        return fail(ctx, "bad ipc step parameters");
    const int b0 = ctx->body_off[scene], nb = ctx->body_off[scene + 1] - b0, nd = 6 * nb;
    if (nb == 0) return 0;
    // static bodies: those pinned by a single-body fixed joint; other joints are not run here
    const int j0 = ctx->joint_off[scene], j1 = ctx->joint_off[scene + 1];
    std::vector<int> stat(nb, 0);
    if (j1 > j0) {
        std::vector<int> jt(j1 - j0), ja(j1 - j0), jb(j1 - j0);
        HIPCHK(hipMemcpy(jt.data(), ctx->d.jtype + j0, 4 * size_t(j1 - j0), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ja.data(), ctx->d.ja + j0, 4 * size_t(j1 - j0), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(jb.data(), ctx->d.jb + j0, 4 * size_t(j1 - j0), hipMemcpyDeviceToHost));
        for (int j = 0; j < j1 - j0; ++j) {
            if (jt[j] != 0 || jb[j] >= 0) return fail(ctx, "ipc step: only single-body fixed joints are supported");
            stat[ja[j] - b0] = 1;
        }
    }
    // work space kept in the context across steps (grown on demand)
    int* dstat = nullptr;
    HIPCHK(grow(ctx->ipc_buf[0], 4 * size_t(nb), &dstat));
    HIPCHK(hipMemcpy(dstat, stat.data(), 4 * size_t(nb), hipMemcpyHostToDevice));
    // pair records of the close pairs' top-K lists
    prep_geometry(ctx);
    nd::launch_broad_phase(ctx->d, ctx->w, 0, std::max(0.0, P.inflation), false, &scene, ctx->stream);
    nd::launch_topk(ctx->d, ctx->w, max_grid_pairs(ctx), ctx->stream);
    int ncl = 0;
    HIPCHK(hipMemcpyAsync(&ncl, ctx->w.n_pairs_total, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int *dcnt = nullptr, *doff = nullptr;
    HIPCHK(grow(ctx->ipc_buf[1], 8 * size_t(std::max(ncl, 1)), &dcnt));
    doff = dcnt + std::max(ncl, 1);
    nd::launch_ipc_emit(ctx->w, ncl, dstat, b0, dcnt, nullptr, nullptr, 0, ctx->stream);
    std::vector<int> cnt(std::max(ncl, 1), 0), off(std::max(ncl, 1), 0);
    if (ncl) HIPCHK(hipMemcpy(cnt.data(), dcnt, 4 * size_t(ncl), hipMemcpyDeviceToHost));
    long long np_ll = 0;
    for (int p = 0; p < ncl; ++p) {
        off[p] = int(np_ll);
        np_ll += cnt[p];
    }
    if (np_ll > NDIPC_IPC_MAX_PAIRS) return fail(ctx, "ipc step: more pair records than NDIPC_IPC_MAX_PAIRS");
    const int np = int(np_ll);
    const size_t n = size_t(std::max(np, 1)), nchunk = (n + 1023) / 1024, nsg = size_t(std::max(ncl, 1));
    nd::IpcEvalBufs B;
    int *ip = nullptr, *offij = nullptr;
    unsigned long long* key = nullptr;
    double *dp = nullptr, *v = nullptr;
    HIPCHK(grow(ctx->ipc_buf[2], 4 * (5 * n + n + 1 + n + 1 + 1 + 1 + 2 * n + 2 * nchunk + n + n), &ip));
    HIPCHK(grow(ctx->ipc_buf[3], 8 * n, &key));
    HIPCHK(grow(ctx->ipc_buf[4], 8 * (n + 12 * n + 144 * n + n + 12 * n + 144 * n + 1 + n), &dp));
    // off-diagonal block pairs (2 per segment), then the PCG's row lists (nb + 1 + 2 per segment)
    HIPCHK(grow(ctx->ipc_buf[5], 4 * (2 * nsg + size_t(nb) + 1 + 2 * nsg), &offij));
    int* pcgi = offij + 2 * nsg;
    // x, x^, xt, dx, g, rhs (nd each), diag (36 nb), offb (36 per segment), ebody, gterm
    // (nb each), sums (4), PCG work (10 nd), the advance's out (1)
    HIPCHK(grow(ctx->ipc_buf[6], 8 * (6 * size_t(nd) + 36 * size_t(nb) + 36 * nsg + 2 * size_t(nb) + 4 + 10 * size_t(nd) + 1),
                &v));
    int* iw = ip;
    int* dpairs = iw;
    iw += 5 * n;
    B.pairs = dpairs;
    B.act = iw; iw += n;
    B.bad = iw; iw += 1;
    B.seg_start = iw; iw += n + 1;
    B.n_seg = iw; iw += 1;
    B.n_act = iw; iw += 1;
    B.seg_bodies = iw; iw += 2 * n;
    B.chunk_cnt = iw; iw += nchunk;
    B.chunk_act = iw; iw += nchunk;
    B.alist = iw; iw += n;
    int* ast = iw;  // the advance's per-pair status
    B.key = key;
    double* w = dp;
    B.pe = w; w += n;
    B.pg = w; w += 12 * n;
    B.pH = w; w += 144 * n;
    B.seg_E = w; w += n;
    B.seg_g = w; w += 12 * n;
    B.seg_H = w; w += 144 * n;
    B.value = w; w += 1;
    double* atoi = w;
    double* x = v;
    double* xh = x + nd;
    double* xt = xh + nd;
    double* dx = xt + nd;
    double* g = dx + nd;
    double* rhs = g + nd;
    double* diag = rhs + nd;
    double* offb = diag + 36 * size_t(nb);
    double* ebody = offb + 36 * nsg;
    double* gterm = ebody + nb;
    double* sums = gterm + nb;
    double* pcgw = sums + 4;
    double* aout = pcgw + 10 * size_t(nd);
    int* piters = B.bad;  // reused: the PCG's iteration count (B.bad is set before every evaluation)
    double prel = 0.0;
    (void)prel;
    if (np) {
        HIPCHK(hipMemcpy(doff, off.data(), 4 * size_t(ncl), hipMemcpyHostToDevice));
        nd::launch_ipc_emit(ctx->w, ncl, dstat, b0, dcnt, doff, dpairs, 1, ctx->stream);
    }
    const double h = dt, h2 = dt * dt;
    nd::launch_ipc_start(ctx->d, b0, nb, h, x, xh, ctx->stream);
    // energy of the moving bodies + h^2 E_ipc at xx (full: also gradient terms and blocks)
    auto evaluate = [&](const double* xx, bool full, double& E) -> int {
        B.x = xx;
        double ev = 0.0;
        if (np) {
            HIPCHK(hipMemcpy(B.bad, &np, 4, hipMemcpyHostToDevice));
            nd::launch_ipc_eval(ctx->d, B, b0, nb, np, P.dhat, P.kappa, 1, ctx->stream, full ? 0 : 1);
        } else {
            const int zero = 0;
            HIPCHK(hipMemcpy(B.n_seg, &zero, 4, hipMemcpyHostToDevice));
        }
        nd::launch_ipc_body(ctx->d, b0, nb, dstat, xx, xh, h2, full ? 0 : 1, B, g, diag, ebody, gterm, ctx->stream);
        nd::launch_ipc_sums(nb, dstat, ebody, gterm, nullptr, nullptr, sums, ctx->stream);
        double sm[2];
        HIPCHK(hipMemcpyAsync(sm, sums, 16, hipMemcpyDeviceToHost, ctx->stream));
        if (np) HIPCHK(hipMemcpyAsync(&ev, B.value, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        E = (sm[0] + h2 * sm[1]) + h2 * ev;
        return 0;
    };
    int newton = 0, trials = 0, pcg_total = 0;
    for (int it = 0; it < P.max_newton; ++it) {
        double E0 = 0.0;
        if (evaluate(x, true, E0)) return -1;
        int nseg = 0;
        HIPCHK(hipMemcpy(&nseg, B.n_seg, 4, hipMemcpyDeviceToHost));
        nd::launch_ipc_offblocks(nd, nseg, dstat, h2, B, g, offij, offb, rhs, ctx->stream);
        nd::launch_pcg_rows(nb, diag, nseg, offij, offb, rhs, P.pcg_tol, P.pcg_max_it, dx, piters, sums + 3, pcgw,
                            pcgi, ctx->stream);
        int pit = 0;
        HIPCHK(hipMemcpyAsync(&pit, piters, 4, hipMemcpyDeviceToHost, ctx->stream));
        nd::launch_ipc_sums(nb, dstat, ebody, gterm, g, dx, sums, ctx->stream);
        double sm[4];
        HIPCHK(hipMemcpyAsync(sm, sums, 32, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        pcg_total += pit;
        const double slope = sm[2], dmax = sm[3];
        if (!(dmax > P.newton_tol)) break;
        if (!(slope < 0.0)) break;
        double alpha = 1.0;
        if (np) {
            const double inf = 1.0 / 0.0;
            HIPCHK(hipMemcpy(aout, &inf, 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(B.bad, &np, 4, hipMemcpyHostToDevice));
            nd::launch_ipc_advance(ctx->d, b0, nb, np, dpairs, x, dx, P.sep, atoi, ast, aout, B.bad, ctx->stream);
            double toi = 0.0;
            HIPCHK(hipMemcpyAsync(&toi, aout, 8, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            alpha = toi < 1.0 ? toi * P.safety : 1.0;
            alpha = std::min(1.0, alpha);
        }
        bool ok = false;
        for (int k = 0; k < P.max_ls; ++k) {
            ++trials;
            nd::launch_ipc_axpy(nd, x, alpha, dx, xt, ctx->stream);
            double Et = 0.0;
            if (evaluate(xt, false, Et)) return -1;
            if (Et <= E0 + 1e-4 * alpha * slope) {
                ok = true;
                break;
            }
            alpha *= 0.5;
        }
        if (!ok) break;
        std::swap(x, xt);
        ++newton;
    }
    nd::launch_ipc_writeback(ctx->d, b0, nb, dstat, x, h, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (stats4) {
        stats4[0] = newton;
        stats4[1] = trials;
        stats4[2] = np;
        stats4[3] = pcg_total;
    }
    return 0;
}

int32_t ndipc_ipc_max_step(ndipc_ctx* ctx, int32_t scene, int32_t np, const int32_t* pairs5, const double* x,
                           const double* dx, double sep, double safety, double* alpha, double* toi) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    if (np < 0 || np > NDIPC_IPC_MAX_PAIRS) return fail(ctx, "pair count out of range");
    if (!(sep >= 0.0 && sep < 1.0)) return fail(ctx, "sep must be in [0, 1)");
    if (!alpha || !toi) return fail(ctx, "null output");
    if (np == 0) {
        *toi = 1.0 / 0.0;
        *alpha = 1.0;
        return 0;
    }
    if (!pairs5 || !x || !dx) return fail(ctx, "null input");
    const int b0 = ctx->body_off[scene], nb = ctx->body_off[scene + 1] - b0;
    const size_t n = size_t(np), nd6 = 6 * size_t(nb);
    ScratchBufs sb;  // freed on every return
    int* ib = nullptr;
    double* db = nullptr;
    HIPCHK(sb.alloc(&ib, 4 * (5 * n + n + 1)));
    HIPCHK(sb.alloc(&db, 8 * (2 * nd6 + n + 1)));
    int* dpairs = ib;
    int* dst = ib + 5 * n;
    int* dbad = dst + n;
    double* dx0 = db;
    double* ddx = db + nd6;
    double* dtoi = ddx + nd6;
    double* dout_t = dtoi + n;
    HIPCHK(hipMemcpy(dpairs, pairs5, 4 * 5 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx0, x, 8 * nd6, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ddx, dx, 8 * nd6, hipMemcpyHostToDevice));
    const double inf = 1.0 / 0.0;
    HIPCHK(hipMemcpy(dout_t, &inf, 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dbad, &np, 4, hipMemcpyHostToDevice));
    nd::launch_ipc_advance(ctx->d, b0, nb, np, dpairs, dx0, ddx, sep, dtoi, dst, dout_t, dbad, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int bad = np;
    double t = 0.0;
    HIPCHK(hipMemcpy(&bad, dbad, 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&t, dout_t, 8, hipMemcpyDeviceToHost));
    if (bad < np) {
        char msg[96];
        std::snprintf(msg, sizeof msg, "ipc pair %d: bad kind, body or primitive index", bad);
        return fail(ctx, msg);
    }
    *toi = t;
    *alpha = t < 1.0 ? t * safety : 1.0;
    return 0;
}

int32_t ndipc_ccd_toi_batch(ndipc_ctx* ctx, int32_t kind, int32_t n, const double* pa, const double* pb,
                            const double* a, const double* b, const double* earliest, double tol, int32_t* hit,
                            double* toi, int32_t* pops) {
    if (!ctx) return -1;
    HIPCHK(hipSetDevice(ctx->device));
    if (kind < 0 || kind > 4) return fail(ctx, "kind must be 0 to 4");
    if (n <= 0) return 0;
    const size_t dbls = size_t(n) * (14 + 14 + 6 + 9 + 1 + 1);
    double* buf = nullptr;
    int* ib = nullptr;
    ScratchBufs sb;  // freed on every return
    HIPCHK(sb.alloc(&buf, 8 * dbls));
    HIPCHK(sb.alloc(&ib, 8 * size_t(n)));
    double* dpa = buf;
    double* dpb = dpa + 14 * size_t(n);
    double* da = dpb + 14 * size_t(n);
    double* db = da + 6 * size_t(n);
    double* de = db + 9 * size_t(n);
    double* dt = de + size_t(n);
    HIPCHK(hipMemcpy(dpa, pa, 8 * 14 * size_t(n), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dpb, pb, 8 * 14 * size_t(n), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(da, a, 8 * 6 * size_t(n), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(db, b, 8 * 9 * size_t(n), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(de, earliest, 8 * size_t(n), hipMemcpyHostToDevice));
    nd::launch_ccd_batch(kind, n, dpa, dpb, da, db, de, tol, ib, dt, ib + n, ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(hit, ib, 4 * size_t(n), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(toi, dt, 8 * size_t(n), hipMemcpyDeviceToHost));
    if (pops) HIPCHK(hipMemcpy(pops, ib + n, 4 * size_t(n), hipMemcpyDeviceToHost));
    return 0;
}

int32_t ndipc_ccd_earliest_toi(ndipc_ctx* ctx, int32_t scene, double dt, double* out) {
    if (check_ctx(ctx)) return -1;
    if (scene < 0 || scene >= ctx->S) return fail(ctx, "scene out of range");
    if (sync_sp_from_dev(ctx)) return -1;
    for (int s = 0; s < ctx->S; ++s) {
        ctx->sp[s].toi = 1.0;
        ctx->sp[s].dt = dt;
        ctx->sp[s].ccd_any = 1;
    }
    const int saved = ctx->sp[scene].enable_ccd;
    ctx->sp[scene].enable_ccd = 1;
    if (sync_sp_to_dev(ctx)) return -1;
    nd::launch_ccd_earliest(ctx->d, ctx->w, max_grid_pairs(ctx), ctx->stream);
    if (sync_sp_from_dev(ctx)) return -1;
    *out = ctx->sp[scene].toi;
    ctx->sp[scene].enable_ccd = saved;
    return sync_sp_to_dev(ctx);
}

int32_t ndipc_pcg_solve(ndipc_ctx* ctx, int32_t nb, const double* diag, int32_t noff, const int32_t* offij,
                        const double* offb, const double* b, double tol, int32_t max_it, double* x, int32_t* iters,
                        double* rel) {
    if (!ctx) return -1;
    HIPCHK(hipSetDevice(ctx->device));
    if (nb <= 0) return fail(ctx, "no blocks");
    const size_t nd6 = 6 * size_t(nb);
    const size_t dbls = 36 * size_t(nb) + 36 * size_t(std::max(noff, 1)) + nd6 + nd6 + 4 * nd6 + 36 * size_t(nb) + 2;
    double* buf = nullptr;
    int* ib = nullptr;
    ScratchBufs sb;  // freed on every return
    HIPCHK(sb.alloc(&buf, 8 * dbls));
    HIPCHK(sb.alloc(&ib, 4 * (2 * size_t(std::max(noff, 1)) + 1)));
    double* dD = buf;
    double* dO = dD + 36 * size_t(nb);
    double* db = dO + 36 * size_t(std::max(noff, 1));
    double* dx = db + nd6;
    double* scr = dx + nd6;
    double* drel = scr + 4 * nd6 + 36 * size_t(nb);
    HIPCHK(hipMemcpy(dD, diag, 8 * 36 * size_t(nb), hipMemcpyHostToDevice));
    if (noff) {
        HIPCHK(hipMemcpy(dO, offb, 8 * 36 * size_t(noff), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ib + 1, offij, 8 * size_t(noff), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(db, b, 8 * nd6, hipMemcpyHostToDevice));
    nd::launch_pcg(nb, dD, noff, ib + 1, dO, db, tol, max_it, dx, ib, drel, scr, ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(x, dx, 8 * nd6, hipMemcpyDeviceToHost));
    if (iters) HIPCHK(hipMemcpy(iters, ib, 4, hipMemcpyDeviceToHost));
    if (rel) HIPCHK(hipMemcpy(rel, drel, 8, hipMemcpyDeviceToHost));
    return 0;
}

int32_t ndipc_pcg_solve_noalloc(ndipc_ctx* ctx, int32_t nb, const double* d_diag, int32_t noff,
                                const int32_t* d_offij, const double* d_offb, const double* d_b, double tol,
                                int32_t max_it, double* d_x, int32_t* d_iters, double* d_rel) {
    if (!ctx) return -1;
    HIPCHK(hipSetDevice(ctx->device));
    if (nb <= 0) return fail(ctx, "pcg_solve_noalloc: no blocks");
    if (nb > NDIPC_PCG_NOALLOC_MAX_BLOCKS)
        return fail(ctx, "pcg_solve_noalloc: " + std::to_string(nb) + " blocks exceed the maximum of " +
                             std::to_string(NDIPC_PCG_NOALLOC_MAX_BLOCKS));
    if (noff < 0 || (noff > 0 && (!d_offij || !d_offb)))
        return fail(ctx, "pcg_solve_noalloc: bad off-diagonal blocks");
    if (!d_diag || !d_b || !d_x || !d_iters || !d_rel) return fail(ctx, "pcg_solve_noalloc: null pointer");
    double* scr = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_ptr[ctx->device] : nullptr;
    if (!scr) return fail(ctx, "pcg_solve_noalloc: scratch not set up on this device");
    // the kernel writes every work vector before reading it: nothing to clear
    nd::launch_pcg(nb, d_diag, noff, d_offij, d_offb, d_b, tol, max_it, d_x, d_iters, d_rel, scr, ctx->stream);
    HIPCHK(hipGetLastError());
    return 0;
}

}  // extern "C"
namespace {
int snapshot_state(ndipc_ctx* ctx, double* s, double* jst, SceneParams* spb);
int restore_state(ndipc_ctx* ctx, const double* s, const double* jst, const SceneParams* spb);
}  // namespace
extern "C" {

// Launch shapes of the persistent kernels (work-groups of k_topk, the narrow phase,
// k_scene_min_flag and k_ccd_fv): from one saved state, each candidate runs trial_steps
// steps between HIP events (kernels enqueued one by one), best of two runs; one knob at
// a time, the others at their best so far. The state is restored afterwards. Results
// do not depend on these shapes, only the speed does.
int32_t ndipc_autotune(ndipc_ctx* ctx, const double* dt, int32_t trial_steps, int32_t* out4, double* out_ms) {
    if (check_ctx(ctx)) return -1;
    if (trial_steps < 1) return fail(ctx, "trial_steps must be positive");
    if (upload_dt(ctx, dt)) return -1;
    const size_t N = size_t(ctx->N), J = size_t(std::max(ctx->J, 1));
    double *sst = nullptr, *sjt = nullptr;
    SceneParams* ssp = nullptr;
    if (hipMalloc(&sst, 8 * 19 * std::max<size_t>(N, 1)) != hipSuccess ||
        hipMalloc(&sjt, 8 * 7 * J) != hipSuccess || hipMalloc(&ssp, sizeof(SceneParams) * ctx->S) != hipSuccess) {
        (void)hipFree(sst); (void)hipFree(sjt); (void)hipFree(ssp);
        return fail(ctx, "autotune: allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = snapshot_state(ctx, sst, sjt, ssp);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(ctx, "autotune: event creation failed");
    const bool graph = ctx->use_graph, timing = ctx->timing;
    ctx->use_graph = false;
    ctx->timing = false;
    int* knob[4] = {&ctx->w.grid_topk, &ctx->w.grid_np, &ctx->w.grid_smf, &ctx->w.grid_ccd};
    const int cand[4][5] = {{512, 1024, 2048, 4096, 8192},
                            {256, 512, 1024, 2048, 4096},
                            {256, 512, 1024, 2048, 4096},
                            {512, 1024, 2048, 4096, 8192}};
    const int defaults[4] = {4096, 4096, 2048, 4096};
    double best_ms = 1e300;
    for (int k = 0; k < 4 && !rc; ++k) {
        int best = defaults[k];
        double kbest = 1e300;
        for (int c = 0; c < 5 && !rc; ++c) {
            *knob[k] = cand[k][c];
            double t = 1e300;
            for (int rep = 0; rep < 2 && !rc; ++rep) {
                if ((rc = restore_state(ctx, sst, sjt, ssp))) break;
                if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = fail(ctx, "autotune: event"); break; }
                if ((rc = run_steps(ctx, trial_steps, nullptr))) break;
                float ms = 0.f;
                if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    rc = fail(ctx, "autotune: event timing");
                    break;
                }
                t = std::min(t, double(ms) / trial_steps);
            }
            if (t < kbest) {
                kbest = t;
                best = cand[k][c];
            }
        }
        *knob[k] = best;
        best_ms = kbest;
    }
    ctx->use_graph = graph;
    ctx->timing = timing;
    if (!rc) rc = restore_state(ctx, sst, sjt, ssp);
    drop_graph(ctx);  // recaptured with the chosen shapes
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sst); (void)hipFree(sjt); (void)hipFree(ssp);
    if (rc) return rc;
    for (int k = 0; k < 4; ++k)
        if (out4) out4[k] = *knob[k];
    if (out_ms) *out_ms = best_ms;
    return 0;
}

int32_t ndipc_set_launch_grids(ndipc_ctx* ctx, const int32_t* grids4) {
    if (check_ctx(ctx)) return -1;
    if (!grids4) return fail(ctx, "set_launch_grids: null grids");
    for (int k = 0; k < 4; ++k)
        if (grids4[k] < 0 || grids4[k] > (1 << 20)) return fail(ctx, "set_launch_grids: work-group count out of range");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->w.grid_topk = grids4[0];
    ctx->w.grid_np = grids4[1];
    ctx->w.grid_smf = grids4[2];
    ctx->w.grid_ccd = grids4[3];
    drop_graph(ctx);  // recaptured with these shapes
    return 0;
}

int32_t ndipc_use_graph(ndipc_ctx* ctx, int32_t on) {
    if (!ctx) return -1;
    ctx->use_graph = on != 0;
    return 0;
}

int32_t ndipc_set_step_mode(ndipc_ctx* ctx, int32_t adaptive, int32_t lookahead) {
    if (check_ctx(ctx)) return -1;
    if (lookahead < 0) return fail(ctx, "set_step_mode: lookahead must be >= 0");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (adaptive < 0 || adaptive > 2) return fail(ctx, "set_step_mode: mode must be 0, 1 or 2");
    if ((adaptive != 0) != ctx->adaptive) drop_graph(ctx);
    ctx->adaptive = adaptive != 0;
    ctx->async_steps = adaptive == 2;
    ctx->round_lookahead = lookahead;
    return 0;
}

int32_t ndipc_timing_enable(ndipc_ctx* ctx, int32_t on) {
    if (!ctx) return -1;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->tim.clear();
    g_event_next = 0;  // pooled events are reused
    ctx->timing = on != 0;
    g_timing_ctx = ctx;
    return 0;
}

int32_t ndipc_timing_get(ndipc_ctx* ctx, const char* kernel, double* avg_ms, int32_t* count) {
    if (!ctx) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    auto it = ctx->tim.find(kernel);
    if (it == ctx->tim.end() || it->second.ev.empty()) {
        *avg_ms = 0.0;
        *count = 0;
        return 0;
    }
    double tot = 0.0;
    for (auto& e : it->second.ev) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
    }
    *count = int32_t(it->second.ev.size());
    *avg_ms = tot / double(*count);
    return 0;
}

int32_t ndipc_round_ops(ndipc_ctx* ctx, const double* a, const double* b, int32_t n, double* out6) {
    if (!ctx) return -1;
    HIPCHK(hipSetDevice(ctx->device));
    if (n <= 0) return 0;
    double* buf = nullptr;
    ScratchBufs sb;  // freed on every return
    HIPCHK(sb.alloc(&buf, 8 * size_t(n) * 8));
    HIPCHK(hipMemcpy(buf, a, 8 * size_t(n), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(buf + n, b, 8 * size_t(n), hipMemcpyHostToDevice));
    nd::launch_round_ops(buf, buf + n, n, buf + 2 * size_t(n), ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out6, buf + 2 * size_t(n), 8 * 6 * size_t(n), hipMemcpyDeviceToHost));
    return 0;
}

int32_t ndipc_elemfn(ndipc_ctx* ctx, int32_t fn, const double* x, const double* y, int32_t n, double* out) {
    if (!ctx) return -1;
    if (fn < 0 || fn > 4) return fail(ctx, "elemfn: unknown function");
    HIPCHK(hipSetDevice(ctx->device));
    if (n <= 0) return 0;
    double* buf = nullptr;
    ScratchBufs sb;  // freed on every return
    HIPCHK(sb.alloc(&buf, 3 * 8 * size_t(n)));
    HIPCHK(hipMemcpy(buf, x, 8 * size_t(n), hipMemcpyHostToDevice));
    if (y) HIPCHK(hipMemcpy(buf + n, y, 8 * size_t(n), hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(buf + n, 0, 8 * size_t(n)));
    nd::launch_elemfn(fn, buf, buf + n, n, buf + 2 * size_t(n), ctx->stream);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(out, buf + 2 * size_t(n), 8 * size_t(n), hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------- adjoint
}  // extern "C"

namespace {
template <typename T>
T* aalloc(ndipc_ctx* ctx, size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    // a failure leaves a null entry, which the caller's check of adj_allocs finds
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
        ctx->adj_allocs.push_back(nullptr);
        return nullptr;
    }
    ctx->adj_allocs.push_back(p);
    if (hipMemset(p, 0, n * sizeof(T)) != hipSuccess) {
        ctx->adj_allocs.push_back(nullptr);
        return nullptr;
    }
    return static_cast<T*>(p);
}

// Saves the full state (poses, velocities, accelerations, joint multipliers and
// stiffnesses, scene parameters) to s[19N], jst[7J], spb[S]; restore_state reverses it.
int snapshot_state(ndipc_ctx* ctx, double* s, double* jst, SceneParams* spb) {
    const size_t N = size_t(ctx->N), J = size_t(std::max(ctx->J, 1));
    hipStream_t st = ctx->stream;
    const Dev& d = ctx->d;
    double* src[6] = {d.p, d.q, d.v, d.w, d.al, d.aa};
    const size_t len[6] = {3 * N, 4 * N, 3 * N, 3 * N, 3 * N, 3 * N};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpyAsync(s, src[k], 8 * len[k], hipMemcpyDeviceToDevice, st));
        s += len[k];
    }
    HIPCHK(hipMemcpyAsync(jst, d.jlambda, 8 * 6 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(jst + 6 * J, d.jmu, 8 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(spb, d.sp, sizeof(SceneParams) * ctx->S, hipMemcpyDeviceToDevice, st));
    return 0;
}
int restore_state(ndipc_ctx* ctx, const double* s, const double* jst, const SceneParams* spb) {
    const size_t N = size_t(ctx->N), J = size_t(std::max(ctx->J, 1));
    hipStream_t st = ctx->stream;
    const Dev& d = ctx->d;
    double* dst[6] = {d.p, d.q, d.v, d.w, d.al, d.aa};
    const size_t len[6] = {3 * N, 4 * N, 3 * N, 3 * N, 3 * N, 3 * N};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpyAsync(dst[k], s, 8 * len[k], hipMemcpyDeviceToDevice, st));
        s += len[k];
    }
    HIPCHK(hipMemcpyAsync(d.jlambda, jst, 8 * 6 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(d.jmu, jst + 6 * J, 8 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(d.sp, spb, sizeof(SceneParams) * ctx->S, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return sync_sp_from_dev(ctx);
}

int save_checkpoint(ndipc_ctx* ctx, int c) {
    const size_t N = size_t(ctx->N), J = size_t(std::max(ctx->J, 1));
    hipStream_t st = ctx->stream;
    double* s = ctx->ck_state + size_t(c) * 19 * N;
    const Dev& d = ctx->d;
    double* src[6] = {d.p, d.q, d.v, d.w, d.al, d.aa};
    const size_t len[6] = {3 * N, 4 * N, 3 * N, 3 * N, 3 * N, 3 * N};
    for (int k = 0; k < 6; ++k) {
        const hipError_t e = hipMemcpyAsync(s, src[k], 8 * len[k], hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) {
            char buf[256];
            snprintf(buf, sizeof buf, "checkpoint copy %d: dst %p src %p bytes %zu base %p: %s", k, (void*)s,
                     (void*)src[k], 8 * len[k], (void*)ctx->ck_state, hipGetErrorString(e));
            return fail(ctx, buf);
        }
        s += len[k];
    }
    double* jst = ctx->ck_joint + size_t(c) * 7 * J;
    HIPCHK(hipMemcpyAsync(jst, d.jlambda, 8 * 6 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(jst + 6 * J, d.jmu, 8 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(ctx->ck_sp + size_t(c) * ctx->S, d.sp, sizeof(SceneParams) * ctx->S,
                          hipMemcpyDeviceToDevice, st));
    return 0;
}

int load_checkpoint(ndipc_ctx* ctx, int c) {
    const size_t N = size_t(ctx->N), J = size_t(std::max(ctx->J, 1));
    hipStream_t st = ctx->stream;
    const double* s = ctx->ck_state + size_t(c) * 19 * N;
    const Dev& d = ctx->d;
    double* dst[6] = {d.p, d.q, d.v, d.w, d.al, d.aa};
    const size_t len[6] = {3 * N, 4 * N, 3 * N, 3 * N, 3 * N, 3 * N};
    for (int k = 0; k < 6; ++k) {
        HIPCHK(hipMemcpyAsync(dst[k], s, 8 * len[k], hipMemcpyDeviceToDevice, st));
        s += len[k];
    }
    const double* jst = ctx->ck_joint + size_t(c) * 7 * J;
    HIPCHK(hipMemcpyAsync(d.jlambda, jst, 8 * 6 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(d.jmu, jst + 6 * J, 8 * J, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(d.sp, ctx->ck_sp + size_t(c) * ctx->S, sizeof(SceneParams) * ctx->S,
                          hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return sync_sp_from_dev(ctx);
}
}  // namespace

extern "C" {

int32_t ndipc_adjoint_forward(ndipc_ctx* ctx, const double* dt, int32_t n_steps, int32_t checkpoint_every) {
    if (check_ctx(ctx)) return -1;
    if (n_steps < 1 || checkpoint_every < 1) return fail(ctx, "n_steps and checkpoint_every must be positive");
    if (upload_dt(ctx, dt)) return -1;
    free_adjoint(ctx);
    const size_t N = size_t(ctx->N), S = size_t(ctx->S), J = size_t(std::max(ctx->J, 1));
    const int L = std::min(checkpoint_every, n_steps);
    const int nseg = (n_steps + checkpoint_every - 1) / checkpoint_every;
    ctx->adj_every = checkpoint_every;
    ctx->adj_steps = n_steps;
    ctx->adj_nck = nseg + 1;
    ctx->adj_dt.assign(dt, dt + S);
    ctx->ck_state = aalloc<double>(ctx, size_t(ctx->adj_nck) * 19 * N);
    ctx->ck_joint = aalloc<double>(ctx, size_t(ctx->adj_nck) * 7 * J);
    ctx->ck_sp = aalloc<SceneParams>(ctx, size_t(ctx->adj_nck) * S);
    int alm_max = 1;
    for (auto& q : ctx->sp) alm_max = std::max(alm_max, q.alm_max_iters);
    nd::Adj& a = ctx->adj;
    a.L = L;
    a.A = alm_max;
    const size_t LA = size_t(L) * alm_max;
    a.s_state = aalloc<double>(ctx, size_t(L) * 13 * N);
    a.s_x = aalloc<double>(ctx, LA * 6 * N);
    a.s_sp = aalloc<SceneParams>(ctx, LA * S);
    a.s_cint = aalloc<int>(ctx, LA * S * nd::kMaxContacts * 5);
    a.s_cdbl = aalloc<double>(ctx, LA * S * nd::kMaxContacts * 7);
    a.s_cext = aalloc<int>(ctx, LA * S * nd::kMaxContacts * 2);
    a.s_nf = aalloc<double>(ctx, LA * S * nd::kMaxContacts);
    a.s_jl = aalloc<double>(ctx, LA * 6 * J);
    a.s_jm = aalloc<double>(ctx, LA * J);
    a.s_jm_post = aalloc<double>(ctx, LA * J);
    a.s_nit = aalloc<int>(ctx, LA * S);
    a.carry = aalloc<double>(ctx, 6 * N);
    a.ap = aalloc<double>(ctx, 3 * N); a.aq = aalloc<double>(ctx, 4 * N);
    a.av = aalloc<double>(ctx, 3 * N); a.aw = aalloc<double>(ctx, 3 * N);
    a.bp = aalloc<double>(ctx, 3 * N); a.bq = aalloc<double>(ctx, 4 * N);
    a.bv = aalloc<double>(ctx, 3 * N); a.bw = aalloc<double>(ctx, 3 * N);
    a.lbar = aalloc<double>(ctx, 6 * J);
    a.g_mass = aalloc<double>(ctx, N);
    a.g_mu = aalloc<double>(ctx, S);
    a.xbar = aalloc<double>(ctx, 6 * N);
    a.eta = aalloc<double>(ctx, 6 * N);
    a.res = aalloc<double>(ctx, 6 * N);
    a.tmp = aalloc<double>(ctx, 6 * N);
    a.kry = aalloc<double>(ctx, 25 * 6 * N);
    a.cvjp = aalloc<double>(ctx, S * nd::kMaxContacts * 15);
    a.gvjp = aalloc<double>(ctx, S * nd::kMaxContacts * 14);
    a.ejac_c = aalloc<double>(ctx, S * nd::kMaxContacts * 144);
    a.ejac_j = aalloc<double>(ctx, 144 * J);
    a.ye_c = aalloc<double>(ctx, S * nd::kMaxContacts * 12);
    a.ye_j = aalloc<double>(ctx, 12 * J);
    a.yea_c = aalloc<double>(ctx, S * nd::kMaxContacts * 12);
    a.yea_j = aalloc<double>(ctx, 12 * J);
    a.dbase = aalloc<double>(ctx, 36 * N);
    a.jvjp = aalloc<double>(ctx, 8 * J);
    a.jxbar = aalloc<double>(ctx, 12 * J);
    a.dense = aalloc<double>(ctx, S * 2 * nd::kDenseRows * nd::kDenseRows);
    a.fail = aalloc<int>(ctx, S);
    a.fail_info = aalloc<double>(ctx, 5 * S);
    for (void* p : ctx->adj_allocs)
        if (!p) {
            free_adjoint(ctx);
            return fail(ctx, "adjoint allocation failed");
        }
    for (int seg = 0; seg < nseg; ++seg) {
        if (save_checkpoint(ctx, seg)) return -1;
        const int len = std::min(checkpoint_every, n_steps - seg * checkpoint_every);
        if (run_steps(ctx, len, nullptr)) return -1;
    }
    if (save_checkpoint(ctx, nseg)) return -1;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int32_t ndipc_adjoint_backward(ndipc_ctx* ctx, const double* dl_dp, const double* dl_dq, const double* dl_dv,
                               const double* dl_dw, double* g_p0, double* g_q0, double* g_v0, double* g_w0,
                               double* g_mass, double* g_mu) {
    if (check_ctx(ctx)) return -1;
    if (ctx->adj_nck == 0) return fail(ctx, "ndipc_adjoint_forward has not run");
    const size_t N = size_t(ctx->N), S = size_t(ctx->S);
    hipStream_t st = ctx->stream;
    nd::Adj a = ctx->adj;
    const double* in[4] = {dl_dp, dl_dq, dl_dv, dl_dw};
    double* dst[4] = {a.ap, a.aq, a.av, a.aw};
    const size_t len[4] = {3 * N, 4 * N, 3 * N, 3 * N};
    for (int k = 0; k < 4; ++k) {
        if (in[k]) HIPCHK(hipMemcpyAsync(dst[k], in[k], 8 * len[k], hipMemcpyHostToDevice, st));
        else HIPCHK(hipMemsetAsync(dst[k], 0, 8 * len[k], st));
    }
    HIPCHK(hipMemsetAsync(a.g_mass, 0, 8 * N, st));
    HIPCHK(hipMemsetAsync(a.g_mu, 0, 8 * S, st));
    HIPCHK(hipMemsetAsync(a.lbar, 0, 8 * 6 * size_t(std::max(ctx->J, 1)), st));
    HIPCHK(hipMemsetAsync(a.fail, 0, 4 * S, st));
    HIPCHK(hipMemsetAsync(a.carry, 0, 8 * 6 * N, st));
    if (upload_dt(ctx, ctx->adj_dt.data())) return -1;
    const int E = ctx->adj_every, nseg = ctx->adj_nck - 1;
    for (int seg = nseg - 1; seg >= 0; --seg) {
        if (load_checkpoint(ctx, seg)) return -1;
        const int len = std::min(E, ctx->adj_steps - seg * E);
        if (run_steps(ctx, len, &a)) return -1;
        for (int t = len - 1; t >= 0; --t) {
            nd::launch_adj_step(ctx->d, ctx->w, a, t, seg * E + t, st);
            std::swap(a.ap, a.bp);
            std::swap(a.aq, a.bq);
            std::swap(a.av, a.bv);
            std::swap(a.aw, a.bw);
        }
    }
    HIPCHK(hipGetLastError());
    double* outp[4] = {g_p0, g_q0, g_v0, g_w0};
    const double* srcp[4] = {a.ap, a.aq, a.av, a.aw};
    for (int k = 0; k < 4; ++k)
        if (outp[k]) HIPCHK(hipMemcpyAsync(outp[k], srcp[k], 8 * len[k], hipMemcpyDeviceToHost, st));
    if (g_mass) HIPCHK(hipMemcpyAsync(g_mass, a.g_mass, 8 * N, hipMemcpyDeviceToHost, st));
    if (g_mu) HIPCHK(hipMemcpyAsync(g_mu, a.g_mu, 8 * S, hipMemcpyDeviceToHost, st));
    std::vector<int> fl(S);
    std::vector<double> fi(5 * S);
    HIPCHK(hipMemcpyAsync(fl.data(), a.fail, 4 * S, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(fi.data(), a.fail_info, 8 * 5 * S, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    // the adjoint sweep leaves the scene in its recorded state; return to the end state
    if (load_checkpoint(ctx, nseg)) return -1;
    for (size_t s = 0; s < S; ++s)
        if (fl[s]) {
            const double* f = fi.data() + 5 * s;
            char buf[256];
            snprintf(buf, sizeof buf,
                     "adjoint: the transposed solve did not converge in scene %zu (step %d, ALM iteration %d, "
                     "direct solve code %d residual %.3e, GMRES residual %.3e)",
                     s, int(f[0]), int(f[1]), int(f[2]), f[3], f[4]);
            return fail(ctx, buf);
        }
    return 0;
}

}  // extern "C"
