// Internal device data model of the ndipc library.
//
// HBM layout (all FP64 unless noted; N bodies over all scenes, S scenes):
//   per body      : mesh offsets (int), local AABB [6], mass, inertia [9], pose p[3] q[4],
//                   velocities v[3] w[3], accelerations al[3] aa[3]
//   mesh pools    : body-frame vertices [3 per vertex], faces [3 int], unique edges [2 int]
//                   (shared by every body that uses the same mesh)
//   per scene     : SceneParams (settings + per-step control flags)
//   per step      : world AABBs [6/body], LBVH per scene, body pairs (i<j, per scene,
//                   sorted), 6 top-K lists per close pair [6 x 32 int], contacts per
//                   scene (up to 40000 raw, 2000 kept), Newton vectors [6/body].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dmath.h"

namespace nd {

constexpr int kTopK = 32;             // ContactAssembler.cpp:221-223
constexpr int kMaxRawContacts = 40000;  // IPCSolver.cpp:413
constexpr int kMaxContacts = 2000;    // IPCSolver.cpp:414
constexpr int kMaxPrims = 65535;      // per-mesh vertex/edge/face limit (16-bit list indices)
constexpr int kTkLdsPrims = 2048;     // longest top-K list / mesh of the LDS-resident paths
constexpr int kMaxSceneBodies = 4096; // per-scene LBVH limit (LDS bitonic sort)
constexpr int kHitWords = 48;         // 3072 candidate bits per pair (2x1024 FV + 1024 EE)
constexpr int kSceneScratch = 65536;  // per-scene scratch doubles (sort keys, barrier b'')
constexpr int kDenseBodies = 64;      // largest coupled component factored directly
constexpr int kDenseRows = 6 * kDenseBodies;
constexpr int kBpCache = 128;         // broad-phase hits of a body kept by the counting pass
constexpr int kLsBatch = 20;          // line-search trials evaluated together, at most
constexpr int kLsBatchLarge = 8;      // ... for scenes whose 20 trials need several LDS rounds

// Line-search helper work-groups per scene (small batches only; solver.hip,
// k_newton_helper): one per trial of a batch, one for the line search's slopes and one
// for the objective at the current iterate.
constexpr int kHelpers = kLsBatch + 2;

// Mailbox between a scene's Newton work-group and its helpers. Zero-initialised; the
// per-step fields are cleared by k_step_begin before the helpers of a step start.
struct HelperBox {
    int epoch;           // Newton launches opened in this step
    int bseq;            // batches posted in this step (numbers never repeat within a step)
    int K;               // trials in the posted batch
    int job;             // 0: line-search trials (and slopes); 1: gradient and diagonal blocks
    int want_slopes;     // 1: the posted batch also asks for the slopes
    int want_f;          // 1: ... and for the objective and contact values at x
    int quit;            // 1: the current launch's helpers leave
    int dead;            // 1: this step's helpers are gone (a wait ran out on either side)
    int closed;          // 1: the step's Newton launches are over; the helpers leave
    int bopen;           // bseq when the current launch opened (later batches are its own)
    int hep[kHelpers];   // per helper: the epoch it last served
    int ack[kHelpers];   // per helper: the batch it last evaluated
    int busy[kHelpers];  // per helper: the batch it took up (set before it re-checks quit)
    // bounds of every wait (wall_clock64 ticks), copied from HelperCfg at step begin
    unsigned long long wait_ticks, epoch_ticks, idle_ticks, straggle_ticks;
    double alpha;        // step length of trial 0 of the posted batch
    double val[kLsBatch], cval[kLsBatch];
    double slope[2];     // g.dx and its contact part cg.dx
    double fx[2];        // objective and contact value at x
    double gm[kHelpers]; // job 1, per helper: max |g| over its bodies
    int okf[kHelpers];   // job 1, per helper: 0 if a singleton's 6x6 factor failed
    // since upload (never reset): batches the helpers answered, waits the scene gave up,
    // and give-ups that found a helper still busy after straggle_ticks
    int n_helped, n_giveup, n_straggle, pad;
};

// Wait bounds of the helper protocol (host-settable, read at every step's begin), in
// wall_clock64 ticks (100 MHz on gfx950), and their defaults.
constexpr unsigned long long kHelperWaitTicks = 500000;        // 5 ms
constexpr unsigned long long kHelperEpochTicks = 500000;       // 5 ms
constexpr unsigned long long kHelperIdleTicks = 20000000;      // 200 ms
constexpr unsigned long long kHelperStraggleTicks = 10000000;  // 100 ms
struct HelperCfg {
    unsigned long long wait_ticks;      // scene: for an answer
    unsigned long long epoch_ticks;     // helper: for the next Newton launch to open
    unsigned long long idle_ticks;      // helper: for the next batch inside a launch
    unsigned long long straggle_ticks;  // scene, after giving up: for a helper mid-job
};

struct SceneParams {
    // settings
    double dhat, kappa, mu, ccd_safety;
    double alm_ctol, alm_dtol, alm_ctcol, alm_trig, alm_ratio;
    int enable_friction, enable_ccd, enable_adaptive, alm_max_iters;
    int barrier_initialized, pad0, vv_needed, pad1;
    double kappa_min;
    // per-step control
    double dt, adt, f_mu, min_pair_dist, toi, prune_margin, ccd_infl, dhat_cand;
    int needs_forms, has_barrier, has_friction, alm_done;
    int alm_iters, newton_iters, converged, n_contacts;
    int nf_count, kappa_flag, ccd_any, solve_fail;
    double viol, dual, cres;
    double evals;  // narrow-phase candidate evaluations in this step
    double n_fv, n_ee;  // FV / EE candidates built this step
    double evals_total; // candidate evaluations since upload
    // Newton work counters of the last step
    int ls_trials, solves_direct, solves_pcg, pcg_iters;
    int comp_max;
    // asynchronous stepping of a batch: 0 the scene starts a step, 1 it is inside one, 2 it
    // has taken all its steps; the steps it has still to start or finish
    int as_state, as_steps_left, pad4;
    double newton_total, trials_total;  // since upload
    double newton_tol;                  // NewtonSolver tolerance (NewtonSolver.cpp:10-95: 1e-6)
    int newton_max_iters, pad5;         // NewtonSolver iteration cap (20)
};

struct Dev {
    int S, N;
    int max_nb;            // largest scene body count (host-side launch shapes)
    int max_nv, max_nf;    // largest mesh vertex / face counts (host-side LDS sizes)
    // scene structure
    const int* body_off;   // [S+1]
    const int* body_scene; // [N]
    const int* body_id;    // [N]
    const int* v_off;      // [N] into verts (vertices)
    const int* nv;         // [N]
    const int* f_off;      // [N] into faces
    const int* nf;         // [N]
    const int* e_off;      // [N] into edges
    const int* ne;         // [N]
    const double* verts;   // pool, 3 per vertex
    const int* faces;      // pool, 3 per face
    const int* edges;      // pool, 2 per edge
    const double* lbox;    // [N*6] local AABB (min, max)
    const double* max_edge;// [N]
    const double* vbound;  // [N] largest |coordinate| of the body's mesh vertices (+inf if one is not finite)
    const double* mass;    // [N]
    const double* inertia; // [N*9]
    // joints
    const int* joint_off;  // [S+1]
    const int* jtype;      // [J]
    const int* ja;         // [J] global body
    const int* jb;         // [J] global body or -1
    const int* bj_off;     // [N+1] joints of each body: bj_ent[bj_off[b] .. bj_off[b+1])
    const int* bj_ent;     // [2J] (j << 1) | side, j ascending (side 0: body a, 1: body b)
    const double* jdata;   // [J*16]
    double* jmu;           // [J]
    double* jlambda;       // [J*6]
    double* jC;            // [J*6] scratch
    // allowed pairs (keys sorted within each scene)
    const int* allowed_off;                 // [S+1]
    const unsigned long long* allowed_key;  // pair keys (lo << 32 | hi) of body ids
    const double* allowed_mu;
    // state
    double *p, *q, *v, *w, *al, *aa;
    SceneParams* sp;
};

// Per-step work buffers.
struct Work {
    // launch shape of the persistent kernels (work-groups; 0: the built-in default),
    // chosen by ndipc_autotune; results do not depend on them
    int grid_topk, grid_np, grid_smf, grid_ccd;
    double* wbox;            // [N*6] world AABB at step start
    double* pose_p;          // [N*3] pose used by the current narrow phase
    double* pose_q;          // [N*4]
    // LBVH (per scene segment at body_off)
    int* leaf_body;          // [N] sorted leaf -> global body
    int* node_child;         // [N*2] internal node children (>= 0 internal, < 0: ~leaf)
    int* node_range;         // [N*2]
    double* node_box;        // [N*6]
    int* nleaves;            // [S]
    int* bvh_age;            // [S] LBVH builds and refits since the last full build
    // Newton line-search helpers: 1 when every scene gets kHelpers helper work-groups
    // beside its Newton work-group (S * (1 + kHelpers) fit the CUs at one per CU)
    int newton_helpers;
    // asynchronous stepping (a copy of Work passed to the kernels of a batch iteration):
    // the prefix's stateful kernels act only on scenes starting a step (as_state 0), the
    // write-back only on scenes whose ALM loop ended (as_state 1, alm_done)
    int as_mode;
    // face-vertex CCD: the close pairs whose queries are evaluated (the screen's list)
    int* ccd_list;   // [pair capacity]
    int* ccd_nlist;  // [1]
    double ccd_bound;  // the screen's input bound (1e50; negative: every pair is evaluated)
    HelperBox* hbox;         // [S]
    HelperCfg* hcfg;         // [1] wait bounds of the helper protocol
    int* hsingle;            // [N] with helpers: 1 if the body is alone in its component
    // body pairs
    long long* pair_cap_off; // [S+1] capacity offsets (n(n-1)/2 per scene)
    int* pair_cnt;           // [S]
    int* pair_off;           // [S+1] compact offsets (scan of pair_cnt)
    int* pairs;              // [2 * pair capacity] global bodies, compact
    int* pair_scene;         // [pair capacity]
    int* n_pairs_total;      // [1]
    int* prox_cnt;           // [S]
    int* prox_off;           // [S+1]
    int* prox_pairs;         // [2 * pair capacity]
    int* prox_scene;         // [pair capacity]
    int* n_prox_total;       // [1]
    // top-K lists per pair: [pair][6][32]; order vA, vB, eA, eB, fA, fB
    int* lists;
    int* lcnt;               // [pair][6]
    // narrow phase
    unsigned long long* hits;// [pair][kHitWords]
    int* hit_fv;             // [pair]
    int* hit_ee;             // [pair]
    int* hit_off_fv;         // [pair]
    int* hit_off_ee;         // [pair]
    int* raw_cnt;            // [S]
    int* cint;               // [S * 40000 * 5]
    double* cdbl;            // [S * 40000 * 7]
    int* cext;               // [S * 40000 * 2] second edge vertices (EE); -1 -1 FV, -2 -2 VV
    int* sort_idx;           // [S * kSceneScratch] scratch for the 2000-cap sort
    double* sort_key;        // [S * kSceneScratch] sort keys / Newton scratch
    // solver vectors, 6 per body
    double *x, *xhat, *xprev, *vprev, *g, *dx, *xt, *cg, *r, *z, *pv, *Ap, *tmp;
    double* D;               // [N*36]
    double* Dinv;            // [N*36]
    double* nf;              // [S*2000] normal forces
    double* cblk;            // [S*2000*4*36] friction contact blocks AA, BB, AB, BA (scaled by h^2)
    double* cgrad;           // [S*2000*24] barrier gA gB, friction gA gB
    int* cactive;            // [S*2000]
    int* csr_off;            // [N+1 + S] body -> contact entries
    int* csr_ent;            // [S*2000*2]
    double* jblk;            // [J*3*36] hinge blocks AA, BB, AB
    double* jgrad;           // [J*12]
    double* dense;           // [S][kDenseRows^2] component LDL^T factors
    int* flag;               // [S] scene-min flag scratch
    // meshes above kTkLdsPrims vertices, edges or faces (null otherwise): per work-group
    // slabs in HBM for the top-K lists (tk_slab_bytes(tk_slab_m) each) and for the
    // scene-min flag's world and candidate vertices and active faces
    unsigned char* tk_slab;
    size_t tk_slab_stride;
    int tk_slab_m, tk_slab_blocks;
    unsigned char* smf_slab;
    size_t smf_slab_stride;
    int smf_slab_blocks;
    unsigned int* round_seq; // [1] ALM rounds reported since upload
    unsigned int* post_count; // [1] work-groups of the running ALM post-pass that finished
    int* bcnt;               // [2N] per-body pair counts (candidates, prox)
    int* boff;               // [2N] per-body offsets within the scene
    int* bp_cache;           // [N * kBpCache] the counting pass's pair and prox hits of each body
};

// Launch wrappers (defined in the .hip files).
void launch_world_boxes(const Dev& d, const Work& w, hipStream_t st);
void launch_lbvh_build(const Dev& d, const Work& w, hipStream_t st);
// mode 0: count + emit close pairs (inflation r, box gap <= r) into w.pairs
// mode 1: min box gap over pairs at inflation r into sp.min_pair_dist; emit prox pairs
//         whose gap <= sp.prune_margin into w.prox_pairs
// mode 2: any allowed pair at inflation sp-dependent (ccd) into sp.ccd_any
void launch_broad_phase(const Dev& d, const Work& w, int mode, double r, bool use_scene_r, const int* scene_mask,
                        hipStream_t st);
void launch_broad_phase_step(const Dev& d, const Work& w, hipStream_t st);
int geom_configure();  // kernel attributes, once per context creation (no stream capture)
int geom_configure_smf();  // the scene-min kernel's LDS limit; nonzero on failure
void launch_topk(const Dev& d, const Work& w, int max_pairs, hipStream_t st);
void launch_pose_from_state(const Dev& d, const Work& w, const double* x, hipStream_t st);
void launch_narrow_phase(const Dev& d, const Work& w, int max_pairs, bool gate_alm, hipStream_t st,
                         bool gather = true);
void launch_scene_min_flag(const Dev& d, const Work& w, int max_prox, hipStream_t st, int ungated = 0);
void launch_ccd_earliest(const Dev& d, const Work& w, int max_pairs, hipStream_t st);
void launch_ccd_batch(int kind, int n, const double* pa, const double* pb, const double* a, const double* b,
                      const double* earliest, double tol, int* hit, double* toi, int* pops, hipStream_t st);
void launch_round_ops(const double* a, const double* b, int n, double* out, hipStream_t st);
void launch_elemfn(int fn, const double* x, const double* y, int n, double* out, hipStream_t st);
// n <= kTkLdsPrims: `slab` unused; else tk_slab_bytes(n) bytes of device memory
void launch_select_topk_scores(const double* scores, int n, int k, int* out, int* out_n, unsigned char* slab,
                               hipStream_t st);
size_t tk_slab_bytes(int M);                          // top-K HBM slab for lists of up to M entries
size_t smf_slab_bytes(int NV, int NF);                // scene-min HBM slab (meshes above kTkLdsPrims)

// Adjoint of a step (implicit-function theorem at every converged Newton point of the
// step's ALM loop). Slot k*A + a holds ALM iteration a of step k of a checkpoint
// segment: the objective of its Newton solve (scene parameters, contacts, lagged normal
// forces, joint multipliers and stiffnesses), the converged variable and the hardened
// stiffness used in the multiplier update; s_state holds the step-start state.
struct Adj {
    int L, A;               // steps per segment, ALM iterations per step
    double* s_state;        // [L][13N] p(3N) q(4N) v(3N) w(3N)
    double* s_x;            // [L*A][6N]
    SceneParams* s_sp;      // [L*A][S] (alm_done = 1: the scene did not solve in that slot)
    int* s_cint;            // [L*A][S*kMaxContacts*5]
    double* s_cdbl;         // [L*A][S*kMaxContacts*7]
    int* s_cext;            // [L*A][S*kMaxContacts*2]
    double* s_nf;           // [L*A][S*kMaxContacts]
    double* s_jl;           // [L*A][6J] multipliers before the solve
    double* s_jm;           // [L*A][J] stiffness of the solve
    double* s_jm_post;      // [L*A][J] stiffness of the multiplier update after it
    int* s_nit;             // [L*A][S] Newton iterations of the solve (0: no step accepted)
    double* carry;          // [6N] cotangent passed to the previous solve's x*
    double *ap, *aq, *av, *aw;  // cotangent of the state after the step
    double *bp, *bq, *bv, *bw;  // cotangent of the state before the step
    double* lbar;           // [6J] cotangent of the joint multipliers (updated in place)
    double* g_mass;         // [N]
    double* g_mu;           // [S]
    double* xbar;           // [6N] cotangent of the converged variable of a solve
    double* eta;            // [6N] A^-T xbar
    double* res;            // [6N] refinement residual / correction
    double* tmp;            // [6N]
    double* kry;            // [25][6N] Krylov basis of the transposed solve
    double* cvjp;           // [S*kMaxContacts][15]
    double* gvjp;           // [S*kMaxContacts][14] cotangent through the contact geometry
    double* ejac_c;         // [S*kMaxContacts][144] exact friction Jacobian (h^2), transposed
    double* ejac_j;         // [J][144] exact hinge Jacobian, transposed
    double* ye_c;           // [S*kMaxContacts][12] element products of a matrix-vector product
    double* ye_j;           // [J][12]
    double* yea_c;          // [S*kMaxContacts][12] the same with absolute values
    double* yea_j;          // [J][12]
    double* dbase;          // [N][36] inertia + fixed-joint + barrier diagonal blocks
    double* jvjp;           // [J][8] cotangent of the reference rotations
    double* jxbar;          // [J][12] cotangent of x from the multiplier update
    double* dense;          // [S][2][kDenseRows^2] a component's A^T and its LU factors
    int* fail;              // [S] 1: a transposed solve did not converge
    double* fail_info;      // [S][5] first failure: step, ALM iteration, direct-solve code,
                            // its relative residual, GMRES relative residual
};
void launch_adj_snapshot(const Dev& d, const Work& w, const Adj& a, int slot, hipStream_t st);
void launch_adj_snapshot_post(const Dev& d, const Work& w, const Adj& a, int slot, hipStream_t st);
void launch_adj_step(const Dev& d, const Work& w, const Adj& a, int step, int gstep, hipStream_t st);

void launch_step_begin(const Dev& d, const Work& w, const double* dt, hipStream_t st);
void launch_kappa(const Dev& d, const Work& w, hipStream_t st);
void launch_ccd_dt(const Dev& d, const Work& w, hipStream_t st);
void launch_pack(const Dev& d, const Work& w, hipStream_t st);
void launch_contacts_setup(const Dev& d, const Work& w, hipStream_t st);
void launch_newton(const Dev& d, const Work& w, hipStream_t st);
// k_newton's line-search helpers (w.newton_helpers) for the n_launch Newton launches of
// one step, to run beside the step's kernels
void launch_newton_helpers(const Dev& d, const Work& w, hipStream_t st);
void launch_helpers_close(const Dev& d, const Work& w, hipStream_t st);
// The ALM post-pass of a round; with gather, the narrow phase's contact gather first
// (np_cap_gather), and with hflag the round's report to the host (round number, every
// scene's ALM loop ended) by the last work-group to finish.
void launch_alm_post(const Dev& d, const Work& w, hipStream_t st, bool gather, unsigned int* hflag);
void launch_finalize(const Dev& d, const Work& w, hipStream_t st);
// asynchronous stepping: every scene has n_steps steps to take, starting one
void launch_async_begin(const Dev& d, const Work& w, int n_steps, hipStream_t st);
// asynchronous stepping: the write-back of the scenes whose ALM loop ended, then their
// state (the next step, or done)
void launch_async_advance(const Dev& d, const Work& w, hipStream_t st);
void launch_barrier_init_ccd(const Dev& d, const Work& w, hipStream_t st);
// which 0: DistanceBarrierForm, 1: FrictionForm, on the contacts stored for scene s
void launch_forms_eval(const Dev& d, const Work& w, int s, int which, double dhat, double kappa, double mu,
                       double eps, double dt, const double* x, double* value, double* grad, double* hess, double* nf,
                       hipStream_t st);
// ipc mode over np pairs of one scene (bodies b0 .. b0 + nb - 1): per-pair terms,
// segments of consecutive pairs of one body pair, per-segment sums, total energy
struct IpcEvalBufs {
    const int* pairs;              // [5 np] kind, body, body, primitive, primitive
    const double* x;               // [6 nb] Newton variable of the scene
    unsigned long long* key;       // [np]
    int* act;                      // [np]
    int* bad;                      // [1] first invalid pair, np if none (set before launch)
    double *pe, *pg, *pH;          // [np], [12 np], [144 np]
    int* seg_start;                // [np + 1]
    int* n_seg;                    // [1]
    int* n_act;                    // [1] pairs within dhat
    int *chunk_cnt, *chunk_act;    // [ceil(np / 1024)] per-chunk segment starts / active pairs
    int* alist;                    // [np] the pairs within dhat, in pair order
    int* seg_bodies;               // [2 np]
    double *seg_E, *seg_g, *seg_H; // [np], [12 np], [144 np]
    double* value;                 // [1]
};
// value_only: the energies alone (line searches): no gradient or Hessian terms
void launch_ipc_eval(const Dev& d, const IpcEvalBufs& B, int b0, int nb, int np, double dhat, double kappa, int psd,
                     hipStream_t st, int value_only = 0);
// ipc-mode step pieces (ipc.hip): pair records from the top-K lists of the close pairs
// (phase 0: counts per close pair into cnt; phase 1: records from the offsets off), the
// start and inertia point, per-body terms and blocks, off-diagonal blocks and rhs, the
// ordered sums, x + alpha dx, the write-back
void launch_ipc_emit(const Work& w, int ncl, const int* stat, int b0, int* cnt, const int* off, int* out, int phase,
                     hipStream_t st);
void launch_ipc_start(const Dev& d, int b0, int nb, double h, double* x, double* xh, hipStream_t st);
void launch_ipc_body(const Dev& d, int b0, int nb, const int* stat, const double* x, const double* xh, double h2,
                     int energy_only, const IpcEvalBufs& B, double* g, double* diag, double* ebody, double* gterm,
                     hipStream_t st);
void launch_ipc_offblocks(int nd, int nseg_max, const int* stat, double h2, const IpcEvalBufs& B, const double* g,
                          int* offij, double* offb, double* rhs, hipStream_t st);
void launch_ipc_sums(int nb, const int* stat, const double* ebody, const double* gterm, const double* g,
                     const double* dx, double* out, hipStream_t st);
void launch_ipc_axpy(int nd, const double* x, double alpha, const double* dx, double* xt, hipStream_t st);
void launch_ipc_writeback(const Dev& d, int b0, int nb, const int* stat, const double* x, double h, hipStream_t st);
// ipc mode maxStep: conservative advancement of every pair from x along dx; out_t the
// least advance (+inf: none; must hold +inf before the launch), out_bad the first bad
// pair (np: none; must hold np before the launch)
void launch_ipc_advance(const Dev& d, int b0, int nb, int np, const int* pairs, const double* x, const double* dx,
                        double sep, double* toi, int* st, double* out_t, int* out_bad, hipStream_t stream);
// launch_pcg's solve with per-row block lists (iscratch: nb + 1 + 2 noff ints)
void launch_pcg_rows(int nb, const double* diag, int noff, const int* offij, const double* offb, const double* b,
                     double tol, int max_it, double* x, int* iters, double* rel, double* scratch, int* iscratch,
                     hipStream_t st);
void launch_pcg(int nb, const double* diag, int noff, const int* offij, const double* offb, const double* b,
                double tol, int max_it, double* x, int* iters, double* rel, double* scratch, hipStream_t st);

// Kernel timing hooks (ctx.hip): kernels that bench.py rooflines are bracketed with
// HIP events on the launching stream while timing is enabled.
void timing_begin(const char* name, hipStream_t st);
void timing_end(const char* name, hipStream_t st);

#define ND_CHECK_LAUNCH() (void)hipGetLastError()

// ---- shared device helpers
__device__ __forceinline__ V3 ld3(const double* a, int i) { return v3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }
__device__ __forceinline__ Q ld4(const double* a, int i) { return quat(a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]); }
__device__ __forceinline__ void st3(double* a, int i, V3 v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
__device__ __forceinline__ void st4(double* a, int i, Q q) { a[4 * i] = q.w; a[4 * i + 1] = q.x; a[4 * i + 2] = q.y; a[4 * i + 3] = q.z; }
__device__ __forceinline__ Box3 ldbox(const double* b, int i) {
    Box3 r;
    r.mn = v3(b[6 * i], b[6 * i + 1], b[6 * i + 2]);
    r.mx = v3(b[6 * i + 3], b[6 * i + 4], b[6 * i + 5]);
    return r;
}
__device__ __forceinline__ unsigned long long pair_key(int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return ((unsigned long long)(unsigned)lo << 32) | (unsigned)hi;
}
// IPCSolver::isContactPairAllowed (IPCSolver.cpp:283-289): empty list allows all.
__device__ __forceinline__ bool pair_allowed(const Dev& d, int s, int ida, int idb) {
    const int lo = d.allowed_off[s], hi = d.allowed_off[s + 1];
    if (lo == hi) return true;
    const unsigned long long k = pair_key(ida, idb);
    int a = lo, b = hi;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (d.allowed_key[m] < k) a = m + 1;
        else b = m;
    }
    return a < hi && d.allowed_key[a] == k;
}
__device__ __forceinline__ double allowed_mu(const Dev& d, int s, int ida, int idb) {
    const int lo = d.allowed_off[s], hi = d.allowed_off[s + 1];
    const unsigned long long k = pair_key(ida, idb);
    int a = lo, b = hi;
    while (a < b) {
        const int m = (a + b) >> 1;
        if (d.allowed_key[m] < k) a = m + 1;
        else b = m;
    }
    if (a < hi && d.allowed_key[a] == k) return d.allowed_mu[a];
    return -1.0;
}
// World position of body-local vertex lv of global body b at pose (p, q).
// Truncation of a scene's raw contacts to the 2000 that std::sort by distance leaves in
// front (IPCSolver.cpp:571-577). Per-scene scratch (at most kMaxRawContacts = 40000
// entries): in sort_key the distances (320,000 B) and the partition positions lpos, rpos
// (80,000 B each); in sort_idx the 16-bit positions (80,000 B), the block starts (40,000
// B) and kept, the raw positions of the kept contacts in their sorted order (written by
// k_np_cap_order when the raw count exceeds 2000).
constexpr size_t kCapLposB = 8 * size_t(kMaxRawContacts);
constexpr size_t kCapRposB = kCapLposB + 2 * size_t(kMaxRawContacts);
constexpr size_t kCapStartB = 2 * size_t(kMaxRawContacts);
constexpr size_t kCapKeptB = (kCapStartB + size_t(kMaxRawContacts) + 255) & ~size_t(255);
static_assert(kCapRposB + 2 * size_t(kMaxRawContacts) <= sizeof(double) * size_t(kSceneScratch), "cap scratch");
static_assert(kCapKeptB + sizeof(int) * kMaxContacts <= sizeof(int) * size_t(kSceneScratch), "cap scratch");
static_assert(kMaxRawContacts <= 65535, "16-bit positions");

// The scene's contact count after the narrow phase, and above 2000 the gather of the kept
// contacts into the first 2000 slots (through sort_key, free once the order is known).
// Called by every thread of one work-group per scene (any size).
__device__ __forceinline__ void np_cap_gather(const Work& w, SceneParams& sp, int s) {
    const int n = w.raw_cnt[s];
    if (n <= kMaxContacts) {
        if (threadIdx.x == 0) sp.n_contacts = n;
        return;
    }
    const int* kept = reinterpret_cast<const int*>(reinterpret_cast<const unsigned char*>(
                                                       w.sort_idx + (long long)s * kSceneScratch) + kCapKeptB);
    const long long cb = (long long)s * kMaxRawContacts;
    int* ti = reinterpret_cast<int*>(w.sort_key + (long long)s * kSceneScratch);  // 2000 * 5 ints
    int* te = ti + 5 * kMaxContacts;                                                  // 2000 * 2 ints
    double* td = w.sort_key + (long long)s * kSceneScratch + 8192;                   // 2000 * 7 doubles
    for (int i = threadIdx.x; i < kMaxContacts; i += blockDim.x) {
        const int src = kept[i];
        for (int k = 0; k < 7; ++k) td[7 * i + k] = w.cdbl[7 * (cb + src) + k];
        for (int k = 0; k < 5; ++k) ti[5 * i + k] = w.cint[5 * (cb + src) + k];
        for (int k = 0; k < 2; ++k) te[2 * i + k] = w.cext[2 * (cb + src) + k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMaxContacts; i += blockDim.x) {
        for (int k = 0; k < 7; ++k) w.cdbl[7 * (cb + i) + k] = td[7 * i + k];
        for (int k = 0; k < 5; ++k) w.cint[5 * (cb + i) + k] = ti[5 * i + k];
        for (int k = 0; k < 2; ++k) w.cext[2 * (cb + i) + k] = te[2 * i + k];
    }
    if (threadIdx.x == 0) sp.n_contacts = kMaxContacts;
}

// Kernels of the step prefix that change a scene's per-step state skip, in asynchronous
// stepping, the scenes that are inside a step or done.
__device__ __forceinline__ bool as_skip(const Work& w, const SceneParams& sp) { return w.as_mode && sp.as_state != 0; }

__device__ __forceinline__ V3 world_vertex(const Dev& d, int b, int lv, V3 p, Q q) {
    return p + rotate(q, ld3(d.verts, d.v_off[b] + lv));
}

// Work-group reductions (blockDim.x == 256).
template <typename T, typename Op>
__device__ __forceinline__ T wg_reduce(T v, T* sh, Op op) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = op(sh[t], sh[t + s]);
        __syncthreads();
    }
    const T r = sh[0];
    __syncthreads();
    return r;
}
struct OpAdd { __device__ double operator()(double a, double b) const { return a + b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return a < b ? b : a; } };
struct OpMin { __device__ double operator()(double a, double b) const { return b < a ? b : a; } };
struct OpAddI { __device__ int operator()(int a, int b) const { return a + b; } };
struct OpOrI { __device__ int operator()(int a, int b) const { return a | b; } };

// Exclusive scan over blockDim.x ints in shared memory; returns the total.
// Exclusive scan over the work-group with wavefront shuffles: two barriers. blockDim.x
// must be a multiple of 64 and at most 64 * 64; sh holds one int per wavefront.
__device__ __forceinline__ int wg_scan_shfl(int v, int* sh, int* out_excl) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) sh[wave] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int k = 0; k < nw; ++k) {
        const int c = sh[k];
        if (k < wave) before += c;
        total += c;
    }
    __syncthreads();
    *out_excl = before + incl - v;
    return total;
}

// Exclusive scan over the work-group (any multiple of 64 threads); sh: at least
// blockDim.x / 64 ints of LDS. Returns the total.
__device__ __forceinline__ int wg_exclusive_scan(int v, int* sh, int* out_excl) {
    return wg_scan_shfl(v, sh, out_excl);
}

}  // namespace nd
