/*
 * ndipc — MI355X (gfx950) implementation of NexDynIPC's contact stage of the
 * implicit-Euler Newton step. C ABI: plain pointers and sizes, int32 status
 * (0 = ok, < 0 = error, message via ndipc_last_error), FP64 everywhere.
 *
 * Data layout of the host arrays passed in:
 *   - bodies of all scenes are concatenated; scene s owns bodies
 *     [scene_body_off[s], scene_body_off[s+1]); scene-local index = global - off;
 *   - poses: p[3n], q[4n] as (w, x, y, z), velocities v[3n], w[3n];
 *   - the Newton variable of a step is x[6n] = [p, theta] per body, theta the
 *     incremental rotation vector (nexdynipc/src/Dynamics/IPCSolver.cpp:846-857).
 *   - contacts: ints[5] = (bodyA, bodyB, type, primA, primB) with scene-local body
 *     indices, dbl[7] = (point xyz, normal xyz, distance) — the fields of
 *     Dynamics::ContactPair (nexdynipc/include/Dynamics/Forms/FrictionForm.h:13-22).
 *
 * One context per host thread; every call is ordered on the context's HIP stream.
 */
#ifndef NDIPC_H
#define NDIPC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ndipc_ctx ndipc_ctx;

/* Scene batch descriptor. Mirrors World + RigidBody + MeshShape + joints of the
 * reference (nexdynipc/include/Dynamics/World.h:11-36, RigidBody.h:11-31,
 * MeshShape.h, Joints/FixedJoint.h, Joints/HingeJoint.h) and the allowed-pair filter
 * of IPCSolver::addAllowedContactPair (IPCSolver.cpp:275-289). */
typedef struct ndipc_scene_desc {
    int32_t n_scenes;
    int32_t n_bodies;
    const int32_t* scene_body_off;   /* [n_scenes+1] */
    const int32_t* body_id;          /* [n_bodies] RigidBody::id */
    const int32_t* vert_off;         /* [n_bodies] first vertex of the body's mesh */
    const int32_t* n_verts;          /* [n_bodies] 0: body has no MeshShape */
    const int32_t* face_off;         /* [n_bodies] first face */
    const int32_t* n_faces;          /* [n_bodies] */
    const double* verts;             /* [total_v*3] body-frame vertices */
    const int32_t* faces;            /* [total_f*3] body-local vertex ids */
    const double* mass;              /* [n_bodies] */
    const double* inertia;           /* [n_bodies*9] body-frame inertia, row-major */
    int32_t n_joints;
    const int32_t* scene_joint_off;  /* [n_scenes+1] */
    const int32_t* joint_type;       /* [n_joints] 0 FixedJoint, 1 HingeJoint */
    const int32_t* joint_body;       /* [n_joints*2] scene-local bodies (second = -1 for fixed) */
    const double* joint_data;        /* [n_joints*16] fixed: p(3) q(w,x,y,z); hinge: anchorA anchorB axisA axisB */
    const double* joint_mu;          /* [n_joints] initial ALM stiffness (Joint.h:44 default 1e8) */
    int32_t n_allowed;
    const int32_t* scene_allowed_off;/* [n_scenes+1] */
    const int32_t* allowed_pairs;    /* [n_allowed*2] body ids */
    const double* allowed_mu;        /* [n_allowed] friction override, < 0: none */
} ndipc_scene_desc;

/* IPCSolver settings (nexdynipc/include/Dynamics/IPCSolver.h:38-79,94-124). */
typedef struct ndipc_params {
    double dhat;                 /* setContactDistanceThreshold */
    double kappa;                /* setContactStiffness, applied if > 0 */
    double friction_mu;          /* setFrictionCoefficient (per-scene values: ndipc_set_scene_mu) */
    int32_t enable_friction;
    int32_t enable_ccd;
    int32_t enable_adaptive_barrier;
    int32_t alm_max_iters;
    double alm_constraint_tol;
    double alm_dual_tol;
    double alm_contact_tol;
    double alm_hardening_trigger;
    double alm_hardening_ratio;
    double ccd_safety_factor;
    double newton_tol;           /* NewtonSolver gradient tolerance (default 1e-6; <= 0: default) */
    int32_t newton_max_iters;    /* NewtonSolver iteration cap (default 20; <= 0: default) */
    int32_t reserved;
} ndipc_params;

/* ---- context and data ---------------------------------------------------------- */
/* A context on HIP device `device` with its own stream. Returns 0, or -1 (out is NULL),
 * -2 (no such device), -3 (device not selectable), -4 (stream creation failed), -5 (the
 * kernels' LDS limits could not be raised: the device lacks the 160 KiB of LDS per
 * work-group of gfx950). */
int32_t ndipc_create(int32_t device, ndipc_ctx** out);
/* Frees the context and its device memory (a null context is a no-op). Returns 0, or
 * -1 when the device or the stream reported an error during teardown (the context is
 * freed either way). */
int32_t ndipc_destroy(ndipc_ctx* ctx);
const char* ndipc_last_error(const ndipc_ctx* ctx);
/* Stream the context launches on (hipStream_t). */
void* ndipc_stream(ndipc_ctx* ctx);

/* Replaces World::addBody/addJoint + ContactAssembler::build's per-mesh copies
 * (ContactAssembler.cpp:124-154): meshes, unique edges and local AABBs are uploaded
 * once and stay resident. A mesh may have up to 65535 vertices, unique edges and faces
 * and a scene up to 4096 bodies; larger inputs are rejected with an error. */
int32_t ndipc_upload_scenes(ndipc_ctx* ctx, const ndipc_scene_desc* desc);
/* Same settings for every scene (IPCSolver setters). */
int32_t ndipc_configure(ndipc_ctx* ctx, const ndipc_params* params);
/* Per-scene friction coefficients (setFrictionCoefficient per scene). */
int32_t ndipc_set_scene_mu(ndipc_ctx* ctx, const double* mu);
/* RigidBody state; any pointer may be NULL (left unchanged / not read back). */
int32_t ndipc_set_state(ndipc_ctx* ctx, const double* p, const double* q, const double* v, const double* w,
                        const double* a_lin, const double* a_ang);
int32_t ndipc_get_state(ndipc_ctx* ctx, double* p, double* q, double* v, double* w, double* a_lin,
                        double* a_ang);
/* AdaptiveBarrier state per scene (AdaptiveBarrier.h:22-87) and the once-per-solver
 * initialisation flag (IPCSolver.cpp:817-822). */
int32_t ndipc_get_barrier(ndipc_ctx* ctx, double* kappa, double* dhat, int32_t* initialized);
int32_t ndipc_set_barrier(ndipc_ctx* ctx, const double* kappa, const double* dhat, const int32_t* initialized);
/* Joint multipliers (6 per joint) and ALM stiffness (Joint.h:113-115). */
int32_t ndipc_get_joint_state(ndipc_ctx* ctx, double* lambda6, double* mu);
int32_t ndipc_set_joint_state(ndipc_ctx* ctx, const double* lambda6, const double* mu);
/* RigidBody::setMass for every body (RigidBody.h:17-31); the body-frame inertia is
 * left as uploaded. */
int32_t ndipc_set_mass(ndipc_ctx* ctx, const double* mass);

/* ---- whole step ---------------------------------------------------------------- */
/* IPCSolver::step(World&, dt) for every scene (IPCSolver.cpp:808-1231), n_steps
 * times; dt[s] per scene. Runs entirely on the device; no host synchronisation. */
int32_t ndipc_step(ndipc_ctx* ctx, const double* dt, int32_t n_steps);
/* Per-scene statistics of the last step, 8 doubles: alm iterations, newton
 * iterations, contacts after the last narrow phase, FV candidates, EE candidates,
 * narrow-phase candidate evaluations, kappa doubled (0/1), earliest TOI. */
int32_t ndipc_last_stats(ndipc_ctx* ctx, double* out8);
/* Per-scene residuals of the last step's last ALM iteration, 3 doubles: the largest
 * joint constraint violation, the dual residual and the contact residual
 * (IPCSolver::lastMaxConstraintViolation / lastDualResidual / lastContactResidual,
 * IPCSolver.h:74-79, set at IPCSolver.cpp:1156-1158). */
int32_t ndipc_last_residuals(ndipc_ctx* ctx, double* out3);
/* Narrow-phase candidate evaluations (FV + EE, VV when the fallback runs) per scene
 * since upload: the "contact pairs" of the throughput metric. */
int32_t ndipc_eval_counter(ndipc_ctx* ctx, double* out);
/* Close body pairs (over all scenes) of the last candidate build. */
int32_t ndipc_pair_count(ndipc_ctx* ctx, int32_t* out);

/* ---- function-level entry points (current state) ------------------------------ */
/* ContactAssembler::detectBodyPairs (ContactAssembler.cpp:156-178) via a device LBVH.
 * Pairs (i<j, scene-local) of scene `scene` in the reference's order: i ascending,
 * then j descending (the order of BVH::query's stack walk). */
int32_t ndipc_detect_body_pairs(ndipc_ctx* ctx, int32_t scene, double inflation, int32_t* out_pairs,
                                int32_t cap, int32_t* n_pairs);
/* ContactAssembler::estimateMinBodyPairDistance (ContactAssembler.cpp:180-212). */
int32_t ndipc_min_body_pair_distance(ndipc_ctx* ctx, int32_t scene, double inflation, double* out);
/* ContactAssembler::buildCandidateSet (ContactAssembler.cpp:214-343): the candidate
 * tuples of scene `scene` in the reference's struct layouts (VV 4, EV 5, EE 6, FV 4
 * ints, body ids and body-local primitive ids). counts[4] = vv, ev, ee, fv. Same
 * tuples in the same order as the reference built with libstdc++. */
int32_t ndipc_build_candidates(ndipc_ctx* ctx, int32_t scene, double dhat, int32_t* counts, int32_t* vv,
                               int32_t* ev, int32_t* ee, int32_t* fv, int32_t cap);
/* select_top_k_indices (ContactAssembler.cpp:67-88) on given scores: the indices kept
 * by std::nth_element, in the order std::sort leaves them (libstdc++'s introselect and
 * introsort, ties included). n <= 65535 and min(n, k) <= 2048. */
int32_t ndipc_select_top_k(ndipc_ctx* ctx, const double* scores, int32_t n, int32_t k, int32_t* out);
/* IPCSolver::detectContacts (x == NULL, IPCSolver.cpp:394-580) and
 * detectContactsAtState (x = Newton variable of scene, IPCSolver.cpp:582-806) on the
 * candidates of the last ndipc_build_candidates of that scene. */
int32_t ndipc_narrow_phase(ndipc_ctx* ctx, int32_t scene, const double* x, int32_t cap, int32_t* n,
                           int32_t* ints, double* dbl);
/* computeSceneMinDistance(...) < 1e-9 (AdaptiveBarrier.h:130-162, IPCSolver.cpp:826-830):
 * the only use of the scene minimum distance; exact pruning by body AABB gap. */
int32_t ndipc_scene_min_distance_flag(ndipc_ctx* ctx, int32_t scene, int32_t* out_flag);
/* DistanceBarrierForm::value/gradient/hessian/computeNormalForces
 * (DistanceBarrierForm.cpp:11-87) at x for given contacts. hess: dense 6n x 6n
 * row-major (may be NULL). */
int32_t ndipc_barrier_eval(ndipc_ctx* ctx, int32_t scene, int32_t nc, const int32_t* ints, const double* dbl,
                           const double* x, double dhat, double kappa, double* value, double* grad, double* hess,
                           double* normal_forces);
/* FrictionForm::value/gradient/hessian (FrictionForm.cpp:22-151) with lagged normal
 * forces. */
int32_t ndipc_friction_eval(ndipc_ctx* ctx, int32_t scene, int32_t nc, const int32_t* ints, const double* dbl,
                            const double* normal_forces, const double* x, double mu, double eps, double dt,
                            double* value, double* grad, double* hess);
/* ipc mode: the primitive-distance barrier of rigid IPC that replaces the centre-
 * distance barrier in the reference's plan (documents/2026.02.28/
 * NexDynIPC_接触链路rigid-ipc完整替换实施规划.md:37-81; the reference has no code for it
 * yet). np pairs of scene-local bodies, 5 int32 each: kind 0 point-triangle (body of the
 * vertex, body of the face, vertex, face) or 1 edge-edge (body and unique-edge index of
 * edge a, then of edge b); x is the scene's Newton variable (6 per body: p, theta about
 * the stored orientation). Per pair E = kappa b(d) (b: Barrier.cpp:11-46) of the
 * primitive distance d, with its gradient and Hessian (projected to PSD when psd != 0)
 * folded into the two bodies' DOFs (Gauss-Newton in the rigid map). Consecutive pairs of
 * one unordered body pair form a segment: per segment seg_bodies[2] (lower body first),
 * seg_g[12] and seg_H[144] (row-major, lower body's 6 DOFs first); value = total energy.
 * n_seg receives the segment count; at most cap segments are written; n_active (may be
 * NULL) the number of pairs within dhat. Fails (naming the first bad pair) on a bad
 * kind, body or primitive index. np <= NDIPC_IPC_MAX_PAIRS. */
#define NDIPC_IPC_MAX_PAIRS (1 << 22)
int32_t ndipc_ipc_barrier_eval(ndipc_ctx* ctx, int32_t scene, int32_t np, const int32_t* pairs5, const double* x,
                               double dhat, double kappa, int32_t psd, double* value, int32_t cap, int32_t* n_seg,
                               int32_t* seg_bodies, double* seg_g, double* seg_H, int32_t* n_active);
/* ipc mode, OptimizationProblem::maxStep (NewtonSolver.h:42, called at
 * NewtonSolver.cpp:53-54) over the pairs of ndipc_ipc_barrier_eval's format: per pair,
 * conservative advancement along the line-search path x + t dx (positions p + t dp,
 * orientations apply_theta(q, theta + t dtheta)), t in [0, 1], keeping sep (in [0, 1))
 * of the pair's current distance: the pair's distance provably stays above sep * d(0)
 * up to its advance (at most 1000 steps per pair; a pair that would need more reports
 * where it stopped, which is still safe). toi = the least advance (+inf if every pair
 * clears the whole step); alpha = toi * safety when toi < 1 (the form of
 * IPCSolver::computeMaxStep, IPCSolver.cpp:385-391), else 1. Fails naming the first
 * bad pair. */
int32_t ndipc_ipc_max_step(ndipc_ctx* ctx, int32_t scene, int32_t np, const int32_t* pairs5, const double* x,
                           const double* dx, double sep, double safety, double* alpha, double* toi);
/* ipc mode time step of one scene: IPCSolver::step's implicit-Euler Newton solve
 * (IPCSolver.cpp:137-262, NewtonSolver.cpp:10-95) with the ipc barrier in place of the
 * centre-distance barrier and maxStep wired into the line search, for scenes whose only
 * joints pin single bodies (those bodies are static). Objective as the reference's:
 * sum over moving bodies of 1/2 m |x_p - x^_p|^2 + 1/2 dtheta^T I dtheta, + h^2 (gravity
 * + E_ipc), x^ = (p + h v, h w). Pairs: the top-K lists of the close pairs at `inflation`
 * (faces x vertices both ways, edges x edges), fixed for the step. Each Newton iteration
 * solves H dx = -g by block-Jacobi PCG (pcg_tol, pcg_max_it), stops when |dx|_inf <=
 * newton_tol or dx does not descend, starts the line search at min(1, maxStep) (sep,
 * safety; ndipc_ipc_max_step) and halves until Armijo (1e-4) holds (max_ls trials).
 * Writes p, q, v, w of the moving bodies. stats4 (may be NULL): Newton iterations,
 * line-search trials, pair records, PCG iterations. */
typedef struct ndipc_ipc_params {
    double dhat, kappa, sep, safety, newton_tol, pcg_tol, inflation;
    int32_t max_newton, max_ls, pcg_max_it;
} ndipc_ipc_params;
int32_t ndipc_ipc_step(ndipc_ctx* ctx, int32_t scene, double dt, const ndipc_ipc_params* params, int32_t* stats4);
/* Interval TOI batch (TimeOfImpact.cpp:8-150): kind 0 edge-vertex, 1 edge-edge,
 * 2 face-vertex as the reference computes them; the ipc mode's 3 face-vertex (boxes that
 * meet the triangle u, v >= 0, u + v <= 1 are kept, where the reference's predicate
 * drops the root box) and 4 edge-edge, both in body-relative form (relative translation
 * linear in t over each box, rotated vectors as slerp-endpoint hulls widened by the
 * arc's distance from its chord), searched earliest first; when the 1000-pop cap stops
 * them they report an impact at the least t.lo still unexplored (a lower bound). Per query:
 * poseA[14], poseB[14] = p0(3) q0(4) p1(3) q1(4); a_pts[6] (vertex or edge), b_pts[9]
 * (edge or face). out: hit (1/0, -1 on NaN, -2 on stack overflow), toi, pops. */
int32_t ndipc_ccd_toi_batch(ndipc_ctx* ctx, int32_t kind, int32_t n, const double* poseA, const double* poseB,
                            const double* a_pts, const double* b_pts, const double* earliest, double tol,
                            int32_t* hit, double* toi, int32_t* pops);
/* CCDSystem::computeEarliestTOI over the FV candidates of the last candidate build
 * (CCD.h:22-82), rotation frozen, t1 = p + v dt. */
int32_t ndipc_ccd_earliest_toi(ndipc_ctx* ctx, int32_t scene, double dt, double* out);
/* Block-Jacobi PCG on a block-sparse SPD system: n_blocks 6x6 diagonal blocks,
 * n_off off-diagonal blocks (row block, col block) stored once for the upper
 * triangle. Solves A x = b; returns iterations and final relative residual. */
int32_t ndipc_pcg_solve(ndipc_ctx* ctx, int32_t n_blocks, const double* diag, int32_t n_off, const int32_t* off_ij,
                        const double* off_blocks, const double* b, double tol, int32_t max_it, double* x,
                        int32_t* iters, double* rel_res);
/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops), the same block-Jacobi PCG as ndipc_pcg_solve
 * with every array in device memory (d_off_ij: 2 per off-diagonal block; d_iters: one
 * int32, d_rel: one double) and the solver's work vectors in a fixed array of the
 * library, one per GPU, sized for NDIPC_PCG_NOALLOC_MAX_BLOCKS blocks (1.9 MB). Larger
 * systems are rejected. Enqueued on the context stream (ndipc_stream); returns without
 * synchronising. Calls that overlap on one GPU share that array and race. */
#define NDIPC_PCG_NOALLOC_MAX_BLOCKS 4096
int32_t ndipc_pcg_solve_noalloc(ndipc_ctx* ctx, int32_t n_blocks, const double* d_diag, int32_t n_off,
                                const int32_t* d_off_ij, const double* d_offb, const double* d_b, double tol,
                                int32_t max_it, double* d_x, int32_t* d_iters, double* d_rel);

/* ---- adjoint (no reference counterpart: the reference has no differentiation) ------ */
/* n_steps of ndipc_step, saving the full state (poses, velocities, accelerations, joint
 * multipliers and stiffnesses, barrier state) every checkpoint_every steps for the
 * reverse sweep. Leaves the scenes at the end state like ndipc_step. */
int32_t ndipc_adjoint_forward(ndipc_ctx* ctx, const double* dt, int32_t n_steps, int32_t checkpoint_every);
/* Reverse sweep of the last ndipc_adjoint_forward. In: the cotangent of the end state
 * (dL/dp[3n], dL/dq[4n], dL/dv[3n], dL/dw[3n]; NULL = 0). Out (NULL = not wanted): dL/d
 * of the start state p, q, v, w, of every body mass, and of every scene's friction
 * coefficient. Each step is differentiated through its ALM solves, with the
 * implicit-function theorem at every solve's Newton point (the exact Jacobian of the
 * Newton gradient solved on the device) and through the multiplier updates; a solve whose
 * Newton loop accepted no step passes its cotangent to its starting point. Held fixed:
 * the contact set with each contact's normal and point, the lagged normal forces, the
 * barrier stiffness and the ALM hardening. Segments between checkpoints are recomputed
 * forward (bitwise the recorded run) before they are swept back. Fails (and names the
 * scene, step and ALM iteration) if a transposed solve does not reach a componentwise
 * backward error of 1e-12. No reference counterpart: the reference has no adjoint. */
int32_t ndipc_adjoint_backward(ndipc_ctx* ctx, const double* dl_dp, const double* dl_dq, const double* dl_dv,
                               const double* dl_dw, double* g_p0, double* g_q0, double* g_v0, double* g_w0,
                               double* g_mass, double* g_mu);

/* ---- diagnostics ---------------------------------------------------------------- */
/* Newton work, 8 per scene: of the last step Newton iterations, line-search trials,
 * direct (component LDL^T) solves, PCG solves, PCG iterations, largest coupled
 * component in bodies; since upload Newton iterations and line-search trials.
 * No reference counterpart (NewtonSolver.cpp:10-95 only logs). */
int32_t ndipc_solver_counters(ndipc_ctx* ctx, double* out8);
/* Line-search helper work-groups (single scenes and small batches), 4 per scene:
 * whether the scenes have helpers, and since upload the batches they answered, the
 * waits a scene gave up (it then did the work itself) and the give-ups that found a
 * helper still busy after the straggle bound. No reference counterpart. */
int32_t ndipc_helper_counters(ndipc_ctx* ctx, int32_t* out4);
/* Wait bounds of the helper protocol in device clock ticks (100 MHz): a scene's wait
 * for an answer, a helper's wait for the next Newton launch and for the next batch, and
 * a scene's wait for a helper already busy with a batch it gave up; a negative value
 * restores the default (5 ms, 5 ms, 200 ms, 100 ms). Taken at the next step's begin.
 * Results never depend on them. No reference counterpart. */
int32_t ndipc_helper_config(ndipc_ctx* ctx, int64_t wait_ticks, int64_t epoch_ticks, int64_t idle_ticks,
                            int64_t straggle_ticks);
/* Average device duration (ms) and count of launches of the named kernel since the
 * last reset, measured with HIP events on the context stream while enabled (steps then
 * enqueue their kernels one by one instead of replaying the step graph). */
int32_t ndipc_timing_enable(ndipc_ctx* ctx, int32_t on);
/* ndipc_step replays captured HIP graphs (default, on = 1: the step's prefix and one ALM
 * round in adaptive mode, the whole step in fixed-round mode) or enqueues the step's
 * kernels one by one (on = 0); both run the same kernels with the same arguments. No
 * reference counterpart (host launch strategy). */
int32_t ndipc_use_graph(ndipc_ctx* ctx, int32_t on);
/* Step modes. adaptive = 1: each step issues ALM rounds only until every scene's ALM
 * loop has ended (a round reports to pinned host memory), at most `lookahead` rounds
 * ahead of the last one reported. adaptive = 2 (default): as 1, and a call of several
 * steps on a batch of several scenes steps the scenes asynchronously: a scene whose ALM
 * loop has ended writes its step back and starts its next step in the next round instead
 * of waiting for the slowest scene (a batch without line-search helpers; otherwise as 1).
 * adaptive = 0 enqueues all alm_max_iters rounds every step (scenes that finished skip
 * them on the device). Same results in every mode. The adjoint's recorded forward always
 * runs the fixed rounds. Environment at create: NDIPC_FIXED_ROUNDS=1,
 * NDIPC_ROUND_LOOKAHEAD=n, NDIPC_ASYNC_STEPS=0 (mode 1). No reference counterpart
 * (IPCSolver.cpp:965-1195 loops on the host, one scene at a time). */
int32_t ndipc_set_step_mode(ndipc_ctx* ctx, int32_t adaptive, int32_t lookahead);
/* Autotuner of the launch shapes (work-group counts of the persistent top-K, narrow-phase,
 * scene-min-distance and CCD kernels) for the uploaded batch: from the current state,
 * each candidate runs trial_steps steps of dt (timed with HIP events), one knob at a
 * time; the state is restored and the fastest shapes are kept (out4: the chosen counts,
 * out_ms: ms per step with them). Results of a step do not depend on the shapes. No
 * reference counterpart. */
int32_t ndipc_autotune(ndipc_ctx* ctx, const double* dt, int32_t trial_steps, int32_t* out4, double* out_ms);
/* Sets the same four launch shapes directly (grids4: top-K, narrow phase, scene-min flag,
 * CCD work-group counts; 0 keeps the default), e.g. shapes tuned on another context at
 * a later state of the same workload. Kept until the next upload. No reference
 * counterpart. */
int32_t ndipc_set_launch_grids(ndipc_ctx* ctx, const int32_t* grids4);
int32_t ndipc_timing_get(ndipc_ctx* ctx, const char* kernel, double* avg_ms, int32_t* count);
/* Directed-rounding primitives of the interval CCD, evaluated on the device:
 * out6 = add_dn, add_up, sub_dn, sub_up, mul_dn, mul_up. */
int32_t ndipc_round_ops(ndipc_ctx* ctx, const double* a, const double* b, int32_t n, double* out6);
/* Elementary functions of the path evaluated on the device, element-wise: fn 0 log(x)
 * (std::log in Barrier.cpp:11-46), 1 sin(x), 2 cos(x) (Eigen AngleAxis and slerp used by
 * IPCSolver.cpp:616-626 and the CCD trajectories), 3 atan2(x, y) (Eigen
 * quaternion -> AngleAxis, FixedJoint.cpp), 4 acos(x) (Eigen slerp). y may be NULL
 * unless fn == 3. */
int32_t ndipc_elemfn(ndipc_ctx* ctx, int32_t fn, const double* x, const double* y, int32_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif
