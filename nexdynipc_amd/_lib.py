"""ctypes binding of the ndipc C ABI (include/ndipc.h) and a batch engine on top.

The library is nexdynipc_amd/_build/libndipc.so, built by nexdynipc_amd/build.py.
There is no CPU fallback: every method runs the HIP kernels or raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libndipc.so")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class SceneDesc(C.Structure):
    """ndipc_scene_desc (include/ndipc.h)."""
    _fields_ = [
        ("n_scenes", C.c_int32), ("n_bodies", C.c_int32),
        ("scene_body_off", i32p), ("body_id", i32p), ("vert_off", i32p), ("n_verts", i32p),
        ("face_off", i32p), ("n_faces", i32p), ("verts", f64p), ("faces", i32p),
        ("mass", f64p), ("inertia", f64p),
        ("n_joints", C.c_int32), ("scene_joint_off", i32p), ("joint_type", i32p), ("joint_body", i32p),
        ("joint_data", f64p), ("joint_mu", f64p),
        ("n_allowed", C.c_int32), ("scene_allowed_off", i32p), ("allowed_pairs", i32p), ("allowed_mu", f64p),
    ]


class Params(C.Structure):
    """ndipc_params (include/ndipc.h)."""
    _fields_ = [
        ("dhat", C.c_double), ("kappa", C.c_double), ("friction_mu", C.c_double),
        ("enable_friction", C.c_int32), ("enable_ccd", C.c_int32), ("enable_adaptive_barrier", C.c_int32),
        ("alm_max_iters", C.c_int32), ("alm_constraint_tol", C.c_double), ("alm_dual_tol", C.c_double),
        ("alm_contact_tol", C.c_double), ("alm_hardening_trigger", C.c_double),
        ("alm_hardening_ratio", C.c_double), ("ccd_safety_factor", C.c_double),
        ("newton_tol", C.c_double), ("newton_max_iters", C.c_int32), ("reserved", C.c_int32),
    ]


class IpcParams(C.Structure):
    _fields_ = [("dhat", C.c_double), ("kappa", C.c_double), ("sep", C.c_double), ("safety", C.c_double),
                ("newton_tol", C.c_double), ("pcg_tol", C.c_double), ("inflation", C.c_double),
                ("max_newton", C.c_int32), ("max_ls", C.c_int32), ("pcg_max_it", C.c_int32)]


EXPORTS = [
    "ndipc_create", "ndipc_destroy", "ndipc_last_error", "ndipc_stream", "ndipc_upload_scenes", "ndipc_configure",
    "ndipc_set_scene_mu", "ndipc_set_state", "ndipc_get_state", "ndipc_get_barrier", "ndipc_set_barrier",
    "ndipc_get_joint_state", "ndipc_step", "ndipc_last_stats", "ndipc_detect_body_pairs",
    "ndipc_min_body_pair_distance", "ndipc_build_candidates", "ndipc_select_top_k", "ndipc_narrow_phase",
    "ndipc_scene_min_distance_flag", "ndipc_barrier_eval", "ndipc_friction_eval", "ndipc_ccd_toi_batch",
    "ndipc_ccd_earliest_toi", "ndipc_pcg_solve", "ndipc_timing_enable", "ndipc_timing_get", "ndipc_round_ops", "ndipc_elemfn", "ndipc_solver_counters", "ndipc_set_joint_state",
    "ndipc_set_mass", "ndipc_adjoint_forward", "ndipc_adjoint_backward", "ndipc_use_graph",
    "ndipc_autotune", "ndipc_set_launch_grids", "ndipc_helper_counters", "ndipc_helper_config",
    "ndipc_set_step_mode", "ndipc_pcg_solve_noalloc", "ndipc_last_residuals", "ndipc_ipc_barrier_eval",
    "ndipc_ipc_max_step", "ndipc_ipc_step",
]


def ptr(a, t):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


_lib = None


def library():
    """Loads libndipc.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run nexdynipc_amd/build.py (hipcc, gfx950) first")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "ndipc_create": ([C.c_int32, C.POINTER(vp)], C.c_int32),
        "ndipc_destroy": ([vp], C.c_int32),
        "ndipc_last_error": ([vp], C.c_char_p),
        "ndipc_stream": ([vp], vp),
        "ndipc_upload_scenes": ([vp, C.POINTER(SceneDesc)], C.c_int32),
        "ndipc_configure": ([vp, C.POINTER(Params)], C.c_int32),
        "ndipc_set_scene_mu": ([vp, f64p], C.c_int32),
        "ndipc_set_state": ([vp] + [f64p] * 6, C.c_int32),
        "ndipc_get_state": ([vp] + [f64p] * 6, C.c_int32),
        "ndipc_get_barrier": ([vp, f64p, f64p, i32p], C.c_int32),
        "ndipc_set_barrier": ([vp, f64p, f64p, i32p], C.c_int32),
        "ndipc_get_joint_state": ([vp, f64p, f64p], C.c_int32),
        "ndipc_step": ([vp, f64p, C.c_int32], C.c_int32),
        "ndipc_last_stats": ([vp, f64p], C.c_int32),
        "ndipc_last_residuals": ([vp, f64p], C.c_int32),
        "ndipc_eval_counter": ([vp, f64p], C.c_int32),
        "ndipc_pair_count": ([vp, i32p], C.c_int32),
        "ndipc_detect_body_pairs": ([vp, C.c_int32, C.c_double, i32p, C.c_int32, i32p], C.c_int32),
        "ndipc_min_body_pair_distance": ([vp, C.c_int32, C.c_double, f64p], C.c_int32),
        "ndipc_build_candidates": ([vp, C.c_int32, C.c_double, i32p, i32p, i32p, i32p, i32p, C.c_int32], C.c_int32),
        "ndipc_select_top_k": ([vp, f64p, C.c_int32, C.c_int32, i32p], C.c_int32),
        "ndipc_narrow_phase": ([vp, C.c_int32, f64p, C.c_int32, i32p, i32p, f64p], C.c_int32),
        "ndipc_scene_min_distance_flag": ([vp, C.c_int32, i32p], C.c_int32),
        "ndipc_barrier_eval": ([vp, C.c_int32, C.c_int32, i32p, f64p, f64p, C.c_double, C.c_double, f64p, f64p,
                                f64p, f64p], C.c_int32),
        "ndipc_ipc_barrier_eval": ([vp, C.c_int32, C.c_int32, i32p, f64p, C.c_double, C.c_double, C.c_int32, f64p,
                                    C.c_int32, i32p, i32p, f64p, f64p, i32p], C.c_int32),
        "ndipc_ipc_max_step": ([vp, C.c_int32, C.c_int32, i32p, f64p, f64p, C.c_double, C.c_double, f64p, f64p],
                               C.c_int32),
        "ndipc_ipc_step": ([vp, C.c_int32, C.c_double, C.POINTER(IpcParams), i32p], C.c_int32),
        "ndipc_friction_eval": ([vp, C.c_int32, C.c_int32, i32p, f64p, f64p, f64p, C.c_double, C.c_double,
                                 C.c_double, f64p, f64p, f64p], C.c_int32),
        "ndipc_ccd_toi_batch": ([vp, C.c_int32, C.c_int32, f64p, f64p, f64p, f64p, f64p, C.c_double, i32p, f64p,
                                 i32p], C.c_int32),
        "ndipc_ccd_earliest_toi": ([vp, C.c_int32, C.c_double, f64p], C.c_int32),
        "ndipc_pcg_solve": ([vp, C.c_int32, f64p, C.c_int32, i32p, f64p, f64p, C.c_double, C.c_int32, f64p, i32p,
                             f64p], C.c_int32),
        "ndipc_pcg_solve_noalloc": ([vp, C.c_int32, vp, C.c_int32, vp, vp, vp, C.c_double, C.c_int32, vp, vp, vp],
                                    C.c_int32),
        "ndipc_timing_enable": ([vp, C.c_int32], C.c_int32),
        "ndipc_use_graph": ([vp, C.c_int32], C.c_int32),
        "ndipc_set_step_mode": ([vp, C.c_int32, C.c_int32], C.c_int32),
        "ndipc_autotune": ([vp, f64p, C.c_int32, i32p, f64p], C.c_int32),
        "ndipc_set_launch_grids": ([vp, i32p], C.c_int32),
        "ndipc_timing_get": ([vp, C.c_char_p, f64p, i32p], C.c_int32),
        "ndipc_round_ops": ([vp, f64p, f64p, C.c_int32, f64p], C.c_int32),
        "ndipc_elemfn": ([vp, C.c_int32, f64p, f64p, C.c_int32, f64p], C.c_int32),
        "ndipc_solver_counters": ([vp, f64p], C.c_int32),
        "ndipc_helper_counters": ([vp, i32p], C.c_int32),
        "ndipc_helper_config": ([vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64], C.c_int32),
        "ndipc_set_joint_state": ([vp, f64p, f64p], C.c_int32),
        "ndipc_set_mass": ([vp, f64p], C.c_int32),
        "ndipc_adjoint_forward": ([vp, f64p, C.c_int32, C.c_int32], C.c_int32),
        "ndipc_adjoint_backward": ([vp] + [f64p] * 10, C.c_int32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def make_desc(sc):
    d = SceneDesc()
    d.n_scenes = sc.n_scenes
    d.n_bodies = sc.n_bodies
    for f in ("scene_body_off", "body_id", "vert_off", "n_verts", "face_off", "n_faces", "faces",
              "scene_joint_off", "joint_type", "joint_body", "scene_allowed_off", "allowed_pairs"):
        setattr(d, f, ptr(getattr(sc, f), i32p))
    for f in ("verts", "mass", "inertia", "joint_data", "joint_mu", "allowed_mu"):
        setattr(d, f, ptr(getattr(sc, f), f64p))
    d.n_joints = sc.n_joints
    d.n_allowed = int(len(sc.allowed_mu))
    return d


def make_params(dhat=1e-3, mu=0.0, friction=True, ccd=True, adaptive=True, alm_max_iters=10, kappa=0.0,
                newton_tol=1e-6, newton_max_iters=20):
    p = Params()
    p.dhat = dhat
    p.kappa = kappa
    p.friction_mu = mu
    p.enable_friction = 1 if friction else 0
    p.enable_ccd = 1 if ccd else 0
    p.enable_adaptive_barrier = 1 if adaptive else 0
    p.alm_max_iters = alm_max_iters
    p.alm_constraint_tol = 1e-4
    p.alm_dual_tol = 1e-3
    p.alm_contact_tol = 1e-3
    p.alm_hardening_trigger = 2.0
    p.alm_hardening_ratio = 2.0
    p.ccd_safety_factor = 0.8
    p.newton_tol = newton_tol
    p.newton_max_iters = newton_max_iters
    return p


ELEMFN = ("log", "sin", "cos", "atan2", "acos")


class NdipcError(RuntimeError):
    pass


class Engine:
    """A device context holding a batch of scenes (all state resident in HBM)."""

    def __init__(self, device=0):
        self.L = library()
        self.ctx = C.c_void_p()
        rc = self.L.ndipc_create(device, C.byref(self.ctx))
        if rc != 0:
            raise NdipcError(f"ndipc_create failed ({rc}): no usable HIP device")
        self.sc = None

    def close(self):
        if self.ctx:
            self.L.ndipc_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        if rc < 0:
            raise NdipcError(f"{what}: {self.L.ndipc_last_error(self.ctx).decode()}")
        return rc

    def stream(self):
        return self.L.ndipc_stream(self.ctx)

    def upload(self, sc, params=None):
        self.sc = sc
        self._desc = make_desc(sc)
        self._ck(self.L.ndipc_upload_scenes(self.ctx, C.byref(self._desc)), "upload")
        if params is None:
            params = make_params(dhat=float(sc.dhat[0]), mu=float(sc.mu[0]), friction=bool(sc.enable_friction[0]))
        self._ck(self.L.ndipc_configure(self.ctx, C.byref(params)), "configure")
        mu = np.ascontiguousarray(sc.mu, dtype=np.float64)
        self._ck(self.L.ndipc_set_scene_mu(self.ctx, ptr(mu, f64p)), "set_scene_mu")
        n = sc.n_bodies
        self.set_state(sc.p, sc.q, sc.v, sc.w, np.zeros((n, 3)), np.zeros((n, 3)))

    def configure(self, params):
        """ndipc_configure: the same IPCSolver settings for every scene."""
        self._ck(self.L.ndipc_configure(self.ctx, C.byref(params)), "configure")

    def set_state(self, p=None, q=None, v=None, w=None, al=None, aa=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (p, q, v, w, al, aa)]
        self._ck(self.L.ndipc_set_state(self.ctx, *[ptr(a, f64p) for a in arrs]), "set_state")

    def get_state(self):
        n = self.sc.n_bodies
        out = [np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)),
               np.zeros((n, 3))]
        self._ck(self.L.ndipc_get_state(self.ctx, *[ptr(a, f64p) for a in out]), "get_state")
        return out

    def barrier(self):
        S = self.sc.n_scenes
        k, d, i = np.zeros(S), np.zeros(S), np.zeros(S, dtype=np.int32)
        self._ck(self.L.ndipc_get_barrier(self.ctx, ptr(k, f64p), ptr(d, f64p), ptr(i, i32p)), "get_barrier")
        return k, d, i

    def set_barrier(self, kappa, dhat, init):
        k = np.ascontiguousarray(kappa, dtype=np.float64)
        d = np.ascontiguousarray(dhat, dtype=np.float64)
        i = np.ascontiguousarray(init, dtype=np.int32)
        self._ck(self.L.ndipc_set_barrier(self.ctx, ptr(k, f64p), ptr(d, f64p), ptr(i, i32p)), "set_barrier")

    def joint_state(self):
        J = self.sc.n_joints
        lam, mu = np.zeros(6 * max(J, 1)), np.zeros(max(J, 1))
        self._ck(self.L.ndipc_get_joint_state(self.ctx, ptr(lam, f64p), ptr(mu, f64p)), "get_joint_state")
        return lam[: 6 * J].reshape(-1, 6), mu[:J]

    def set_joint_state(self, lam=None, mu=None):
        lam = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64)
        mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        self._ck(self.L.ndipc_set_joint_state(self.ctx, ptr(lam, f64p), ptr(mu, f64p)), "set_joint_state")

    def set_mass(self, mass):
        m = np.ascontiguousarray(mass, dtype=np.float64)
        assert m.shape == (self.sc.n_bodies,)
        self._ck(self.L.ndipc_set_mass(self.ctx, ptr(m, f64p)), "set_mass")

    def set_mu(self, mu):
        mu = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (self.sc.n_scenes,)))
        self._ck(self.L.ndipc_set_scene_mu(self.ctx, ptr(mu, f64p)), "set_scene_mu")

    def adjoint_forward(self, n_steps, every=10, dt=None):
        dt = self.sc.dt if dt is None else np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.sc.n_scenes,))
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        self._ck(self.L.ndipc_adjoint_forward(self.ctx, ptr(dt, f64p), int(n_steps), int(every)), "adjoint_forward")

    def adjoint_backward(self, dl_dp=None, dl_dq=None, dl_dv=None, dl_dw=None):
        """Gradients of a loss of the end state with respect to the start state, the
        masses and the per-scene friction coefficients (dict of arrays)."""
        n, S = self.sc.n_bodies, self.sc.n_scenes
        ins = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (dl_dp, dl_dq, dl_dv, dl_dw)]
        out = {"p0": np.zeros((n, 3)), "q0": np.zeros((n, 4)), "v0": np.zeros((n, 3)), "w0": np.zeros((n, 3)),
               "mass": np.zeros(n), "mu": np.zeros(S)}
        args = [ptr(a, f64p) for a in ins] + [ptr(out[k], f64p) for k in ("p0", "q0", "v0", "w0", "mass", "mu")]
        self._ck(self.L.ndipc_adjoint_backward(self.ctx, *args), "adjoint_backward")
        return out

    def step(self, n_steps=1, dt=None):
        dt = self.sc.dt if dt is None else np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.sc.n_scenes,))
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        self._ck(self.L.ndipc_step(self.ctx, ptr(dt, f64p), int(n_steps)), "step")

    def stats(self):
        o = np.zeros((self.sc.n_scenes, 8))
        self._ck(self.L.ndipc_last_stats(self.ctx, ptr(o, f64p)), "last_stats")
        return o

    def last_residuals(self):
        """Per scene: joint violation, dual residual, contact residual of the last step's
        last ALM iteration."""
        o = np.zeros((self.sc.n_scenes, 3))
        self._ck(self.L.ndipc_last_residuals(self.ctx, ptr(o, f64p)), "last_residuals")
        return o

    def solver_counters(self):
        """Per scene: newton iterations, line-search trials, direct solves, PCG solves,
        PCG iterations, largest coupled component (bodies) of the last step; Newton
        iterations and line-search trials since upload."""
        o = np.zeros((self.sc.n_scenes, 8))
        self._ck(self.L.ndipc_solver_counters(self.ctx, ptr(o, f64p)), "solver_counters")
        return o

    def set_step_mode(self, adaptive=True, lookahead=1, asynchronous=True):
        """Adaptive ALM rounds (issue only the rounds the scenes need) or fixed rounds;
        with asynchronous, a multi-step call on a batch lets each scene start its next
        step as soon as its ALM loop ends."""
        mode = (2 if asynchronous else 1) if adaptive else 0
        self._ck(self.L.ndipc_set_step_mode(self.ctx, mode, int(lookahead)), "set_step_mode")

    def helper_counters(self):
        """Per scene: helpers on (0/1), and since upload the batches the helpers answered,
        the waits the scene gave up, and give-ups that outlasted the straggle bound."""
        o = np.zeros((self.sc.n_scenes, 4), dtype=np.int32)
        self._ck(self.L.ndipc_helper_counters(self.ctx, ptr(o, i32p)), "helper_counters")
        return o

    def helper_config(self, wait_ticks=-1, epoch_ticks=-1, idle_ticks=-1, straggle_ticks=-1):
        """Wait bounds of the helper protocol in 100 MHz ticks (-1: default)."""
        self._ck(self.L.ndipc_helper_config(self.ctx, int(wait_ticks), int(epoch_ticks), int(idle_ticks),
                                            int(straggle_ticks)), "helper_config")

    def eval_counter(self):
        o = np.zeros(self.sc.n_scenes)
        self._ck(self.L.ndipc_eval_counter(self.ctx, ptr(o, f64p)), "eval_counter")
        return o

    def pair_count(self):
        o = C.c_int32()
        self._ck(self.L.ndipc_pair_count(self.ctx, C.byref(o)), "pair_count")
        return o.value

    # ---- function level (scene-local indices, the oracle's output formats)
    def body_pairs(self, scene, r):
        n = int(self.sc.scene_body_off[scene + 1] - self.sc.scene_body_off[scene])
        cap = max(n * n, 1)
        out = np.zeros(2 * cap, dtype=np.int32)
        cnt = C.c_int32()
        self._ck(self.L.ndipc_detect_body_pairs(self.ctx, scene, r, ptr(out, i32p), cap, C.byref(cnt)), "pairs")
        return out[: 2 * cnt.value].reshape(-1, 2)

    def min_body_pair_distance(self, scene, r):
        o = C.c_double()
        self._ck(self.L.ndipc_min_body_pair_distance(self.ctx, scene, r, C.byref(o)), "min_pair_distance")
        return o.value

    def candidates(self, scene, dhat):
        counts = np.zeros(4, dtype=np.int32)
        self._ck(self.L.ndipc_build_candidates(self.ctx, scene, dhat, ptr(counts, i32p), None, None, None, None, 0),
                 "candidates")
        cap = max(int(counts.max()), 1)
        vv, ev = np.zeros(4 * cap, dtype=np.int32), np.zeros(5 * cap, dtype=np.int32)
        ee, fv = np.zeros(6 * cap, dtype=np.int32), np.zeros(4 * cap, dtype=np.int32)
        self._ck(self.L.ndipc_build_candidates(self.ctx, scene, dhat, ptr(counts, i32p), ptr(vv, i32p), ptr(ev, i32p),
                                               ptr(ee, i32p), ptr(fv, i32p), cap), "candidates")
        return (vv[: 4 * counts[0]].reshape(-1, 4), ev[: 5 * counts[1]].reshape(-1, 5),
                ee[: 6 * counts[2]].reshape(-1, 6), fv[: 4 * counts[3]].reshape(-1, 4))

    def select_top_k(self, scores, k):
        s = np.ascontiguousarray(scores, dtype=np.float64)
        out = np.zeros(max(k, 1), dtype=np.int32)
        m = self._ck(self.L.ndipc_select_top_k(self.ctx, ptr(s, f64p), len(s), k, ptr(out, i32p)), "select_top_k")
        return out[:m]

    def narrow_phase(self, scene, x=None, cap=40000):
        n = C.c_int32()
        ints = np.zeros(5 * cap, dtype=np.int32)
        dbl = np.zeros(7 * cap)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        self._ck(self.L.ndipc_narrow_phase(self.ctx, scene, ptr(xa, f64p), cap, C.byref(n), ptr(ints, i32p),
                                           ptr(dbl, f64p)), "narrow_phase")
        k = min(n.value, cap)
        return ints[: 5 * k].reshape(-1, 5), dbl[: 7 * k].reshape(-1, 7)

    def scene_min_distance_flag(self, scene):
        f = C.c_int32()
        self._ck(self.L.ndipc_scene_min_distance_flag(self.ctx, scene, C.byref(f)), "scene_min_flag")
        return f.value

    def barrier_eval(self, scene, ints, dbl, x, dhat, kappa, hess=True):
        nd = len(x)
        val = C.c_double()
        g = np.zeros(nd)
        H = np.zeros((nd, nd)) if hess else None
        nf = np.zeros(max(len(ints), 1))
        ints = np.ascontiguousarray(ints, dtype=np.int32)
        dbl = np.ascontiguousarray(dbl, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._ck(self.L.ndipc_barrier_eval(self.ctx, scene, len(ints), ptr(ints, i32p), ptr(dbl, f64p), ptr(x, f64p),
                                           dhat, kappa, C.byref(val), ptr(g, f64p), ptr(H, f64p), ptr(nf, f64p)),
                 "barrier_eval")
        return val.value, g, H, nf[: len(ints)]

    def ipc_barrier_eval(self, scene, pairs, x, dhat, kappa, psd=True, cap=None, with_active=False):
        """ipc mode over pairs (n x 5 int32) of scene at Newton variable x: total energy and
        per segment (lower body, higher body), gradient [12] and Hessian [12, 12] (at most
        cap segments returned; default: one per pair); with_active: also the number of
        pairs within dhat."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 5)
        x = np.ascontiguousarray(x, dtype=np.float64)
        cap = max(len(pr) if cap is None else int(cap), 1)
        val = C.c_double()
        ns = C.c_int32()
        sb = np.zeros(2 * cap, dtype=np.int32)
        sg = np.zeros(12 * cap)
        sH = np.zeros(144 * cap)
        na = C.c_int32()
        self._ck(self.L.ndipc_ipc_barrier_eval(self.ctx, scene, len(pr), ptr(pr, i32p), ptr(x, f64p), dhat, kappa,
                                               1 if psd else 0, C.byref(val), cap, C.byref(ns), ptr(sb, i32p),
                                               ptr(sg, f64p), ptr(sH, f64p), C.byref(na)), "ipc_barrier_eval")
        k = min(ns.value, cap)
        out = (val.value, sb[: 2 * k].reshape(-1, 2), sg[: 12 * k].reshape(-1, 12), sH[: 144 * k].reshape(-1, 12, 12))
        return out + (na.value,) if with_active else out

    def ipc_max_step(self, scene, pairs, x, dx, sep=0.1, safety=0.8):
        """ipc mode maxStep over pairs (n x 5 int32) of scene along dx from x, keeping sep of
        each pair's distance: (alpha, toi)."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 5)
        x = np.ascontiguousarray(x, dtype=np.float64)
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        a, t = C.c_double(), C.c_double()
        self._ck(self.L.ndipc_ipc_max_step(self.ctx, scene, len(pr), ptr(pr, i32p), ptr(x, f64p), ptr(dx, f64p), sep,
                                           safety, C.byref(a), C.byref(t)), "ipc_max_step")
        return a.value, t.value

    def ipc_step(self, scene, dt, dhat=1e-3, kappa=1e4, sep=0.1, safety=0.8, newton_tol=1e-9, pcg_tol=1e-12,
                 inflation=None, max_newton=50, max_ls=30, pcg_max_it=2000):
        """ipc mode time step of one scene; (Newton iterations, line-search trials, pairs,
        PCG iterations)."""
        prm = IpcParams(dhat, kappa, sep, safety, newton_tol, pcg_tol, dhat if inflation is None else inflation,
                        max_newton, max_ls, pcg_max_it)
        st = np.zeros(4, dtype=np.int32)
        self._ck(self.L.ndipc_ipc_step(self.ctx, scene, dt, C.byref(prm), ptr(st, i32p)), "ipc_step")
        return tuple(int(v) for v in st)

    def friction_eval(self, scene, ints, dbl, nf, x, mu, eps, dt, hess=True):
        nd = len(x)
        val = C.c_double()
        g = np.zeros(nd)
        H = np.zeros((nd, nd)) if hess else None
        ints = np.ascontiguousarray(ints, dtype=np.int32)
        dbl = np.ascontiguousarray(dbl, dtype=np.float64)
        nf = np.ascontiguousarray(nf, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._ck(self.L.ndipc_friction_eval(self.ctx, scene, len(ints), ptr(ints, i32p), ptr(dbl, f64p),
                                            ptr(nf, f64p), ptr(x, f64p), mu, eps, dt, C.byref(val), ptr(g, f64p),
                                            ptr(H, f64p)), "friction_eval")
        return val.value, g, H

    def ccd_toi_batch(self, kind, poseA, poseB, a_pts, b_pts, earliest, tol=1e-8):
        n = len(poseA)
        pa = np.ascontiguousarray(poseA, dtype=np.float64).reshape(n, 14)
        pb = np.ascontiguousarray(poseB, dtype=np.float64).reshape(n, 14)
        a = np.zeros((n, 6))
        a[:, : a_pts.shape[1]] = a_pts
        b = np.zeros((n, 9))
        b[:, : b_pts.shape[1]] = b_pts
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(earliest, dtype=np.float64), (n,)))
        hit = np.zeros(n, dtype=np.int32)
        toi = np.zeros(n)
        pops = np.zeros(n, dtype=np.int32)
        self._ck(self.L.ndipc_ccd_toi_batch(self.ctx, kind, n, ptr(pa, f64p), ptr(pb, f64p), ptr(a, f64p),
                                            ptr(b, f64p), ptr(e, f64p), tol, ptr(hit, i32p), ptr(toi, f64p),
                                            ptr(pops, i32p)), "ccd_toi_batch")
        return hit, toi, pops

    def ccd_earliest_toi(self, scene, dt):
        o = C.c_double()
        self._ck(self.L.ndipc_ccd_earliest_toi(self.ctx, scene, dt, C.byref(o)), "ccd_earliest_toi")
        return o.value

    def pcg_solve(self, diag, offij, offb, b, tol=1e-14, max_it=500):
        nb = len(diag)
        d = np.ascontiguousarray(diag, dtype=np.float64)
        oij = np.ascontiguousarray(offij, dtype=np.int32)
        ob = np.ascontiguousarray(offb, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(6 * nb)
        it = C.c_int32()
        rel = C.c_double()
        self._ck(self.L.ndipc_pcg_solve(self.ctx, nb, ptr(d, f64p), len(oij), ptr(oij, i32p), ptr(ob, f64p),
                                        ptr(bb, f64p), tol, max_it, ptr(x, f64p), C.byref(it), C.byref(rel)), "pcg")
        return x, it.value, rel.value

    def pcg_solve_noalloc(self, nb, d_diag, noff, d_offij, d_offb, d_b, tol, max_it, d_x, d_iters, d_rel):
        """ndipc_pcg_solve_noalloc: every argument a device address (int); enqueued on the
        context stream, not synchronised."""
        self._ck(self.L.ndipc_pcg_solve_noalloc(self.ctx, int(nb), d_diag, int(noff), d_offij, d_offb, d_b,
                                                float(tol), int(max_it), d_x, d_iters, d_rel), "pcg_solve_noalloc")

    def autotune(self, trial_steps=3, dt=None):
        """Chooses the persistent kernels' work-group counts for this batch (state is
        restored); returns (counts[topk, narrow phase, scene-min flag, CCD], ms/step)."""
        dt = self.sc.dt if dt is None else np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.sc.n_scenes,))
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        out = np.zeros(4, dtype=np.int32)
        ms = C.c_double()
        self._ck(self.L.ndipc_autotune(self.ctx, ptr(dt, f64p), int(trial_steps), ptr(out, i32p), C.byref(ms)),
                 "autotune")
        return out, ms.value

    def set_launch_grids(self, grids):
        """Sets the four launch shapes of autotune directly (0 keeps a default)."""
        g = np.ascontiguousarray(grids, dtype=np.int32)
        assert g.shape == (4,)
        self._ck(self.L.ndipc_set_launch_grids(self.ctx, ptr(g, i32p)), "set_launch_grids")

    def use_graph(self, on=True):
        """Replay the captured step graph (default) or enqueue kernels one by one."""
        self._ck(self.L.ndipc_use_graph(self.ctx, 1 if on else 0), "use_graph")

    def timing_enable(self, on=True):
        self._ck(self.L.ndipc_timing_enable(self.ctx, 1 if on else 0), "timing_enable")

    def timing_get(self, name):
        avg = C.c_double()
        cnt = C.c_int32()
        self._ck(self.L.ndipc_timing_get(self.ctx, name.encode(), C.byref(avg), C.byref(cnt)), "timing_get")
        return avg.value, cnt.value

    def round_ops(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros((len(a), 6))
        self._ck(self.L.ndipc_round_ops(self.ctx, ptr(a, f64p), ptr(b, f64p), len(a), ptr(out, f64p)), "round_ops")
        return out

    def elemfn(self, fn, x, y=None):
        """fn in ("log", "sin", "cos", "atan2", "acos"); atan2 takes (x, y) = (y_arg, x_arg) of C."""
        code = ELEMFN.index(fn)
        x = np.ascontiguousarray(x, dtype=np.float64)
        yy = np.ascontiguousarray(np.zeros_like(x) if y is None else y, dtype=np.float64)
        out = np.zeros(len(x))
        self._ck(self.L.ndipc_elemfn(self.ctx, code, ptr(x, f64p), ptr(yy, f64p), len(x), ptr(out, f64p)), "elemfn")
        return out
