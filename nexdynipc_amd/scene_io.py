"""Scene and state I/O (include/ndipc_io.h) from Python: the reference's JSON scenes and
OBJ meshes (nexdynipc/src/App/SceneLoader.cpp), the IPCSolver settings Simulation::run
derives from them (nexdynipc/src/App/Simulation.cpp:14-56) and StateExporter's CSV
(nexdynipc/src/App/StateExporter.cpp). The parsing runs in libndipc.so (host C++).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import scenes
from ._lib import Params, SceneDesc, f64p, i32p, library, ptr


class SimSettings(C.Structure):
    """ndipc_sim_settings (include/ndipc_io.h)."""
    _fields_ = [
        ("dt", C.c_double), ("max_time", C.c_double), ("newmark_beta", C.c_double), ("newmark_gamma", C.c_double),
        ("joint_stiffness", C.c_double), ("alm_max_iters", C.c_int32), ("kpi_gate_enabled", C.c_int32),
        ("alm_constraint_tolerance", C.c_double), ("alm_dual_tolerance", C.c_double),
        ("alm_hardening_trigger", C.c_double), ("alm_hardening_ratio", C.c_double),
        ("kpi_v_w_max", C.c_double), ("kpi_t_sat_max", C.c_double), ("kpi_r_dual_max", C.c_double),
        ("contact_enabled", C.c_int32), ("n_contact_pairs", C.c_int32),
        ("contact_thickness", C.c_double), ("contact_stiffness", C.c_double), ("contact_friction", C.c_double),
        ("integrator_type", C.c_char * 32), ("output_dir", C.c_char * 512), ("output_name", C.c_char * 256),
    ]


class IOError_(RuntimeError):
    pass


def _lib():
    L = library()
    if getattr(L, "_io_ready", False):
        return L
    vp = C.c_void_p
    sig = {
        "ndipc_io_last_error": ([], C.c_char_p),
        "ndipc_scene_load": ([C.c_char_p, C.POINTER(vp)], C.c_int32),
        "ndipc_scene_free": ([vp], None),
        "ndipc_scene_settings": ([vp, C.POINTER(SimSettings)], C.c_int32),
        "ndipc_scene_describe": ([vp, C.POINTER(SceneDesc)], C.c_int32),
        "ndipc_scene_params": ([vp, C.POINTER(Params)], C.c_int32),
        "ndipc_scene_state": ([vp, f64p, f64p, f64p, f64p], C.c_int32),
        "ndipc_scene_body_name": ([vp, C.c_int32, C.c_char_p, C.c_int32], C.c_int32),
        "ndipc_scene_unsupported": ([vp, C.c_int32, C.c_char_p, C.c_int32], C.c_int32),
        "ndipc_obj_load": ([C.c_char_p, i32p, f64p, C.c_int32, i32p, i32p, C.c_int32], C.c_int32),
        "ndipc_export_open": ([C.c_char_p, C.c_char_p, C.POINTER(vp)], C.c_int32),
        "ndipc_export_frame": ([vp, vp, C.c_int32, C.c_double] + [f64p] * 6, C.c_int32),
        "ndipc_export_kpi": ([vp, C.c_int32] + [C.c_double] * 5, C.c_int32),
        "ndipc_export_close": ([vp], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    L._io_ready = True
    return L


def _ck(rc, what):
    if rc < 0:
        msg = _lib().ndipc_io_last_error().decode()
        raise IOError_(msg if msg.startswith(what) else f"{what}: {msg}")
    return rc


def load_obj(path):
    """(V[n, 3], F[m, 3]) of an OBJ file as the reference reads it."""
    L = _lib()
    nv, nf = C.c_int32(), C.c_int32()
    _ck(L.ndipc_obj_load(str(path).encode(), C.byref(nv), None, 0, C.byref(nf), None, 0), "obj_load")
    V = np.zeros((nv.value, 3))
    F = np.zeros((nf.value, 3), dtype=np.int32)
    _ck(L.ndipc_obj_load(str(path).encode(), C.byref(nv), ptr(V, f64p), nv.value, C.byref(nf), ptr(F, i32p),
                         nf.value), "obj_load")
    return V, F


@dataclasses.dataclass
class LoadedScene:
    scenes: scenes.Scenes      # one scene, in the upload layout
    params: Params             # the IPCSolver settings Simulation::run applies
    settings: SimSettings
    unsupported: list          # entries the device path does not run
    handle: int                # ndipc_scene_file*

    def close(self):
        if self.handle:
            _lib().ndipc_scene_free(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_scene(path) -> LoadedScene:
    """SceneLoader::load(path, world, config) for a .json scene."""
    L = _lib()
    h = C.c_void_p()
    _ck(L.ndipc_scene_load(str(path).encode(), C.byref(h)), "scene_load")
    st = SimSettings()
    d = SceneDesc()
    pr = Params()
    L.ndipc_scene_settings(h, C.byref(st))
    L.ndipc_scene_describe(h, C.byref(d))
    L.ndipc_scene_params(h, C.byref(pr))
    n, J, A = d.n_bodies, d.n_joints, d.n_allowed

    def arr(p, count, dtype):
        if count <= 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype).copy()

    nvt = int(sum(arr(d.n_verts, n, np.int32))) if n else 0
    nft = int(sum(arr(d.n_faces, n, np.int32))) if n else 0
    p = np.zeros((n, 3))
    q = np.zeros((n, 4))
    v = np.zeros((n, 3))
    w = np.zeros((n, 3))
    L.ndipc_scene_state(h, ptr(p, f64p), ptr(q, f64p), ptr(v, f64p), ptr(w, f64p))
    names = []
    buf = C.create_string_buffer(512)
    for b in range(n):
        L.ndipc_scene_body_name(h, b, buf, 512)
        names.append(buf.value.decode())
    k = L.ndipc_scene_unsupported(h, -1, None, 0)
    unsup = []
    for i in range(k):
        L.ndipc_scene_unsupported(h, i, buf, 512)
        unsup.append(buf.value.decode())
    sc = scenes.Scenes(
        n_scenes=1, scene_body_off=np.array([0, n], dtype=np.int32), body_id=arr(d.body_id, n, np.int32),
        vert_off=arr(d.vert_off, n, np.int32), n_verts=arr(d.n_verts, n, np.int32),
        face_off=arr(d.face_off, n, np.int32), n_faces=arr(d.n_faces, n, np.int32),
        verts=arr(d.verts, 3 * nvt, np.float64), faces=arr(d.faces, 3 * nft, np.int32),
        mass=arr(d.mass, n, np.float64), inertia=arr(d.inertia, 9 * n, np.float64),
        scene_joint_off=np.array([0, J], dtype=np.int32), joint_type=arr(d.joint_type, J, np.int32),
        joint_body=arr(d.joint_body, 2 * J, np.int32), joint_data=arr(d.joint_data, 16 * J, np.float64),
        joint_mu=arr(d.joint_mu, J, np.float64), scene_allowed_off=np.array([0, A], dtype=np.int32),
        allowed_pairs=arr(d.allowed_pairs, 2 * A, np.int32), allowed_mu=arr(d.allowed_mu, A, np.float64),
        p=p, q=q, v=v, w=w, dt=np.array([st.dt]), dhat=np.array([pr.dhat]), mu=np.array([pr.friction_mu]),
        enable_friction=np.array([pr.enable_friction], dtype=np.int32), names=names)
    return LoadedScene(sc, pr, st, unsup, h.value)


class Exporter:
    """StateExporter: <dir>/<name>.csv and <dir>/<name>_kpi.csv (%.17g)."""

    def __init__(self, out_dir, name):
        self.L = _lib()
        self.h = C.c_void_p()
        _ck(self.L.ndipc_export_open(str(out_dir).encode(), str(name).encode(), C.byref(self.h)), "export_open")

    def frame(self, loaded: LoadedScene, frame, time, p, q, v, w, al, aa):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (p, q, v, w, al, aa)]
        _ck(self.L.ndipc_export_frame(self.h, loaded.handle, int(frame), float(time), *[ptr(a, f64p) for a in arrs]),
            "export_frame")

    def kpi(self, frame, time, v_w, t_sat, r_dual, viol):
        _ck(self.L.ndipc_export_kpi(self.h, int(frame), float(time), v_w, t_sat, r_dual, viol), "export_kpi")

    def close(self):
        if self.h:
            self.L.ndipc_export_close(self.h)
            self.h = None
