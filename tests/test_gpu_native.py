"""The drop-in paths against the oracle on the GPU:

  - tests/native/run_scene (C++, include/ndipc.h + include/ndipc_io.h only): a JSON
    scene in the reference's format with OBJ meshes, run as the reference's
    main -> Simulation::run would, its per-frame state CSV (%.17g) bitwise equal to the
    oracle's steps of the same loaded scene;
  - nexdynipc_amd.IPCSolver on a World (the reference's solver interface): bitwise equal
    to the oracle step by step, and unchanged by a re-upload in the middle of a run
    (joint multipliers and barrier state carried over).
"""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scene_io, scenes
from nexdynipc_amd.solver import FixedJoint, IPCSolver, RigidBody, World

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "tests", "native", "_build", "run_scene")


def _oracle_params(pr):
    return O.Params(dhat=pr.dhat, kappa=pr.kappa, friction_mu=pr.friction_mu, enable_friction=pr.enable_friction,
                    enable_ccd=pr.enable_ccd, enable_adaptive_barrier=pr.enable_adaptive_barrier,
                    alm_max_iters=pr.alm_max_iters, alm_constraint_tol=pr.alm_constraint_tol,
                    alm_dual_tol=pr.alm_dual_tol, alm_contact_tol=pr.alm_contact_tol,
                    alm_hardening_trigger=pr.alm_hardening_trigger, alm_hardening_ratio=pr.alm_hardening_ratio,
                    ccd_safety_factor=pr.ccd_safety_factor)


def _write_obj(path, V, F):
    with open(path, "w") as f:
        for v in V:
            f.write("v %.17g %.17g %.17g\n" % tuple(v))
        for t in F:
            f.write("f %d %d %d\n" % tuple(t + 1))


def test_native_runner_matches_oracle(tmp_path):
    """C1's bodies as a JSON scene (the loader reads positions only, so the cubes start
    upright), 60 steps through the C++ runner."""
    assert os.path.exists(RUNNER), "tests/native/_build/run_scene missing: run __graft_entry__.build()"
    c1 = scenes.make_config("c1")
    Vg, Fg = scenes.box_mesh(10.0, 1.0, 10.0)
    Vc, Fc = scenes.box_mesh(1.0, 1.0, 1.0)
    _write_obj(tmp_path / "ground.obj", Vg, Fg)
    _write_obj(tmp_path / "cube.obj", Vc, Fc)
    bodies = []
    for b in range(c1.n_bodies):
        I = c1.inertia.reshape(-1, 3, 3)[b]
        bodies.append({"id": b, "name": f"body{b}", "mass": float(c1.mass[b]), "position": c1.p[b].tolist(),
                       "inertia": I.tolist(), "mesh": "ground.obj" if b == 0 else "cube.obj"})
    scene = {"case_name": "c1json", "settings": {"dt": 0.01, "max_time": 0.6, "integrator_type": "ImplicitEuler",
                                                  "joint_stiffness": 1e8, "alm_max_iters": 10},
             "bodies": bodies,
             "joints": [{"type": "fixed", "body": 0, "position": c1.p[0].tolist()}],
             "interactions": {"contact": {"global_params": {"default_contact_thickness": 1e-3}}}}
    path = tmp_path / "c1.json"
    path.write_text(json.dumps(scene))
    out = tmp_path / "out"
    r = subprocess.run([RUNNER, str(path), "--out", str(out), "--name", "c1json"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "60 steps" in r.stdout
    loaded = scene_io.load_scene(path)
    o = O.OracleScene(loaded.scenes, 0, _oracle_params(loaded.params))
    with open(out / "c1json.csv") as f:
        rows = list(csv.DictReader(f))
    n = loaded.scenes.n_bodies
    assert len(rows) == 61 * n
    for k in range(61):
        if k:
            o.step(0.01)
        p, q = o.get_state()[:2]
        for b in range(n):
            row = rows[k * n + b]
            assert int(row["frame_idx"]) == k and int(row["body_id"]) == b
            got_p = [float(row[c]) for c in ("pos_x", "pos_y", "pos_z")]
            got_q = [float(row[c]) for c in ("qw", "qx", "qy", "qz")]
            assert got_p == p[b].tolist() and got_q == q[b].tolist(), (k, b)
    # the cubes fell onto the ground and rest there
    assert float(rows[-1]["pos_y"]) < 1.2
    with open(out / "c1json_kpi.csv") as f:
        assert len(list(csv.reader(f))) == 62


def _c1_world(sc):
    w = World()
    for b in range(sc.n_bodies):
        vo, nv, fo, nf = sc.vert_off[b], sc.n_verts[b], sc.face_off[b], sc.n_faces[b]
        w.addBody(RigidBody(id=int(sc.body_id[b]), name=f"b{b}", position=sc.p[b].copy(),
                            orientation=sc.q[b].copy(), velocity=sc.v[b].copy(), angular_velocity=sc.w[b].copy(),
                            mass=float(sc.mass[b]), inertia_body=sc.inertia.reshape(-1, 3, 3)[b].copy(),
                            shape=(sc.verts.reshape(-1, 3)[vo:vo + nv], sc.faces.reshape(-1, 3)[fo:fo + nf])))
    for j in range(sc.n_joints):
        w.addJoint(FixedJoint(int(sc.body_id[sc.joint_body[2 * j]]), sc.joint_data[16 * j:16 * j + 3].copy(),
                              sc.joint_data[16 * j + 3:16 * j + 7].copy()))
    return w


def test_ipc_solver_interface_matches_oracle():
    sc = scenes.make_config("c1")
    world = _c1_world(sc)
    solver = IPCSolver()
    solver.enableFriction(True)
    solver.setFrictionCoefficient(0.3)
    o = O.OracleScene(sc)
    try:
        for k in range(40):
            solver.step(world, 0.01)
            o.step(0.01)
            p, q, v, w = o.get_state()[:4]
            assert np.array_equal(np.array([b.position for b in world.bodies]), p), k
            assert np.array_equal(np.array([b.orientation for b in world.bodies]), q), k
            assert np.array_equal(np.array([b.velocity for b in world.bodies]), v), k
        assert solver.lastMaxConstraintViolation() >= 0.0 and solver.lastDualResidual() >= 0.0
    finally:
        solver.close()


def test_ipc_solver_reupload_keeps_the_trajectory():
    """A structural change forces a new upload; the joints' multipliers and stiffness
    and the barrier state are carried over, so the run continues bit for bit."""
    sc = scenes.make_config("c1")
    wa, wb = _c1_world(sc), _c1_world(sc)
    sa, sb = IPCSolver(), IPCSolver()
    try:
        for k in range(30):
            sa.step(wa, 0.01)
            if k == 12:  # the same pair filter written differently: a new structure signature
                sb.addAllowedContactPair(0, 1)
                sb.clearAllowedContactPairs()
                sb._sig = None
            sb.step(wb, 0.01)
            for x, y in zip(wa.bodies, wb.bodies):
                assert np.array_equal(x.position, y.position) and np.array_equal(x.orientation, y.orientation), k
        for ja, jb in zip(wa.joints, wb.joints):
            assert np.array_equal(ja.lambda_, jb.lambda_) and ja.stiffness == jb.stiffness
    finally:
        sa.close()
        sb.close()
