"""Scene and state I/O (include/ndipc_io.h, nexdynipc_amd/csrc/scene_io.cpp) on the CPU:
the reference's own contact example (tests/golden/basic_contact.json, a copy of
assets/examples/example_basic_contact/basic_contact.json; its OBJ meshes are not in the
reference snapshot, so synthetic ones stand in), the loader rules of SceneLoader.cpp and
Simulation.cpp, the OBJ reader, the JSON reader, and the CSV exporter's round trip."""
import csv
import json
import os
import shutil

import numpy as np
import pytest

from nexdynipc_amd import scene_io, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "nexdynipc_amd", "_build", "libndipc.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libndipc.so not built")


def _write_obj(path, V, F, quads=False):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("# synthetic mesh\n")
        for v in V:
            f.write("v %.17g %.17g %.17g\n" % tuple(v))
        for t in F:
            f.write("f %d/1/1 %d//2 %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))


@pytest.fixture
def basic_contact(tmp_path):
    d = tmp_path / "case"
    d.mkdir()
    shutil.copy(os.path.join(HERE, "golden", "basic_contact.json"), d / "basic_contact.json")
    V1, F1 = scenes.box_mesh(2.0, 0.3, 2.0)
    V2, F2 = scenes.icosphere_mesh(1, 0.1)
    _write_obj(str(d / "models/obj/basic_contact/Body1.obj"), V1, F1)
    _write_obj(str(d / "models/obj/basic_contact/Body2.obj"), V2, F2)
    return d / "basic_contact.json", (V1, F1, V2, F2)


def test_reference_contact_example_loads_as_the_reference_reads_it(basic_contact):
    path, (V1, F1, V2, F2) = basic_contact
    s = scene_io.load_scene(path)
    sc = s.scenes
    assert sc.n_bodies == 2 and list(sc.body_id) == [3, 4] and sc.names == ["Body1", "Body2"]
    assert np.array_equal(sc.mass, [62.8, 4.11025038844665])
    assert np.array_equal(sc.inertia.reshape(2, 3, 3)[0], np.diag([0.261666666666667, 1.04666666666667,
                                                                    0.889666666666667]))
    assert np.array_equal(sc.p, [[0.0, -0.15, 0.0], [0.0, 0.0, 0.0]])
    assert np.array_equal(sc.q, [[1, 0, 0, 0], [1, 0, 0, 0]]) and not sc.v.any() and not sc.w.any()
    # meshes as written (the OBJ reader keeps the file's vertices and faces)
    assert list(sc.n_verts) == [len(V1), len(V2)] and list(sc.n_faces) == [len(F1), len(F2)]
    assert np.array_equal(sc.verts.reshape(-1, 3)[: len(V1)], V1)
    assert np.array_equal(sc.faces.reshape(-1, 3)[len(F1):], F2)
    # the fixed joint on body id 3 (world index 0) gets the scene's joint_stiffness
    # (default 1e8 replaced, Simulation.cpp:20-24)
    assert list(sc.joint_type) == [0] and list(sc.joint_body) == [0, -1]
    assert np.array_equal(sc.joint_data[:7], [0.0, -0.15, 0.0, 1.0, 0.0, 0.0, 0.0]) and sc.joint_mu[0] == 1e7
    # the contact pair by name, its friction the solver's coefficient
    assert list(sc.allowed_pairs) == [3, 4] and list(sc.allowed_mu) == [0.2]
    p = s.params
    assert p.dhat == 2e-4 and p.kappa == 1e6 and p.enable_adaptive_barrier == 0
    assert p.friction_mu == 0.2 and p.enable_friction == 1 and p.enable_ccd == 1 and p.alm_max_iters == 6
    st = s.settings
    assert st.dt == 0.001 and st.max_time == 1.0 and st.integrator_type == b"ImplicitNewmark"
    assert st.joint_stiffness == 1e7 and st.contact_stiffness == 1e6 and st.n_contact_pairs == 1
    assert st.output_dir == b"output/basic_contact" and st.output_name == b"basic_contact"
    assert s.unsupported == []


def test_loader_rules(tmp_path):
    """Defaults, joint signatures (duplicates skipped, 'revolute' read as 'hinge', the
    stiffness not part of the signature), fixed-joint orientation order [x, y, z, w],
    unknown body ids, unsupported entries, disabled pairs and the output-name rules."""
    V, F = scenes.box_mesh(1, 1, 1)
    _write_obj(str(tmp_path / "cube.obj"), V, F)
    scene = {
        "bodies": [{"id": 7, "name": "A", "mesh": "cube.obj", "position": [0, 1, 0]},
                   {"id": 9, "mesh": "missing.obj"}, {"name": "C", "mass": 2.5}],
        "joints": [
            {"type": "revolute", "body_a": 7, "body_b": 9, "anchor_a": [0.5, 0, 0], "anchor_b": [-0.5, 0, 0],
             "axis_a": [0, 0, 1], "axis_b": [0, 0, 1]},
            {"type": "hinge", "body_a": 7, "body_b": 9, "anchor_a": [0.5, 0, 0], "anchor_b": [-0.5, 0, 0],
             "axis_a": [0, 0, 1], "axis_b": [0, 0, 1], "stiffness": 5.0},
            {"type": "fixed", "body": 123, "position": [1, 2, 3], "orientation": [0.1, 0.2, 0.3, 0.9]},
            {"type": "spherical", "body_a": 7, "body_b": 9, "anchor_a": [0, 0, 0], "anchor_b": [0, 0, 0]},
        ],
        "constraints": [{"type": "fixed", "body": 9, "position": [0, 0, 0], "stiffness": 3e4}],
        "interactions": {"contact": {"pairs": [
            {"object1": "A", "object2": 9, "friction": 0.4},
            {"object1": "A", "object2": "C", "enabled": False, "friction": 0.9},
            {"object1": "nobody", "object2": "A"}]}},
    }
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(scene))
    s = scene_io.load_scene(path)
    sc = s.scenes
    assert list(sc.body_id) == [7, 9, -1] and sc.names == ["A", "Body", "C"]
    assert np.array_equal(sc.mass, [1.0, 1.0, 2.5])
    assert np.array_equal(sc.inertia.reshape(3, 3, 3)[1], np.eye(3))
    assert list(sc.n_verts) == [8, 0, 0]
    # revolute == hinge: the second entry is a duplicate; spherical is not run here
    assert list(sc.joint_type) == [1, 0, 0]
    assert list(sc.joint_body) == [0, 1, 0, -1, 1, -1]  # unknown id 123 -> index 0
    assert np.array_equal(sc.joint_data[16:23], [1, 2, 3, 0.9, 0.1, 0.2, 0.3])
    assert np.array_equal(sc.joint_mu, [1000.0, 1000.0, 3e4])  # SimulationConfig joint_stiffness 1000
    assert list(sc.allowed_pairs) == [7, 9] and list(sc.allowed_mu) == [0.4]
    # the mean friction of the enabled pairs that give one (SceneLoader.cpp:683-696)
    assert s.params.friction_mu == 0.4 and s.params.enable_adaptive_barrier == 1
    assert s.params.dhat == 1e-3 and s.params.kappa == 0.0
    assert s.settings.output_name == b"rules" and s.settings.output_dir == b"output/rules"
    assert any("spherical" in u for u in s.unsupported) and any("missing.obj" in u for u in s.unsupported)


def test_obj_reader(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# c\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 1\n"
                 "f 1/1/1 2/2/2 3/3/3 4/4/4\n"   # a quad: two fan triangles
                 "f 1 2 x\n"                      # unparsable token: line dropped
                 "f 1 2\n"                        # too short
                 "f -1 2 5\n"                     # negative index: triangle dropped
                 "f 1//3 3//1 5\n")
    V, F = scene_io.load_obj(p)
    assert V.shape == (5, 3)
    assert F.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 4]]
    (tmp_path / "empty.obj").write_text("v 0 0 0\n")
    with pytest.raises(RuntimeError, match="no valid mesh"):
        scene_io.load_obj(tmp_path / "empty.obj")


@pytest.mark.parametrize("text,msg", [
    ('{"bodies": [1, 2,]}', "bad number"),
    ('{"a": "x\\q"}', "bad escape"),
    ('{"a": 01}', "expected"),
    ('{"a": 1} x', "trailing"),
    ('{"a": "\\u12"}', "escape"),
    ('[1, 2', "expected"),
])
def test_json_reader_rejects_malformed_input(tmp_path, text, msg):
    p = tmp_path / "bad.json"
    p.write_text(text)
    with pytest.raises(RuntimeError, match=msg):
        scene_io.load_scene(p)


def test_json_reader_values(tmp_path):
    """Escapes, unicode, exponents and duplicate keys (the last wins)."""
    p = tmp_path / "v.json"
    p.write_text('{"bodies": [{"name": "a\\"b\\\\c\\u00e9\\ud83d\\ude00", "mass": 1.5e-3, "mass": -2.25E+2,'
                 ' "position": [1e-300, -0.0, 12345678901234567890]}]}')
    s = scene_io.load_scene(p)
    assert s.scenes.names == ['a"b\\cé\U0001F600']
    assert s.scenes.mass[0] == -225.0
    assert s.scenes.p[0].tolist() == [1e-300, -0.0, 12345678901234567890.0]


def test_exporter_round_trip(tmp_path, basic_contact):
    path, _ = basic_contact
    s = scene_io.load_scene(path)
    rng = np.random.default_rng(0)
    n = s.scenes.n_bodies
    ex = scene_io.Exporter(tmp_path / "out", "run")
    frames = []
    for k in range(3):
        st = [rng.normal(size=(n, 3)), rng.normal(size=(n, 4))] + [rng.normal(size=(n, 3)) for _ in range(4)]
        frames.append(st)
        ex.frame(s, k, 1e-3 * k, *st)
        ex.kpi(k, 1e-3 * k, 0.0, 0.0, rng.normal(), rng.normal())
    ex.close()
    with open(tmp_path / "out" / "run.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == ("frame_idx,time,body_id,body_name,pos_x,pos_y,pos_z,qx,qy,qz,qw,vx,vy,vz,ax,ay,az,"
                       "wx,wy,wz,alphax,alphay,alphaz").split(",")
    assert len(rows) == 1 + 3 * n
    for k, (p, q, v, w, al, aa) in enumerate(frames):
        for b in range(n):
            r = rows[1 + k * n + b]
            assert int(r[0]) == k and float(r[1]) == 1e-3 * k and int(r[2]) == s.scenes.body_id[b]
            assert r[3] == s.scenes.names[b]
            vals = np.array([float(x) for x in r[4:]])
            ref = np.concatenate([p[b], q[b, [1, 2, 3, 0]], v[b], al[b], w[b], aa[b]])
            assert np.array_equal(vals, ref)  # %.17g round-trips exactly
    with open(tmp_path / "out" / "run_kpi.csv") as f:
        krows = list(csv.reader(f))
    assert krows[0] == ["frame_idx", "time", "V_w", "T_sat", "R_dual", "max_constraint_violation"]
    assert len(krows) == 4


@pytest.mark.parametrize("m", [1, 2, 5, 24])
def test_grid_box_mesh_is_a_closed_outward_box(m):
    """The gridded box: vertex and face counts, Euler characteristic 2, every edge shared
    by two faces that traverse it in opposite directions, the face planes exactly at
    +-size/2 and the enclosed volume and inertia those of the box."""
    sx, sy, sz = 1.5, 0.1, 0.75
    V, F = scenes.grid_box_mesh(sx, sy, sz, m)
    assert len(V) == 6 * (m - 1) ** 2 + 12 * (m - 1) + 8 and len(F) == 12 * m * m
    directed = {}
    for a, b, c in F:
        for u, v in ((a, b), (b, c), (c, a)):
            assert (u, v) not in directed
            directed[(u, v)] = 1
    assert all((v, u) in directed for (u, v) in directed)
    n_edges = len(directed) // 2
    assert len(V) - n_edges + len(F) == 2
    for k, h in enumerate((sx, sy, sz)):
        assert set(np.unique(np.abs(V[:, k])).tolist()) >= {0.5 * h}
        assert np.abs(V[:, k]).max() == 0.5 * h
    mass, inertia = scenes.polyhedron_mass_properties(V, F)
    mb, Ib = scenes.box_mass_properties(sx, sy, sz)
    assert abs(mass - mb) <= 1e-12 * mb
    assert np.allclose(inertia, Ib, rtol=1e-10, atol=1e-12 * Ib.max())
