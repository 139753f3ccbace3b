"""Meshes above 2048 vertices, edges or faces on the GPU against the oracle. Their top-K
lists and the scene-min flag's vertex arrays do not fit the LDS-resident kernels and run
from per-work-group HBM slabs; lists of the small meshes of the same scene stay in LDS.

- the standalone selection on score lists of 2049 to 65535 entries (ties, distinct
  scores, and the McIlroy adversary inputs of tests/golden at 4096 and 16384 entries
  through test_gpu_parity), and its limits;
- the 2000-contact cap with tied distances across the cut (std::sort's order);
- body pairs, candidates, the narrow phase and the scene-min flag on a scene of a
  gridded plate (3,458 vertices, 10,368 edges, 6,912 faces) and two subdivision-4
  icospheres (2,562 / 7,680 / 5,120), bitwise;
- time steps of that scene at two scales, every state bitwise with the same Newton,
  line-search and contact counts: millimetre bodies whose centres are within dhat (the
  barrier and the 2000-contact cap active) and metre bodies a small gap apart;
- the ipc-mode step on the metre scene (8,576 pair records), bitwise.
"""
import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scenes

pytestmark = pytest.mark.gpu


def test_select_top_k_long_lists(engine_factory):
    e = engine_factory(scenes.make_config("c1"))
    rng = np.random.default_rng(12)
    for trial in range(40):
        n = [2049, 4097, 32768, 65535][trial] if trial < 4 else int(rng.integers(2049, 65536))
        k = int(rng.integers(1, 2049)) if trial % 2 else int(rng.integers(1, 40))
        if trial % 3 == 0:
            s = rng.permutation(n).astype(np.float64)
        else:
            s = rng.integers(0, 1 + trial % 9, size=n).astype(np.float64) * 0.25
        assert O.select_top_k(s, k).tolist() == e.select_top_k(s, k).tolist(), (trial, n, k)
    with pytest.raises(RuntimeError, match="65535"):
        e.select_top_k(np.zeros(65536), 4)
    with pytest.raises(RuntimeError, match="2048"):
        e.select_top_k(np.zeros(3000), 2049)
    # the context still works
    s = rng.permutation(5000).astype(np.float64)
    assert O.select_top_k(s, 32).tolist() == e.select_top_k(s, 32).tolist()


def _macro():
    return scenes.large_mesh_scene(scale=1.0, gap=2e-4, dt=5e-3)


def test_large_mesh_geometry(engine_factory):
    sc = _macro()
    assert max(sc.n_verts) > 2048 and max(sc.n_faces) > 2048
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    for r in (0.0, 1e-3, 0.05):
        np.testing.assert_array_equal(o.body_pairs(r), e.body_pairs(0, r))
        assert o.min_body_pair_distance(r) == e.min_body_pair_distance(0, r)
    ro = o.candidates(1e-3)
    rg = e.candidates(0, 1e-3)
    assert sum(len(x) for x in ro) > 1000
    for a, b in zip(ro, rg):
        np.testing.assert_array_equal(a, b)
    rng = np.random.default_rng(8)
    for pert in (None, 1e-4):
        xs = None
        if pert is not None:
            xs = np.zeros(6 * o.n)
            for k in range(3):
                xs[k::6] = sc.p[:, k]
            xs += rng.normal(0, pert, size=xs.shape)
        oi, od = o.narrow_phase(xs)
        gi, gd = e.narrow_phase(0, xs)
        assert len(oi) > 0
        np.testing.assert_array_equal(oi, gi)
        assert np.array_equal(od.view(np.int64), gd.view(np.int64))
    assert e.scene_min_distance_flag(0) == int(o.scene_min_distance() < 1e-9)


def test_contact_cap_keeps_the_reference_order_of_tied_distances(engine_factory):
    """More than 2000 raw contacts, many of them at the same distance across the cut: the
    kept 2000 and their order are those of std::sort (libstdc++'s introsort) by distance,
    bitwise, as the oracle keeps them."""
    sc = scenes.large_mesh_scene()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    dhat = float(sc.dhat[0])
    o.candidates(dhat)
    e.candidates(0, dhat)
    o.set_narrow_raw(True)
    raw_i, raw_d = o.narrow_phase(None)
    o.set_narrow_raw(False)
    cut = np.sort(raw_d[:, 6])[1999]
    assert len(raw_d) > 2000 and (raw_d[:, 6] == cut).sum() > 1 and (raw_d[:, 6] < cut).sum() < 2000
    oi, od = o.narrow_phase(None)
    gi, gd = e.narrow_phase(0, None)
    assert len(oi) == 2000
    np.testing.assert_array_equal(oi, gi)
    assert np.array_equal(od.view(np.int64), gd.view(np.int64))


@pytest.mark.parametrize("order", ["box_on_plate", "plate_on_box"])
def test_large_mesh_scene_min_flag_touching(engine_factory, order):
    """A box resting exactly on the gridded plate (its bottom vertices on the plate's top
    faces), and the plate resting exactly on a box (the plate's bottom vertices on the
    box's top face): the flag is set either way, with the large mesh as the vertex side
    or as the face side of the test (vertices of body i against faces of body j > i)."""
    V, F = scenes.grid_box_mesh(1.5, 0.5, 1.5, 24)
    m, I = scenes.polyhedron_mass_properties(V, F)
    ident = np.array([1.0, 0.0, 0.0, 0.0])
    plate = lambda y: scenes.BodySpec("grid_plate", V, F, m, I, np.array([0.0, y, 0.0]), ident.copy(),
                                      np.zeros(3), np.zeros(3))
    box = lambda y: scenes._box_body("unit_box", 0.5, 0.5, 0.5, [0.25, y, -0.5])
    ground = scenes._box_body("far_ground", 1.0, 1.0, 1.0, [20.0, -5.0, 0.0], static=True)
    if order == "box_on_plate":
        bodies = [ground, box(0.5), plate(0.0)]  # box bottom at 0.25 = plate top
    else:
        bodies = [ground, plate(0.5), box(0.0)]  # plate bottom at 0.25 = box top
    sc = scenes.pack([scenes.SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.0)])
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    ref = o.scene_min_distance() < 1e-9
    assert ref
    assert e.scene_min_distance_flag(0) == 1


@pytest.mark.parametrize("name,steps", [("micro", 8), ("macro", 8)])
def test_large_mesh_steps(engine_factory, name, steps):
    sc = scenes.large_mesh_scene() if name == "micro" else _macro()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    dt = float(sc.dt[0])
    contacts = 0
    for k in range(steps):
        o.step(dt)
        e.step(1)
        for a, b in zip(o.get_state()[:4], e.get_state()[:4]):
            assert np.array_equal(a, b), k
        so, sg = np.asarray(o.stats())[:3], e.stats()[0][:3]
        assert np.array_equal(sg, so), (k, so, sg)
        contacts += int(so[2])
    assert contacts > 0


def test_large_mesh_ipc_step(engine_factory):
    sc = _macro()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    kw = dict(dhat=1e-3, kappa=1e7, inflation=2e-3)
    for k in range(3):
        so = o.ipc_step(5e-3, **kw)
        sg = e.ipc_step(0, 5e-3, **kw)
        assert so == sg, (k, so, sg)
        for a, b in zip(o.get_state()[:4], e.get_state()[:4]):
            assert np.array_equal(a, b), k
