"""Pins the CPU oracle against the reference's own known-answer tests, closed forms,
brute force on small inputs and invariants (CPU only).

Reference tests restated (file:line in W-Moorer/NexDynIPC):
  tests/physics/test_ccd.cpp:79-125 (root finder), :163-210 / :294-326 (adaptive barrier),
  :230-256 (empty CCDSystem), :258-292 (edge-vertex TOI);
  tests/physics/test_friction.cpp:15-84 (potential), :299-326 / :388-417 (FD checks);
  tests/physics/test_contact_regression.cpp:424-461 (pairs, dhat);
  tests/geometry/test_bvh.cpp:328-381,437-469 (BVH == brute force);
  tests/geometry/test_aabb.cpp:136-140 (touching boxes overlap);
  tests/math/test_interval.cpp (outward rounding).
"""
import math

import os

import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scenes


# ---------------------------------------------------------------- interval / root finder
def test_root_finder_kats():
    r, lo, hi = O.root_finder_kat(0)  # x - 0.5
    assert r == 1 and abs(lo - 0.5) <= 1e-6 and abs(hi - 0.5) <= 1e-6
    r, lo, hi = O.root_finder_kat(1)  # x^2 - 0.25
    assert r == 1 and abs(lo - 0.5) <= 1e-6
    r, _, _ = O.root_finder_kat(2)  # x^2 + 1: no root
    assert r == 0


def test_directed_rounding_brackets_exact_result():
    from fractions import Fraction

    rng = np.random.default_rng(0)
    a = rng.normal(size=500) * np.exp(rng.uniform(-30, 30, 500))
    b = rng.normal(size=500) * np.exp(rng.uniform(-30, 30, 500))
    out = O.round_ops(a, b)
    for i in range(len(a)):
        fa, fb = Fraction(float(a[i])), Fraction(float(b[i]))
        for k, exact in ((0, fa + fb), (2, fa - fb), (4, fa * fb)):
            dn, up = out[i, k], out[i, k + 1]
            assert Fraction(dn) <= exact <= Fraction(up)
            # tight: at most one ulp apart, equal when exact
            assert up == dn or np.nextafter(dn, np.inf) == up
            if Fraction(dn) == exact:
                assert dn == up


# ---------------------------------------------------------------- CCD
def _pose(p0, p1, q0=(1, 0, 0, 0), q1=(1, 0, 0, 0)):
    return np.array(list(p0) + list(q0) + list(p1) + list(q1), dtype=np.float64)


def test_edge_vertex_toi_kat():
    # tests/physics/test_ccd.cpp:258-292: B moves (2,0,0) -> 0, edge (0,+-1,0), tol 1e-6
    r, toi, pops = O.ccd_toi(0, _pose((0, 0, 0), (0, 0, 0)), _pose((2, 0, 0), (0, 0, 0)),
                             [0, 0, 0], [0, -1, 0, 0, 1, 0], 1.0, 1e-6)
    assert r == 1 and 0.0 <= toi <= 1.0
    # the translational hull does not shrink with t, so the capped DFS ends just below 1
    assert pops == 1000 and toi > 0.99


def test_separated_trajectories_never_impact():
    rng = np.random.default_rng(1)
    for kind in (0, 1):
        for _ in range(20):
            pa = _pose(rng.normal(0, 0.1, 3), rng.normal(0, 0.1, 3))
            pb = _pose(rng.normal(0, 0.1, 3) + [10, 0, 0], rng.normal(0, 0.1, 3) + [10, 0, 0])
            r, toi, _ = O.ccd_toi(kind, pa, pb, rng.normal(0, 0.2, 6), rng.normal(0, 0.2, 9), 1.0, 1e-8)
            assert r == 0 and math.isinf(toi)


def test_face_vertex_toi_never_reports_impact():
    # the initial box fails the u + v <= 1 domain test (TimeOfImpact.cpp:130-147)
    rng = np.random.default_rng(2)
    for _ in range(50):
        r, toi, pops = O.ccd_toi(2, _pose(rng.normal(0, 0.1, 3), rng.normal(0, 0.1, 3)),
                                 _pose(rng.normal(0, 0.1, 3), rng.normal(0, 0.1, 3)),
                                 rng.normal(0, 0.2, 3), rng.normal(0, 0.2, 9), 1.0, 1e-8)
        assert r == 0 and pops == 1


# ---------------------------------------------------------------- potentials
def test_barrier_potential_closed_form_and_fd():
    dh, k = 1e-3, 1e9
    for d in (1e-6, 1e-4, 5e-4, 9.99e-4):
        v, g, h = O.barrier_potential(d, dh, k)
        assert v == pytest.approx(-k * (d - dh) ** 2 * math.log(d / dh), rel=1e-14)
        e = 1e-7 * d
        vp, gp, _ = O.barrier_potential(d + e, dh, k)
        vm, gm, _ = O.barrier_potential(d - e, dh, k)
        assert g == pytest.approx((vp - vm) / (2 * e), rel=1e-6)
        assert h == pytest.approx((gp - gm) / (2 * e), rel=1e-6)
        assert v > 0 and g < 0 and h > 0
    assert O.barrier_potential(1e-3, dh, k).tolist() == [0.0, 0.0, 0.0]
    assert math.isinf(O.barrier_potential(0.0, dh, k)[0])


def test_friction_potential_kats():
    # test_friction.cpp:15-84
    v, g, H = O.friction_potential([0.1, 0, 0], 10.0, 0.5)
    assert v > 0 and g[0] > 0 and abs(g[1]) < 1e-10
    assert O.friction_potential([0.1, 0, 0], 0.0, 0.5)[0] == 0.0
    assert O.friction_potential([0.2, 0, 0], 10.0, 0.5)[0] > v
    _, _, H = O.friction_potential([0.1, 0.05, 0], 10.0, 0.5)
    assert np.allclose(H, H.T, atol=1e-10)
    assert np.linalg.eigvalsh(H).min() >= -1e-10
    # test_friction.cpp:299-326 and :388-417: central FD, h = 1e-6
    u = np.array([0.1, 0.05, 0.02])
    _, g, H = O.friction_potential(u, 10.0, 0.5, 1e-8)
    for i in range(3):
        up, um = u.copy(), u.copy()
        up[i] += 1e-6
        um[i] -= 1e-6
        vp, gp, _ = O.friction_potential(up, 10.0, 0.5, 1e-8)
        vm, gm, _ = O.friction_potential(um, 10.0, 0.5, 1e-8)
        assert abs(g[i] - (vp - vm) / 2e-6) <= 1e-5
        assert np.max(np.abs(H[:, i] - (gp - gm) / 2e-6)) <= 1e-4
    # regularised branch
    _, g0, H0 = O.friction_potential([0, 0, 0], 10.0, 0.5, 1e-8)
    assert np.all(g0 == 0) and np.allclose(H0, (0.5 * 10 / 1e-8) * np.eye(3))


# ---------------------------------------------------------------- broad phase / candidates
def _random_box_scene(n, seed):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        s = rng.uniform(0.1, 1.0, 3)
        b = scenes._box_body(f"b{i}", *s, rng.uniform(-5, 5, 3), scenes.random_tilt(rng, 30))
        specs.append(b)
    return scenes.pack([scenes.SceneSpec(specs, dt=1e-2, dhat=1e-3, mu=0.0)])


def _world_aabbs(sc):
    boxes = []
    for b in range(sc.n_bodies):
        vo, nv = sc.vert_off[b], sc.n_verts[b]
        V = sc.verts.reshape(-1, 3)[vo:vo + nv]
        mn, mx = V.min(0), V.max(0)
        corners = np.array([[x, y, z] for x in (mn[0], mx[0]) for y in (mn[1], mx[1]) for z in (mn[2], mx[2])])
        W = np.array([scenes.rotate(sc.q[b], c) for c in corners]) + sc.p[b]
        boxes.append((W.min(0), W.max(0)))
    return boxes


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_body_pairs_equal_brute_force(seed):
    # test_bvh.cpp:328-381,437-469 — BVH query == all-pairs overlap
    sc = _random_box_scene(60, seed)
    o = O.OracleScene(sc)
    boxes = _world_aabbs(sc)
    for r in (0.0, 0.05, 0.5):
        got = {tuple(p) for p in o.body_pairs(r).tolist()}
        ref, near = set(), set()
        for i in range(len(boxes)):
            for j in range(i + 1, len(boxes)):
                a0, a1 = boxes[i][0] - r, boxes[i][1] + r
                b0, b1 = boxes[j][0] - r, boxes[j][1] + r
                if np.all(a0 <= b1) and np.all(a1 >= b0):
                    ref.add((i, j))
                # world boxes here are recomputed in numpy: a pair whose boxes touch
                # within 1e-9 on some axis may fall either way
                if np.min(np.abs(np.concatenate([a0 - b1, a1 - b0]))) < 1e-9:
                    near.add((i, j))
        assert len(ref) > 0
        assert got - near == ref - near


def test_touching_boxes_overlap_and_cube_ground_candidates():
    # test_aabb.cpp:136-140 + the 868-candidate cube-ground count of the survey
    sc = scenes.make_config("c1")
    sc.q[1] = [1, 0, 0, 0]
    sc.p[1] = [0, 0.5, 0]  # resting exactly on the ground
    sc.p[2:, 1] += 100.0   # the other cubes out of reach
    o = O.OracleScene(sc)
    pairs = o.body_pairs(0.0).tolist()
    assert [0, 1] in pairs
    vv, ev, ee, fv = o.candidates(1e-3)
    n_ground = sum(1 for c in fv if 0 in (c[0], c[1]))
    assert len(vv) == 64 and len(ev) == 288 and len(ee) == 324 and len(fv) == 192
    assert n_ground == 192


def test_select_top_k_without_ties_is_a_full_sort():
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = int(rng.integers(33, 2000))
        s = rng.permutation(n).astype(np.float64) * 0.5
        got = O.select_top_k(s, 32)
        assert got.tolist() == np.argsort(s, kind="stable")[:32].tolist()


def test_select_top_k_adversarial_inputs_keep_the_smallest():
    """On the adversary inputs (tests/golden/topk_adversarial.npz: libstdc++'s heap
    fallbacks taken) the kept scores are still the k smallest, in ascending order, with
    distinct indices."""
    adv = np.load(os.path.join(os.path.dirname(__file__), "golden", "topk_adversarial.npz"))
    assert len(adv.files) == 76
    for name in adv.files:
        s = adv[name].astype(np.float64)
        k = int(name.split("_k")[1])
        got = O.select_top_k(s, k)
        assert len(set(got.tolist())) == k
        assert np.array_equal(s[got], np.sort(s)[:k]), name


def test_scene_min_distance_brute_force_small():
    # vertices of body i are tested against faces of bodies j > i (AdaptiveBarrier.h:130-162)
    sc = scenes.make_config("c1")
    sc.q[1:3] = [1, 0, 0, 0]
    sc.p[1] = [3.0, 2.0, 3.0]
    sc.p[2] = [3.0, 1.0, 3.0]
    o = O.OracleScene(sc)
    assert o.scene_min_distance() < 1e-9
    sc.p[1] = [3.0, 2.0 + 1e-6, 3.0]
    o2 = O.OracleScene(sc)
    assert o2.scene_min_distance() == pytest.approx(1e-6, rel=1e-6)


# ---------------------------------------------------------------- adaptive barrier / step
def test_adaptive_barrier_initialisation_once():
    sc = scenes.make_config("c1")
    o = O.OracleScene(sc)
    k0, d0, init0 = o.barrier()
    assert k0 == 0.0 and init0 == 0
    o.step(1e-2)
    k1, d1, init1 = o.barrier()
    masses = sc.mass
    P = sc.p
    diag = np.linalg.norm(P.max(0) - P.min(0))
    assert init1 == 1 and d1 == 1e-3
    assert k1 == pytest.approx(1e11 * masses.mean() / diag ** 2, rel=1e-12)
    o.step(1e-2)
    assert o.barrier()[0] == k1


def test_free_fall_closed_form():
    # no active contact: implicit Euler free fall, v_{n+1} = v_n - g h, x_{n+1} = x_n + h v_{n+1}
    sc = scenes.make_config("c1")
    o = O.OracleScene(sc)
    h = 1e-2
    p0 = sc.p.copy()
    for _ in range(20):
        o.step(h)
    p, q, v, w, _, _ = o.get_state()
    n = 20
    expect_y = p0[1:, 1] - 9.81 * h * h * n * (n + 1) / 2
    assert np.max(np.abs(p[1:, 1] - expect_y)) <= 1e-9
    assert np.max(np.abs(v[1:, 1] + 9.81 * h * n)) <= 1e-9
    # the static ground stays put (FixedJoint with ALM)
    assert np.max(np.abs(p[0] - p0[0])) <= 1e-6


def test_micro_scene_activates_barrier_and_friction():
    # millimetre cubes with centre separations below dhat: contacts at the start state
    # activate the centre-distance barrier and give non-zero lagged normal forces
    sc = scenes.pack([scenes.micro_contact_scene(0)])
    o = O.OracleScene(sc)
    o.candidates(float(sc.dhat[0]))
    ints, dbl = o.narrow_phase(None)
    assert len(ints) > 0
    x = np.zeros(6 * o.n)
    for k in range(3):
        x[k::6] = sc.p[:, k]
    val, g, H, nf = o.barrier_eval(ints, dbl, x, 1e-3, 1e11)
    assert val > 0 and np.any(nf > 0)
    assert np.allclose(H, H.T)
    # the first step runs a stiff barrier solve (many Newton iterations)
    o.step(float(sc.dt[0]))
    assert o.stats()[1] > 4
