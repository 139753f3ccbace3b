"""Pins of the CPU oracle's narrow phase, forms, Newton / line-search / ALM loop and
helpers against independent computations (CPU only): closed forms, brute force with
different numerics, finite differences, numpy / scipy solvers and the reference's own
known-answer tests.

Reference tests restated (file:line in W-Moorer/NexDynIPC):
  tests/physics/test_friction.cpp:264-296 (tangent displacement follows the iterate),
  :328-386 (FrictionForm rotational gradient, central FD, tolerance 6e-2);
  tests/physics/test_ccd.cpp:163-210, :294-326 (AdaptiveBarrier growth and cap);
  tests/math/test_interval.cpp:52-336 (interval algebra of the operations the CCD uses);
  tests/geometry/test_bvh.cpp:112-275 (box queries, broad phase with inflation).
"""
import collections
import math
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scenes


# ------------------------------------------------------------------ segment closest points
def _seg_min_kkt(p1, q1, p2, q2):
    """Minimum of |p1 + s d1 - p2 - t d2| over [0,1]^2 by enumerating the KKT cases of the
    convex quadratic (interior stationary point, the four edges, the corners)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, b, e = d1 @ d1, d1 @ d2, d2 @ d2
    c, f = d1 @ r, d2 @ r

    def dist(s, t):
        return float(np.linalg.norm(p1 + s * d1 - p2 - t * d2))

    cands = [(0.0, 0.0), (0.0, 1.0), (1.0, 0.0), (1.0, 1.0)]
    det = a * e - b * b
    if det > 1e-14 * max(a * e, 1e-300):
        s = (b * f - c * e) / det
        t = (a * f - b * c) / det
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append((s, t))
    for s in (0.0, 1.0):  # t minimising for fixed s
        if e > 0:
            cands.append((s, min(1.0, max(0.0, (b * s + f) / e))))
    for t in (0.0, 1.0):
        if a > 0:
            cands.append((min(1.0, max(0.0, (b * t - c) / a)), t))
    return min(dist(s, t) for s, t in cands)


def _on_segment(c, p, q, tol=1e-12):
    d = q - p
    L2 = d @ d
    if L2 == 0:
        return np.linalg.norm(c - p) <= tol
    s = (c - p) @ d / L2
    return -1e-12 <= s <= 1 + 1e-12 and np.linalg.norm(p + s * d - c) <= tol * max(1.0, math.sqrt(L2))


def test_closest_points_random_pairs_reach_the_kkt_minimum():
    rng = np.random.default_rng(11)
    for _ in range(3000):
        p1, q1, p2, q2 = (rng.normal(0, 1, 3) for _ in range(4))
        if rng.random() < 0.2:  # nearly parallel pairs
            q2 = p2 + (q1 - p1) * rng.uniform(0.3, 2) + rng.normal(0, 1e-9, 3)
        c1, c2 = O.closest_points_segments(p1, q1, p2, q2)
        assert _on_segment(c1, p1, q1, 1e-11) and _on_segment(c2, p2, q2, 1e-11)
        got = np.linalg.norm(c1 - c2)
        ref = _seg_min_kkt(p1, q1, p2, q2)
        d1, d2 = q1 - p1, q2 - p2
        if abs((d1 @ d1) * (d2 @ d2) - (d1 @ d2) ** 2) > 1e-12:
            assert got <= ref + 1e-11 and got >= ref - 1e-11, (got, ref)
        else:  # the reference's parallel branch fixes s = 0 first: the minimum up to the
            # segments' tiny relative angle
            assert ref - 1e-11 <= got <= ref + 1e-6, (got, ref)


def test_closest_points_brute_force_grid():
    rng = np.random.default_rng(12)
    s = np.linspace(0, 1, 401)
    for _ in range(40):
        p1, q1, p2, q2 = (rng.normal(0, 1, 3) for _ in range(4))
        A = p1[None, :] + s[:, None] * (q1 - p1)[None, :]
        B = p2[None, :] + s[:, None] * (q2 - p2)[None, :]
        grid = np.min(np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2))
        c1, c2 = O.closest_points_segments(p1, q1, p2, q2)
        got = np.linalg.norm(c1 - c2)
        assert got <= grid + 1e-12 and grid - got <= 0.02  # the grid can only overshoot, by less than its step


def test_closest_points_closed_forms_and_degenerate_branches():
    g = 3.7e-4
    # crossed segments at a known gap
    c1, c2 = O.closest_points_segments(np.array([-1., 0, 0]), np.array([1., 0, 0]), np.array([0., g, -1]),
                                       np.array([0., g, 1]))
    assert np.allclose(c1, [0, 0, 0], atol=1e-15) and np.allclose(c2, [0, g, 0], atol=1e-15)
    # exactly parallel: the denominator branch takes s = 0, then t < 0 is clamped and s
    # recomputed (IPCSolver.cpp:69-80)
    c1, c2 = O.closest_points_segments(np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0.5, 1, 0]),
                                       np.array([2., 1, 0]))
    assert np.array_equal(c1, [0.5, 0, 0]) and np.array_equal(c2, [0.5, 1, 0])
    # both segments points
    a, b = np.array([1., 2, 3]), np.array([4., 5, 6])
    c1, c2 = O.closest_points_segments(a, a, b, b)
    assert np.array_equal(c1, a) and np.array_equal(c2, b)
    # first segment a point: projection onto the second, clamped
    c1, c2 = O.closest_points_segments(np.array([0.25, 1, 0]), np.array([0.25, 1, 0]), np.array([0., 0, 0]),
                                       np.array([1., 0, 0]))
    assert np.array_equal(c1, [0.25, 1, 0]) and np.allclose(c2, [0.25, 0, 0], atol=1e-16)
    c1, c2 = O.closest_points_segments(np.array([3., 1, 0]), np.array([3., 1, 0]), np.array([0., 0, 0]),
                                       np.array([1., 0, 0]))
    assert np.array_equal(c2, [1, 0, 0])
    # second segment a point
    c1, c2 = O.closest_points_segments(np.array([0., 0, 0]), np.array([2., 0, 0]), np.array([-1., 3, 0]),
                                       np.array([-1., 3, 0]))
    assert np.array_equal(c1, [0, 0, 0]) and np.array_equal(c2, [-1, 3, 0])


# ------------------------------------------------------------------ point in triangle
def _bary_signed_area(p, a, b, c):
    n = np.cross(b - a, c - a)
    A = n @ n
    l0 = np.cross(b - p, c - p) @ n / A
    l1 = np.cross(c - p, a - p) @ n / A
    l2 = np.cross(a - p, b - p) @ n / A
    return np.array([l0, l1, l2])


def test_point_in_triangle_against_signed_areas():
    rng = np.random.default_rng(13)
    checked = 0
    for _ in range(4000):
        a, b, c = (rng.normal(0, 1, 3) for _ in range(3))
        if np.linalg.norm(np.cross(b - a, c - a)) < 1e-3:
            continue
        lam = rng.uniform(-0.3, 1.3, 3)
        lam /= lam.sum()
        p = lam[0] * a + lam[1] * b + lam[2] * c
        ref = _bary_signed_area(p, a, b, c)
        m = ref.min()
        if -1e-7 < m + 1e-8 < 1e-7:  # within rounding of the reference's -1e-8 tolerance
            continue
        assert O.point_in_triangle(p, a, b, c) == (m >= -1e-8), (ref, m)
        checked += 1
    assert checked > 3000


def test_point_in_triangle_boundary_cases():
    a, b, c = np.array([0., 0, 0]), np.array([2., 0, 0]), np.array([0., 1, 0])
    for p in (a, b, c, (a + b) / 2, (b + c) / 2, (c + a) / 2, (a + b + c) / 3):
        assert O.point_in_triangle(p, a, b, c)
    # just outside each edge by more than the tolerance, just inside it by less
    assert not O.point_in_triangle(np.array([1., -1e-6, 0]), a, b, c)
    assert O.point_in_triangle(np.array([1., -1e-9, 0]), a, b, c)
    assert not O.point_in_triangle(np.array([-1e-6, 0.5, 0]), a, b, c)
    assert not O.point_in_triangle(np.array([1.0 + 1e-5, 0.5, 0]), a, b, c)
    # degenerate (collinear) triangle
    assert not O.point_in_triangle(np.array([0.5, 0, 0]), a, b, np.array([4., 0, 0]))


def test_safe_normalized_branches():
    assert np.allclose(O.safe_normalized([3., 4, 0], [0, 0, 1]), [0.6, 0.8, 0], atol=1e-16)
    assert np.allclose(O.safe_normalized([1e-13, 0, 0], [0, 0, 2]), [0, 0, 1])
    assert np.array_equal(O.safe_normalized([0., 0, 0], [0, 0, 0]), [0, 1, 0])


# ------------------------------------------------------------------ narrow phase, brute force
def _mesh_world(sc, b):
    vo, nv = sc.vert_off[b], sc.n_verts[b]
    fo, nf = sc.face_off[b], sc.n_faces[b]
    V = sc.verts.reshape(-1, 3)[vo:vo + nv]
    F = sc.faces.reshape(-1, 3)[fo:fo + nf]
    W = np.array([scenes.rotate(sc.q[b], v) for v in V]) + sc.p[b]
    E = []
    seen = set()
    for f in F:
        for u, w in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            k = (min(u, w), max(u, w))
            if k not in seen:
                seen.add(k)
                E.append(k)
    return W, F, E


def _brute_contacts(sc, bodies_pairs, dhat):
    """Every face-vertex and edge-edge pair of the given body pairs tested with different
    numerics (signed areas, KKT segment distance): multisets of (bodyA, bodyB, type,
    primA, primB) for clear hits, and the ambiguous ones (within rounding of a threshold)."""
    hits, amb = collections.Counter(), collections.Counter()
    dists = {}
    for A, B in bodies_pairs:
        WA, FA, EA = _mesh_world(sc, A)
        WB, FB, EB = _mesh_world(sc, B)
        for (fb, Ff, Wf, vb, Wv) in ((A, FA, WA, B, WB), (B, FB, WB, A, WA)):
            for fi, f in enumerate(Ff):
                v0, v1, v2 = Wf[f[0]], Wf[f[1]], Wf[f[2]]
                n = np.cross(v1 - v0, v2 - v0)
                n = n / np.linalg.norm(n)
                for vi, p in enumerate(Wv):
                    sd = (p - v0) @ n
                    if abs(sd) >= dhat + 1e-12:
                        continue
                    lam = _bary_signed_area(p - sd * n, v0, v1, v2)
                    key = (fb, vb, 0, fi, vi)
                    if abs(abs(sd) - dhat) < 1e-12 or abs(lam.min() + 1e-8) < 1e-9:
                        amb[key] += 1
                    elif lam.min() >= -1e-8:
                        hits[key] += 1
                        dists.setdefault(key, []).append(abs(sd))
        for ea in EA:
            for eb in EB:
                d = _seg_min_kkt(WA[ea[0]], WA[ea[1]], WB[eb[0]], WB[eb[1]])
                key = (A, B, 1, ea[0], eb[0])
                if abs(d - dhat) < 1e-12:
                    amb[key] += 1
                elif d < dhat:
                    hits[key] += 1
                    dists.setdefault(key, []).append(d)
    return hits, amb, dists


def _cube_on_ground(gap, q=None, xz=(0.3, -0.2)):
    q = scenes._IDENT if q is None else q
    ground = scenes._box_body("ground", 10, 1, 10, [0, -0.5, 0], static=True)
    # lowest point of the (possibly tilted) cube at height gap above the ground's top
    V, _ = scenes.box_mesh(1, 1, 1)
    low = min(scenes.rotate(q, v)[1] for v in V)
    cube = scenes._box_body("cube", 1, 1, 1, [xz[0], gap - low, xz[1]], q)
    return scenes.pack([scenes.SceneSpec([ground, cube], dt=1e-2, dhat=1e-3, mu=0.3)])


@pytest.mark.parametrize("gap", [1e-4, 4.4e-4, 9.5e-4])
def test_narrow_phase_cube_at_known_gap(gap):
    """An axis-aligned unit cube centred at a known gap above a static box: the reference's
    face-vertex test finds the cube's 4 bottom vertices over the box's top face, the two
    on the face's diagonal once per triangle (6 contacts), each at distance = gap with
    normal +Y, point = midpoint of the vertex and its projection. Edge-edge contacts and
    the whole set are checked against brute force with different numerics."""
    sc = _cube_on_ground(gap, xz=(0.0, 0.0))
    o = O.OracleScene(sc)
    o.candidates(1e-3)
    ints, dbl = o.narrow_phase(None)
    fv = ints[:, 2] == 0
    assert int(fv.sum()) == 6
    assert np.all(ints[fv, 0] == 0) and np.all(ints[fv, 1] == 1)
    assert np.max(np.abs(dbl[fv, 6] - gap)) <= 1e-15
    assert np.array_equal(dbl[fv, 3:6], np.tile([0.0, 1.0, 0.0], (6, 1)))
    assert np.max(np.abs(dbl[fv, 1] - gap / 2)) <= 1e-15
    hits, amb, dists = _brute_contacts(sc, [(0, 1)], 1e-3)
    got = collections.Counter(tuple(int(v) for v in r) for r in ints)
    assert not amb
    assert got == hits
    for r, d in zip(ints, dbl):
        key = tuple(int(v) for v in r)
        assert min(abs(d[6] - x) for x in dists[key]) <= 1e-12
        # normal oriented toward p_B - p_A, unit length
        assert abs(np.linalg.norm(d[3:6]) - 1) <= 1e-14
        assert d[3:6] @ (sc.p[r[1]] - sc.p[r[0]]) >= 0


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_narrow_phase_tilted_cube_matches_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    q = scenes.random_tilt(rng, 25)
    sc = _cube_on_ground(rng.uniform(1e-4, 9e-4), q)
    o = O.OracleScene(sc)
    o.candidates(1e-3)
    ints, dbl = o.narrow_phase(None)
    hits, amb, dists = _brute_contacts(sc, [(0, 1)], 1e-3)
    got = collections.Counter(tuple(int(v) for v in r) for r in ints)
    for k in amb:
        got.pop(k, None)
    assert got == hits and len(ints) > 0
    for r, d in zip(ints, dbl):
        key = tuple(int(v) for v in r)
        if key in dists:
            assert min(abs(d[6] - x) for x in dists[key]) <= 1e-12


def _overlapping_cubes(n=10, seed=1):
    rng = np.random.default_rng(seed)
    bodies = [scenes._box_body("ground", 10, 1, 10, [0, -0.5, 0], static=True)]
    for _ in range(n):
        bodies.append(scenes._box_body("cube", 1, 1, 1, [rng.uniform(-1e-4, 1e-4), 0.5 + rng.uniform(-1e-4, 1e-4),
                                                          rng.uniform(-1e-4, 1e-4)]))
    return scenes.pack([scenes.SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.3)])


def test_contact_cap_keeps_the_2000_closest():
    """IPCSolver.cpp:571-577: above 2000 contacts the list is sorted by distance and cut.
    A stack of coincident cubes gives 7420 raw contacts; the kept ones are the 2000 of
    smallest distance (as a multiset; which of several equal distances at the cut is kept
    is the sort's tie order, not pinned)."""
    sc = _overlapping_cubes()
    o = O.OracleScene(sc)
    o.candidates(1e-3)
    ints, dbl = o.narrow_phase(None)
    o.set_narrow_raw(True)
    rints, rdbl = o.narrow_phase(None)
    assert len(rints) > 2000 and len(ints) == 2000
    assert np.array_equal(np.sort(dbl[:, 6]), np.sort(rdbl[:, 6])[:2000])
    cut = np.max(dbl[:, 6])
    below = collections.Counter(tuple(r) for r, d in zip(rints.tolist(), rdbl) if d[6] < cut)
    kept = collections.Counter(tuple(r) for r, d in zip(ints.tolist(), dbl) if d[6] < cut)
    assert below == kept
    # the raw list itself: candidate order is FV first, then EE
    types = rints[:, 2]
    assert np.all(np.diff(types) >= 0)


# ------------------------------------------------------------------ forms: finite differences
def _micro_contacts(seed=0):
    sc = scenes.pack([scenes.micro_contact_scene(seed, n_cubes=4, tilt_deg=8.0)])
    o = O.OracleScene(sc)
    o.candidates(float(sc.dhat[0]))
    ints, dbl = o.narrow_phase(None)
    x = np.zeros(6 * o.n)
    for k in range(3):
        x[k::6] = sc.p[:, k]
    return sc, o, ints, dbl, x


def test_barrier_form_value_gradient_hessian_by_central_differences():
    sc, o, ints, dbl, x = _micro_contacts(0)
    dh, kap = 1e-3, 1e11
    v0, g0, H0, nf = o.barrier_eval(ints, dbl, x, dh, kap)
    assert v0 > 0 and np.any(nf > 0)
    rng = np.random.default_rng(3)
    x = x + rng.normal(0, 2e-6, len(x)) * np.tile([1, 1, 1, 0, 0, 0], o.n)
    v0, g0, H0, _ = o.barrier_eval(ints, dbl, x, dh, kap)
    for i in range(len(x)):
        h = 1e-9
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        vp, gp, _, _ = o.barrier_eval(ints, dbl, xp, dh, kap, hess=False)
        vm, gm, _, _ = o.barrier_eval(ints, dbl, xm, dh, kap, hess=False)
        fd = (vp - vm) / (2 * h)
        if i % 6 >= 3:  # the centre-distance barrier does not depend on rotations
            assert g0[i] == 0.0 and vp == vm and np.all(H0[i] == 0.0)
            continue
        assert abs(g0[i] - fd) <= 1e-6 * max(abs(g0[i]), 1e-3 * np.max(np.abs(g0)))
        hcol = (gp - gm) / (2 * h)
        sc_ = max(np.max(np.abs(H0[:, i])), 1e-3 * np.max(np.abs(H0)))
        assert np.max(np.abs(H0[:, i] - hcol)) <= 1e-5 * sc_
    assert np.array_equal(H0, H0.T)
    # b'' > 0 on (0, dhat): the Hessian is positive semidefinite
    assert np.linalg.eigvalsh(H0).min() >= -1e-6 * np.max(np.abs(H0))


def test_friction_form_translation_gradient_and_hessian_by_central_differences():
    sc, o, ints, dbl, x = _micro_contacts(1)
    _, _, _, nf = o.barrier_eval(ints, dbl, x, 1e-3, 1e11, hess=False)
    assert np.any(nf > 0)
    dt, mu, eps = 1e-3, 0.5, 1e-8
    rng = np.random.default_rng(4)
    x = x + rng.normal(0, 1e-6, len(x)) * np.tile([1, 1, 1, 0, 0, 0], o.n)
    v0, g0, H0 = o.friction_eval(ints, dbl, nf, x, mu, eps, dt)
    assert v0 > 0
    for i in range(len(x)):
        if i % 6 >= 3:
            continue
        h = 1e-10
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        vp, gp, _ = o.friction_eval(ints, dbl, nf, xp, mu, eps, dt, hess=False)
        vm, gm, _ = o.friction_eval(ints, dbl, nf, xm, mu, eps, dt, hess=False)
        assert abs(g0[i] - (vp - vm) / (2 * h)) <= 1e-6 * max(abs(g0[i]), 1e-3 * np.max(np.abs(g0)))
        hc = (gp - gm) / (2 * h)
        tr = np.tile([1, 1, 1, 0, 0, 0], o.n).astype(bool)  # translation rows
        assert np.max(np.abs(H0[tr, i] - hc[tr])) <= 1e-5 * np.max(np.abs(H0[tr][:, tr]))
    assert np.allclose(H0, H0.T, rtol=0, atol=1e-12 * np.max(np.abs(H0)))


def _two_body_scene(p0, q0, p1, q1, v0=(0, 0, 0)):
    b0 = scenes._box_body("a", 0.2, 0.2, 0.2, p0, q0, v=np.array(v0, dtype=np.float64))
    b1 = scenes._box_body("b", 0.2, 0.2, 0.2, p1, q1)
    return scenes.pack([scenes.SceneSpec([b0, b1], dt=1e-2, dhat=1e-3, mu=0.5)])


def test_friction_tangent_displacement_follows_the_iterate():
    """test_friction.cpp:264-296: world velocity 100 m/s on body 0 does not enter; at the
    step-start positions the friction energy is 0, after moving body 0 by 2 cm it is > 0."""
    sc = _two_body_scene([0, 0, 0], scenes._IDENT, [1, 0, 0], scenes._IDENT, v0=(100.0, 0, 0))
    o = O.OracleScene(sc)
    ints = np.array([[0, 1, 0, 0, 0]], dtype=np.int32)
    dbl = np.array([[0, 0, 0, 0, 1, 0, 0]], dtype=np.float64)
    x = np.zeros(12)
    x[0:3] = [0, 0, 0]
    x[6:9] = [1, 0, 0]
    v, _, _ = o.friction_eval(ints, dbl, np.array([5.0]), x, 0.5, 1e-8, 0.01, hess=False)
    assert abs(v) <= 1e-12
    x[0] = 0.02
    v, _, _ = o.friction_eval(ints, dbl, np.array([5.0]), x, 0.5, 1e-8, 0.01, hess=False)
    assert v > 0.0


def test_friction_form_rotational_gradient_kat():
    """test_friction.cpp:328-386: bodies rotated 0.4 rad about Y and -0.2 rad about X,
    contact point (0, 0, 1), normal +Y, normal force 5, mu 0.5, eps 1e-4, h 0.01;
    gradient entries 4 and 10 within 6e-2 of central differences (h = 1e-6). The
    reference's lever arm ignores its own rotation dependence, hence the tolerance."""
    qa = scenes.quat_from_axis_angle([0, 1, 0], 0.4)
    qb = scenes.quat_from_axis_angle([1, 0, 0], -0.2)
    sc = _two_body_scene([0, 0, 0], qa, [1, 0, 0], qb)
    o = O.OracleScene(sc)
    ints = np.array([[0, 1, 0, 0, 0]], dtype=np.int32)
    dbl = np.array([[0, 0, 1, 0, 1, 0, 0]], dtype=np.float64)
    nf = np.array([5.0])
    x = np.zeros(12)
    x[6:9] = [1, 0, 0]
    x[4] = 0.005
    x[10] = -0.003
    _, g, _ = o.friction_eval(ints, dbl, nf, x, 0.5, 1e-4, 0.01, hess=False)
    for idx in (4, 10):
        xp, xm = x.copy(), x.copy()
        xp[idx] += 1e-6
        xm[idx] -= 1e-6
        vp = o.friction_eval(ints, dbl, nf, xp, 0.5, 1e-4, 0.01, hess=False)[0]
        vm = o.friction_eval(ints, dbl, nf, xm, 0.5, 1e-4, 0.01, hess=False)[0]
        assert abs(g[idx] - (vp - vm) / 2e-6) <= 6e-2
    # the translational entries are exact
    for idx in (0, 2, 6, 8):
        xp, xm = x.copy(), x.copy()
        xp[idx] += 1e-7
        xm[idx] -= 1e-7
        vp = o.friction_eval(ints, dbl, nf, xp, 0.5, 1e-4, 0.01, hess=False)[0]
        vm = o.friction_eval(ints, dbl, nf, xm, 0.5, 1e-4, 0.01, hess=False)[0]
        assert abs(g[idx] - (vp - vm) / 2e-7) <= 1e-6 * max(1.0, abs(g[idx]))


# ------------------------------------------------------------------ adaptive barrier
def test_adaptive_barrier_kats():
    """test_ccd.cpp:163-210 and :294-326."""
    k, kmin = O.adaptive_barrier(1.0, 10.0, [])
    assert kmin > 0 and k[0] == kmin == 1e11 * 1.0 / 100.0
    k, _ = O.adaptive_barrier(1.0, 10.0, [1e-12])
    assert k[1] > k[0]
    k, _ = O.adaptive_barrier(1.0, 1.0, [1e-10, 1e-10])
    assert k[2] > k[1] > k[0] and k[1] == 2 * k[0]
    k, _ = O.adaptive_barrier(1.0, 1.0, [1e-15] * 100, max_stiffness=1e15)
    assert np.all(k <= 1e15) and k[-1] == 1e15
    # no growth at or above dhat_epsilon = 1e-9
    k, _ = O.adaptive_barrier(2.0, 3.0, [1e-9, 5e-9, 1.0])
    assert np.all(k == k[0])


def test_scene_min_distance_pruned_equals_brute_force():
    """The AABB-pruned computeSceneMinDistance (an oracle option for large scenes) gives
    the unpruned result's comparison with 1e-9, and its value whenever that is below the
    prune margin."""
    cases = []
    sc = scenes.make_config("c1")
    sc.q[1:3] = [1, 0, 0, 0]
    sc.p[1] = [3.0, 2.0, 3.0]
    for dy in (0.0, 5e-10, 2e-9, 1e-6, 3e-3):
        s2 = scenes.make_config("c1")
        s2.q[1:3] = [1, 0, 0, 0]
        s2.p[1] = [3.0, 2.0 + dy, 3.0]
        s2.p[2] = [3.0, 1.0, 3.0]
        cases.append(s2)
    cases.append(_overlapping_cubes(4, seed=2))
    cases.append(_cube_on_ground(0.0))
    for s in cases:
        a, b = O.OracleScene(s), O.OracleScene(s)
        b.set_scene_min_pruned(True)
        ma, mb = a.scene_min_distance(), b.scene_min_distance()
        assert (ma < 1e-9) == (mb < 1e-9)
        if ma < 1e-8:
            assert ma == mb


# ------------------------------------------------------------------ interval algebra
def test_interval_kats():
    """test_interval.cpp:52-336, the operations the CCD uses."""
    i, j = (0.0, 1.0), (4.0, 5.0)
    assert O.interval_op(0, i, j).tolist() == [4.0, 6.0]
    assert O.interval_op(0, i, (10.0, 10.0)).tolist() == [10.0, 11.0]
    assert O.interval_op(1, i, j).tolist() == [-5.0, -3.0]
    assert O.interval_op(1, j, i).tolist() == [3.0, 5.0]
    assert O.interval_op(2, i, j).tolist() == [0.0, 5.0]
    assert O.interval_op(2, (-1.0, 1.0), (2.0, 3.0)).tolist() == [-3.0, 3.0]
    assert O.interval_op(3, (1.0, 5.0))[0] == 4.0
    assert O.interval_op(4, (1.0, 5.0))[0] == 3.0
    assert O.interval_op(5, (1.0, 5.0))[0] == 0.0
    assert O.interval_op(5, (-1.0, 1.0))[0] == 1.0
    assert O.interval_op(5, (0.0, 5.0))[0] == 1.0


def test_interval_products_are_the_tight_outward_hull():
    """Every sign case of the product table: the result is the exact hull of the four
    endpoint products, rounded outward by at most one ulp."""
    rng = np.random.default_rng(21)
    for _ in range(3000):
        a = np.sort(rng.normal(0, 1, 2) * np.exp(rng.uniform(-5, 5)))
        b = np.sort(rng.normal(0, 1, 2) * np.exp(rng.uniform(-5, 5)))
        if rng.random() < 0.1:  # an endpoint at zero
            a[0 if a[1] >= 0 else 1] = 0.0
        r = O.interval_op(2, a, b)
        prods = [Fraction(float(x)) * Fraction(float(y)) for x in a for y in b]
        lo, hi = min(prods), max(prods)
        assert Fraction(r[0]) <= lo and hi <= Fraction(r[1])
        assert r[0] == float(lo) or np.nextafter(r[0], np.inf) >= float(lo)
        assert r[1] == float(hi) or np.nextafter(r[1], -np.inf) <= float(hi)
        s = O.interval_op(0, a, b)
        assert Fraction(s[0]) <= Fraction(a[0]) + Fraction(b[0]) <= Fraction(s[0]) + Fraction(abs(s[0]) + 1e-300) * 2 ** -52 * 2
        d = O.interval_op(1, a, b)
        assert Fraction(d[0]) <= Fraction(a[0]) - Fraction(b[1]) and Fraction(a[1]) - Fraction(b[0]) <= Fraction(d[1])


# ------------------------------------------------------------------ BVH
BOXES4 = [[0, 0, 0, 1, 1, 1], [2, 0, 0, 3, 1, 1], [0, 2, 0, 1, 3, 1], [2, 2, 0, 3, 3, 1]]


def test_bvh_box_query_kats():
    """test_bvh.cpp:112-157."""
    assert O.bvh_query(BOXES4, [0.5, 0.5, 0], [0.6, 0.6, 1]) == [0]
    assert sorted(O.bvh_query(BOXES4, [0, 0, 0], [3, 3, 1])) == [0, 1, 2, 3]
    assert O.bvh_query(BOXES4, [10, 10, 10], [11, 11, 11]) == []
    assert sorted(O.bvh_query(BOXES4, [0.5, 0.5, 0], [2.5, 0.6, 1])) == [0, 1]
    # test_bvh.cpp:159-191, points as degenerate boxes
    two = BOXES4[:2]
    assert O.bvh_query(two, [0.5] * 3, [0.5] * 3) == [0]
    assert O.bvh_query(two, [2.5, 0.5, 0.5], [2.5, 0.5, 0.5]) == [1]
    assert O.bvh_query(two, [5] * 3, [5] * 3) == []


def test_broad_phase_kats():
    """test_bvh.cpp:230-275: overlapping pair found; inflation 0.2 joins boxes 0.1 apart."""
    assert O.broad_phase_boxes([[0, 0, 0, 1, 1, 1], [0.5, 0.5, 0.5, 1.5, 1.5, 1.5], [5, 5, 5, 6, 6, 6]]) == [(0, 1)]
    b = [[0, 0, 0, 1, 1, 1], [1.1, 1.1, 1.1, 2, 2, 2]]
    assert O.broad_phase_boxes(b) == []
    assert O.broad_phase_boxes(b, 0.2) == [(0, 1)]


@pytest.mark.parametrize("seed", [0, 1])
def test_bvh_query_equals_brute_force_on_random_boxes(seed):
    rng = np.random.default_rng(50 + seed)
    n = 300
    c = rng.uniform(-5, 5, (n, 3))
    e = rng.uniform(0.05, 1.0, (n, 3))
    boxes = np.concatenate([c - e, c + e], axis=1)
    for _ in range(50):
        qc, qe = rng.uniform(-5, 5, 3), rng.uniform(0.1, 2.0, 3)
        lo, hi = qc - qe, qc + qe
        ref = [i for i in range(n) if np.all(boxes[i, :3] <= hi) and np.all(boxes[i, 3:] >= lo)]
        assert sorted(O.bvh_query(boxes, lo, hi)) == ref


# ------------------------------------------------------------------ Newton, line search, ALM
def _solves(o, n_steps):
    out = []
    for st in range(int(o.record_info()[0])):
        for so in range(int(o.record_info(st)[0])):
            out.append((st, so))
    return out


@pytest.mark.parametrize("name", ["micro", "chain"])
def test_newton_line_search_and_alm_pins(name):
    """Recorded Newton solves of contact-active (micro cubes, barrier and friction) and
    hinge (free chain) steps, checked against independent numerics:
      - every direction solves H dx = -g to a componentwise backward error of 1e-12
        (pins the LDL^T);
      - the trials are alpha = 1, 1/2, ...; every rejected trial fails and the accepted one
        passes the Armijo test with c1 = 1e-4 plus the contact-merit test
        (LineSearch.cpp:6-58), their values being the objective at x + alpha dx;
      - the accepted iterate is the next iterate, and the last one is x*;
      - a converged solve has |g(x*)|_inf < 1e-6, a positive semidefinite Hessian and no
        lower value at random nearby points (a local minimum: with the clamped barrier
        distance the objective need not be convex, and the stop test is absolute,
        NewtonSolver.cpp:10-95);
      - ALM: mu hardens x2 where |C|_inf > 2e-4 and lambda += mu_post C(x*)
        (IPCSolver.cpp:1084-1110)."""
    if name == "micro":
        sc = scenes.pack([scenes.micro_contact_scene(2, n_cubes=4)])
    else:
        sc = scenes.hinge_chain_scene(n_links=3)
    o = O.OracleScene(sc)
    o.record_steps(float(sc.dt[0]), 3)
    solves = _solves(o, 3)
    assert len(solves) >= 3
    n_iter = n_conv = 0
    prev = None
    for st, so in solves:
        S = o.record_solve(st, so)
        x = S["x_hat"] if so == 0 else prev["x_star"]
        for it in range(S["n_iters"]):
            R = o.record_iteration(st, so, it)
            assert np.array_equal(R["x"], x)
            ob = o.record_objective(st, so, x, hess=True)
            g, H = ob["g"], ob["H"]
            # componentwise backward error of the solve (H spans ~1e-6 (mm-cube masses)
            # to ~1e7 (barrier curvature), so forward errors say little)
            res = np.abs(H @ R["dx"] + g)
            scale = np.abs(H) @ np.abs(R["dx"]) + np.abs(g)
            assert np.all(res <= 1e-12 * scale + 1e-300), np.max(res / (scale + 1e-300))
            assert R["fx"] == ob["f"] and R["fc"] == ob["fc"]
            assert abs(R["descent"] - g @ R["dx"]) <= 1e-12 * (np.abs(g) @ np.abs(R["dx"]))
            tr = R["trials"]
            if R["descent"] > 0:
                assert len(tr) == 0
                break
            alpha = 1.0
            accepted = None
            for k, (a, fn, fcn) in enumerate(tr):
                assert a == alpha
                xn = x + a * R["dx"]
                ot = o.record_objective(st, so, xn)
                assert fn == ot["f"] and fcn == ot["fc"]
                cok = (fcn <= R["fc"] + 1e-4 * a * R["cdesc"]) if R["cdesc"] <= 0 else (fcn <= R["fc"])
                ok = fn <= R["fx"] + 1e-4 * a * R["descent"] and cok
                if k < len(tr) - 1:
                    assert not ok
                else:
                    accepted = ok
                alpha *= 0.5
            assert len(tr) <= 20
            if not accepted:
                break
            x = x + tr[-1, 0] * R["dx"]
            n_iter += 1
        if S["converged"]:
            assert np.array_equal(x, S["x_star"])
        xs = S["x_star"]
        ob = o.record_objective(st, so, xs)
        if S["converged"]:
            n_conv += 1
            assert np.max(np.abs(ob["g"])) < 1e-6
            rng = np.random.default_rng(st * 31 + so)
            for _ in range(20):
                d = rng.normal(0, 1, len(xs)) * 1e-7 * (1 + np.abs(xs))
                assert o.record_objective(st, so, xs + d)["f"] >= ob["f"] - 1e-12 * abs(ob["f"])
            # second-order condition: the Hessian at x* is positive semidefinite
            Hs = o.record_objective(st, so, xs, hess=True)["H"]
            assert np.linalg.eigvalsh(0.5 * (Hs + Hs.T)).min() >= -1e-9 * np.max(np.abs(Hs))
        # ALM bookkeeping
        if len(S["mu"]) and sc.n_joints:
            Cx = ob["C"]
            infn = np.max(np.abs(Cx), axis=1)
            expect_mu = np.where(infn > 2.0 * 1e-4, np.minimum(S["mu"] * 2.0, 1e12), S["mu"])
            assert np.array_equal(S["mu_post"], expect_mu)
            nxt = (st, so + 1) if (st, so + 1) in solves else None
            if nxt is not None:
                N = o.record_solve(*nxt)
                assert np.allclose(N["lam"], S["lam"] + S["mu_post"][:, None] * Cx, rtol=0,
                                   atol=1e-12 * max(1.0, np.max(np.abs(N["lam"]))))
        prev = S
    assert n_iter > 0 and n_conv > 0
