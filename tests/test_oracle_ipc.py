"""Pins of the oracle's `ipc` mode (the primitive-distance barrier of rigid IPC,
nexdynipc_amd/csrc/ipc_contact.h, included by the oracle), each against something the
oracle does not compute the same way:

  - PT / EE distances: the textbook characterisation (the closest point of a triangle is
    the plane projection when it falls inside, else the nearest of the three clamped
    segment points; two segments: the interior line-line solution when both parameters
    fall inside, else the nearest of the four endpoint-to-segment distances) and a dense
    sample of the primitives as an upper bound;
  - the distance type: the squared distance of the reported type equals the true one;
  - E = kappa b(d) against -kappa (d - dhat)^2 ln(d / dhat) (Barrier.cpp:11-46), its
    gradient by central differences of E, its Hessian by central differences of the
    gradient;
  - the PSD projection against numpy's eigh with eigenvalues clamped at 0;
  - the rigid map's world point against p + Rodrigues(theta) R(q) r, its theta Jacobian
    against -[R(q) r]x at theta = 0 and central differences elsewhere;
  - the per-segment body gradient against central differences of the total energy in x,
    the translation blocks of the body Hessian (where the map is linear, so the Gauss-
    Newton fold is exact) against central differences of the gradient, the rotation block
    against J^T H J with J taken by central differences of the map;
  - the segment grouping of consecutive pairs of one unordered body pair.
"""
import math

import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scenes

# ------------------------------------------------------------------ independent geometry
def _seg_point(p, a, b):
    ab = b - a
    L = ab @ ab
    t = 0.0 if L == 0.0 else min(max((p - a) @ ab / L, 0.0), 1.0)
    return a + t * ab, t


def _pt_ref(x):
    p, a, b, c = x[:3], x[3:6], x[6:9], x[9:12]
    n = np.cross(b - a, c - a)
    cands = []
    h = (p - a) @ n / (n @ n)
    proj = p - h * n
    # barycentric coordinates of the projection by areas
    A = n @ n
    u = np.cross(c - b, proj - b) @ n / A
    v = np.cross(a - c, proj - c) @ n / A
    w = 1.0 - u - v
    if u >= 0 and v >= 0 and w >= 0:
        cands.append(np.linalg.norm(p - proj))
    for e0, e1 in ((a, b), (b, c), (c, a)):
        q, _ = _seg_point(p, e0, e1)
        cands.append(np.linalg.norm(p - q))
    return min(cands)


def _ee_ref(x):
    a0, a1, b0, b1 = x[:3], x[3:6], x[6:9], x[9:12]
    d1, d2, r = a1 - a0, b1 - b0, a0 - b0
    M = np.array([[d1 @ d1, -(d1 @ d2)], [-(d1 @ d2), d2 @ d2]])
    cands = []
    if abs(np.linalg.det(M)) > 1e-12 * M[0, 0] * M[1, 1]:
        s, t = np.linalg.solve(M, [-(d1 @ r), d2 @ r])
        if 0 <= s <= 1 and 0 <= t <= 1:
            cands.append(np.linalg.norm(a0 + s * d1 - b0 - t * d2))
    for p, e0, e1 in ((a0, b0, b1), (a1, b0, b1), (b0, a0, a1), (b1, a0, a1)):
        q, _ = _seg_point(p, e0, e1)
        cands.append(np.linalg.norm(p - q))
    return min(cands)


def _sampled_min(kind, x, n=60):
    u = np.linspace(0.0, 1.0, n + 1)
    if kind == 0:
        p, a, b, c = x[:3], x[3:6], x[6:9], x[9:12]
        U, V = np.meshgrid(u, u, indexing="ij")
        m = (U + V) <= 1.0
        pts = a + U[m][:, None] * (b - a) + V[m][:, None] * (c - a)
        return np.min(np.linalg.norm(pts - p, axis=1))
    a0, a1, b0, b1 = x[:3], x[3:6], x[6:9], x[9:12]
    P = a0 + u[:, None] * (a1 - a0)
    Q = b0 + u[:, None] * (b1 - b0)
    return np.min(np.linalg.norm(P[:, None, :] - Q[None, :, :], axis=2))


def _random_pair(kind, rng, scale=1.0):
    x = rng.normal(size=12) * scale
    if kind == 0:  # keep the triangle well shaped
        x[3:] = np.concatenate([[0, 0, 0], [1, 0.1, 0], [0.2, 1, 0]]) * scale + rng.normal(size=9) * 0.2 * scale
    return x


# the squared distance of each type, written out from the vertex order of the types
def _type_d2(kind, t, x):
    v = x.reshape(4, 3)

    def pp(a, b):
        return float((a - b) @ (a - b))

    def pe(p, e0, e1):
        e = e1 - e0
        c = np.cross(e, p - e0)
        return float(c @ c / (e @ e))

    def pl(p, a, b, c):
        n = np.cross(b - a, c - a)
        return float(((p - a) @ n) ** 2 / (n @ n))

    def ll(a0, a1, b0, b1):
        n = np.cross(a1 - a0, b1 - b0)
        return float(((b0 - a0) @ n) ** 2 / (n @ n))

    if kind == 0:
        return [lambda: pp(v[0], v[1]), lambda: pp(v[0], v[2]), lambda: pp(v[0], v[3]),
                lambda: pe(v[0], v[1], v[2]), lambda: pe(v[0], v[2], v[3]), lambda: pe(v[0], v[3], v[1]),
                lambda: pl(*v)][t]()
    return [lambda: pp(v[0], v[2]), lambda: pp(v[0], v[3]), lambda: pp(v[1], v[2]), lambda: pp(v[1], v[3]),
            lambda: pe(v[0], v[2], v[3]), lambda: pe(v[1], v[2], v[3]), lambda: pe(v[2], v[0], v[1]),
            lambda: pe(v[3], v[0], v[1]), lambda: ll(*v)][t]()


@pytest.mark.parametrize("kind", [0, 1])
def test_distance_matches_textbook_and_samples(kind):
    rng = np.random.default_rng(1000 + kind)
    types = set()
    for _ in range(400):
        x = _random_pair(kind, rng)
        d, t = O.ipc_distance(kind, x)
        ref = _pt_ref(x) if kind == 0 else _ee_ref(x)
        assert abs(d - ref) <= 1e-12 * max(1.0, ref), (x.tolist(), d, ref, t)
        assert d <= _sampled_min(kind, x) + 1e-12
        # the squared distance of the reported type is the distance
        assert abs(math.sqrt(max(_type_d2(kind, t, x), 0.0)) - ref) <= 1e-10 * max(1.0, ref), t
        types.add(t)
    assert types == set(range(7 if kind == 0 else 9)), types


def test_distance_special_cases():
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], dtype=float)
    cases = [  # (point, expected distance, expected type)
        ([0.25, 0.25, 0.3], 0.3, 6), ([-1, -1, 0.0], math.sqrt(2), 0), ([2, 0, 0], 1.0, 1), ([0, 4, 4], 5.0, 2),
        ([0.5, -2, 0], 2.0, 3), ([1, 1, 0], math.sqrt(0.5), 4), ([-3, 0.5, 4], 5.0, 5)]
    for p, dref, tref in cases:
        d, t = O.ipc_distance(0, np.concatenate([p, tri]))
        assert t == tref and abs(d - dref) < 1e-15 * max(1, dref), (p, d, t)
    # perpendicular crossing segments: the line-line case
    d, t = O.ipc_distance(1, np.array([-1, 0, 0, 1, 0, 0, 0, -1, 0.5, 0, 1, 0.5]))
    assert t == 8 and d == 0.5
    # parallel segments, overlapping: s = 0 first (as closestPointsOnSegments takes it), t
    # clamps to 0 below the range, s follows to b0's projection: the b0-to-segment-a case
    d, t = O.ipc_distance(1, np.array([0, 0, 0, 2, 0, 0, 1, 0.3, 0, 3, 0.3, 0]))
    assert abs(d - 0.3) < 1e-15 and t == 6
    # endpoint to endpoint
    d, t = O.ipc_distance(1, np.array([0, 0, 0, 1, 0, 0, 2, 1, 0, 3, 1, 0]))
    assert t == 2 and abs(d - math.sqrt(2)) < 1e-15


# ------------------------------------------------------------------ energy, gradient, Hessian
def _barrier_ref(d, dhat, kappa):
    return -kappa * (d - dhat) ** 2 * math.log(d / dhat) if d < dhat else 0.0


def _near_pair(kind, rng, dhat):
    """A pair at distance in (0.2, 0.8) dhat whose type does not change within 1e-6 dhat."""
    while True:
        x = _random_pair(kind, rng)
        d, t = O.ipc_distance(kind, x)
        if d < 1e-3:
            continue
        # move the first vertex (PT) / the second edge (EE) along the closest direction
        target = rng.uniform(0.2, 0.8) * dhat
        xs = x.copy()
        if kind == 0:
            c = _closest_on_tri(x)
            n = (x[:3] - c) / np.linalg.norm(x[:3] - c)
            xs[:3] = c + target * n
        else:
            pa, pb = _closest_ee(x)
            n = (pb - pa) / np.linalg.norm(pb - pa)
            shift = pa + target * n - pb
            xs[6:] = xs[6:] + np.tile(shift, 2)
        d2, t2 = O.ipc_distance(kind, xs)
        if t2 != t or abs(d2 - target) > 1e-12:
            continue
        ok = True
        for _ in range(12):
            dx = rng.normal(size=12) * 1e-6 * dhat
            if O.ipc_distance(kind, xs + dx)[1] != t:
                ok = False
                break
        if ok:
            return xs, t


def _closest_on_tri(x):
    p, a, b, c = x[:3], x[3:6], x[6:9], x[9:12]
    best = None
    for e0, e1 in ((a, b), (b, c), (c, a)):
        q, _ = _seg_point(p, e0, e1)
        if best is None or np.linalg.norm(p - q) < np.linalg.norm(p - best):
            best = q
    n = np.cross(b - a, c - a)
    proj = p - ((p - a) @ n) / (n @ n) * n
    A = n @ n
    u = np.cross(c - b, proj - b) @ n / A
    v = np.cross(a - c, proj - c) @ n / A
    if u >= 0 and v >= 0 and 1 - u - v >= 0 and np.linalg.norm(p - proj) < np.linalg.norm(p - best):
        best = proj
    return best


def _closest_ee(x):
    a0, a1, b0, b1 = x[:3], x[3:6], x[6:9], x[9:12]
    d1, d2, r = a1 - a0, b1 - b0, a0 - b0
    M = np.array([[d1 @ d1, -(d1 @ d2)], [-(d1 @ d2), d2 @ d2]])
    best = None
    s, t = np.linalg.solve(M, [-(d1 @ r), d2 @ r])
    if 0 <= s <= 1 and 0 <= t <= 1:
        best = (a0 + s * d1, b0 + t * d2)
    for p, e0, e1, first in ((a0, b0, b1, True), (a1, b0, b1, True), (b0, a0, a1, False), (b1, a0, a1, False)):
        q, _ = _seg_point(p, e0, e1)
        pair = (p, q) if first else (q, p)
        if best is None or np.linalg.norm(pair[0] - pair[1]) < np.linalg.norm(best[0] - best[1]):
            best = pair
    return best


@pytest.mark.parametrize("kind", [0, 1])
def test_energy_gradient_hessian_by_differences(kind):
    rng = np.random.default_rng(2000 + kind)
    dhat, kappa = 0.05, 3.0e3
    seen = set()
    for _ in range(24):
        x, t = _near_pair(kind, rng, dhat)
        seen.add(t)
        E, g, H, t0, d = O.ipc_pair(kind, x, dhat, kappa)
        assert t0 == t
        assert abs(E - _barrier_ref(d, dhat, kappa)) <= 1e-13 * abs(E)
        h = 1e-7 * dhat
        gfd = np.zeros(12)
        Hfd = np.zeros((12, 12))
        for i in range(12):
            e = np.zeros(12)
            e[i] = h
            Ep, gp = O.ipc_pair(kind, x + e, dhat, kappa)[:2]
            Em, gm = O.ipc_pair(kind, x - e, dhat, kappa)[:2]
            gfd[i] = (Ep - Em) / (2 * h)
            Hfd[:, i] = (gp - gm) / (2 * h)
        assert np.allclose(g, gfd, rtol=1e-5, atol=1e-6 * np.abs(g).max()), (t, g, gfd)
        assert np.allclose(H, Hfd, rtol=1e-4, atol=1e-5 * np.abs(H).max()), t
        assert np.array_equal(H, H.T)
        # translating both primitives together changes nothing
        assert abs(np.tile(np.eye(3), (4, 1)).T @ g).max() <= 1e-9 * np.abs(g).max()
    assert len(seen) >= 3


def test_energy_outside_and_at_contact():
    x = np.array([0.25, 0.25, 0.3, 0, 0, 0, 1, 0, 0, 0, 1, 0.0])
    E, g, H, t, d = O.ipc_pair(0, x, 0.3, 1.0)
    assert E == 0.0 and not g.any() and not H.any() and d == 0.3  # d >= dhat: no energy
    x[2] = 0.0
    E, g, H, t, d = O.ipc_pair(0, x, 0.3, 1.0)
    assert math.isinf(E) and E > 0 and d == 0.0


# ------------------------------------------------------------------ PSD projection
def test_psd_projection_matches_eigh():
    rng = np.random.default_rng(3000)
    for k in range(40):
        Q, _ = np.linalg.qr(rng.normal(size=(12, 12)))
        lam = rng.normal(size=12) * 10.0 ** rng.uniform(-3, 3, size=12)
        if k % 5 == 0:
            lam[:3] = 0.0  # a rank-deficient one
        A = (Q * lam) @ Q.T
        A = 0.5 * (A + A.T)
        P, lmin = O.psd_project12(A)
        w, V = np.linalg.eigh(A)
        ref = (V * np.maximum(w, 0.0)) @ V.T
        scale = np.abs(w).max()
        assert np.abs(P - ref).max() <= 1e-12 * scale, k
        assert abs(lmin - w.min()) <= 1e-12 * scale
        assert np.linalg.eigvalsh(0.5 * (P + P.T)).min() >= -1e-12 * scale
    # a PSD matrix is kept; a negative definite one goes to zero
    B = rng.normal(size=(12, 12))
    S = B @ B.T
    assert np.abs(O.psd_project12(S)[0] - S).max() <= 1e-12 * np.abs(S).max()
    assert not O.psd_project12(-S)[0].any()


# ------------------------------------------------------------------ the rigid map
def _qmat(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _rodrigues(th):
    a = np.linalg.norm(th)
    if a <= 1e-10:
        return np.eye(3)
    k = th / a
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K


def _xw_ref(xb, q, r):
    return xb[:3] + _rodrigues(xb[3:]) @ _qmat(q) @ r


def test_vertex_map_and_jacobian():
    rng = np.random.default_rng(4000)
    for k in range(30):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        r = rng.normal(size=3)
        xb = np.concatenate([rng.normal(size=3), np.zeros(3) if k % 2 == 0 else rng.normal(size=3) * 0.7])
        xw, J = O.vertex_jacobian(xb, q, r)
        assert np.abs(xw - _xw_ref(xb, q, r)).max() <= 1e-14 * (1 + np.abs(xw).max())
        if k % 2 == 0:
            w = _qmat(q) @ r
            assert np.abs(J - np.array([[0, w[2], -w[1]], [-w[2], 0, w[0]], [w[1], -w[0], 0]])).max() <= 1e-15 * (
                1 + np.abs(w).max())
        h = 1e-6
        Jfd = np.zeros((3, 3))
        for i in range(3):
            e = np.zeros(6)
            e[3 + i] = h
            Jfd[:, i] = (_xw_ref(xb + e, q, r) - _xw_ref(xb - e, q, r)) / (2 * h)
        assert np.abs(J - Jfd).max() <= 1e-8 * (1 + np.abs(r).max()), k


# ------------------------------------------------------------------ scene-level fold
def _unique_edges(F):
    seen, E = set(), []
    for f in F:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            k = (min(a, b), max(a, b))
            if k not in seen:
                seen.add(k)
                E.append(k)
    return E


@pytest.fixture(scope="module")
def three_cubes():
    """Ground slab plus two tilted 6e-4 m cubes, placed through x so that the lower
    cube's bottom face and the upper cube's edges sit within dhat of their neighbours.
    Off-centre and tilted so that no vertex lies on a region boundary (the ground's
    face diagonal) and no two edges are parallel: there the Hessian jumps and central
    differences straddle the jump."""
    size = 6e-4
    body = scenes._box_body
    spec = scenes.SceneSpec([body("g", 0.01, size, 0.01, [0, -0.5 * size, 0], static=True),
                             body("c", size, size, size, [1.3e-4, 0.5 * size + 1e-4, -0.7e-4],
                                  q=scenes.quat_from_axis_angle([1.0, 0.2, -0.5], 0.02)),
                             body("c", size, size, size, [1e-4, 1.5 * size + 2.5e-4, 0],
                                  q=scenes.quat_from_axis_angle([0.3, 0.1, 1.0], 0.03))],
                            dt=1e-3, dhat=1e-3, mu=0.0)
    sc = scenes.pack([spec])
    o = O.OracleScene(sc)
    n = sc.n_bodies
    x = np.zeros(6 * n)
    for b in range(n):
        x[6 * b:6 * b + 3] = sc.p[b]
    x[6 * 1 + 3:6 * 1 + 6] = [-1e-3, 5e-4, 2e-3]
    x[6 * 2 + 3:6 * 2 + 6] = [2e-3, -1e-3, 4e-3]
    V = [sc.verts.reshape(-1, 3)[sc.vert_off[b]:sc.vert_off[b] + sc.n_verts[b]] for b in range(n)]
    F = [sc.faces.reshape(-1, 3)[sc.face_off[b]:sc.face_off[b] + sc.n_faces[b]] for b in range(n)]
    return o, sc, x, V, F


def _pairs_near(sc, x, V, F, dhat):
    """Every PT (both directions) and EE pair of bodies (0, 1) and (1, 2) within dhat, in
    the order (body pair, kind, primitives), by brute force over world points."""
    n = sc.n_bodies
    W = [np.array([_xw_ref(x[6 * b:6 * b + 6], sc.q[b], r) for r in V[b]]) for b in range(n)]
    E = [_unique_edges(F[b]) for b in range(n)]
    out = []
    for A, B in ((0, 1), (1, 2)):
        for a, b in ((A, B), (B, A)):
            for vi in range(len(W[a])):
                for fi, f in enumerate(F[b]):
                    xx = np.concatenate([W[a][vi], W[b][f[0]], W[b][f[1]], W[b][f[2]]])
                    if _pt_ref(xx) < dhat:
                        out.append((0, a, b, vi, fi))
        for ea, (i0, i1) in enumerate(E[A]):
            for eb, (j0, j1) in enumerate(E[B]):
                xx = np.concatenate([W[A][i0], W[A][i1], W[B][j0], W[B][j1]])
                if _ee_ref(xx) < dhat:
                    out.append((1, A, B, ea, eb))
    return np.array(out, dtype=np.int32), W, E


def _pair_world(pr, W, F, E):
    k, a, b, i, j = pr
    if k == 0:
        f = F[b][j]
        return np.concatenate([W[a][i], W[b][f[0]], W[b][f[1]], W[b][f[2]]])
    return np.concatenate([W[a][E[a][i][0]], W[a][E[a][i][1]], W[b][E[b][j][0]], W[b][E[b][j][1]]])


def test_scene_energy_gradient_and_hessian(three_cubes):
    o, sc, x, V, F = three_cubes
    dhat, kappa = 3e-4, 1e4
    pairs, W, E = _pairs_near(sc, x, V, F, dhat)
    assert len(pairs) > 20 and set(pairs[:, 0]) == {0, 1}
    val, segb, sg, sH = o.ipc_barrier_eval(pairs, x, dhat, kappa, psd=False)
    # value: the sum over pairs of kappa b(d) at the world points of the independent map
    ref = sum(_barrier_ref(O.ipc_distance(int(p[0]), _pair_world(p, W, F, E))[0], dhat, kappa) for p in pairs)
    assert abs(val - ref) <= 1e-12 * abs(ref)
    assert segb.tolist() == [[0, 1], [1, 2]]

    def value(xx):
        return o.ipc_barrier_eval(pairs, xx, dhat, kappa, psd=False)[0]

    def seg_grad(xx):
        _, sb, g, _ = o.ipc_barrier_eval(pairs, xx, dhat, kappa, psd=False)
        full = np.zeros(6 * sc.n_bodies)
        for (lo, hi), gs in zip(sb, g):
            full[6 * lo:6 * lo + 6] += gs[:6]
            full[6 * hi:6 * hi + 6] += gs[6:]
        return full

    g = seg_grad(x)
    nd = 6 * sc.n_bodies
    gfd = np.zeros(nd)
    Hfd = np.zeros((nd, nd))
    h = np.where(np.arange(nd) % 6 < 3, 1e-10, 1e-7)
    for i in range(nd):
        e = np.zeros(nd)
        e[i] = h[i]
        gfd[i] = (value(x + e) - value(x - e)) / (2 * h[i])
        Hfd[:, i] = (seg_grad(x + e) - seg_grad(x - e)) / (2 * h[i])
    assert np.allclose(g, gfd, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    # assembled body Hessian; the translation rows/columns are exact (the map is linear in p)
    Hs = np.zeros((nd, nd))
    for (lo, hi), Hb in zip(segb, sH):
        idx = np.r_[6 * lo:6 * lo + 6, 6 * hi:6 * hi + 6]
        Hs[np.ix_(idx, idx)] += Hb
    tr = np.array([i for i in range(nd) if i % 6 < 3])
    assert np.allclose(Hs[:, tr], Hfd[:, tr], rtol=1e-4, atol=1e-5 * np.abs(Hs).max())
    # the rotation block: J^T H J with J from central differences of the independent map
    for s, (lo, hi) in enumerate(segb):
        Hsum = np.zeros((12, 12))
        for p in pairs:
            if sorted((p[1], p[2])) != [lo, hi]:
                continue
            xw = _pair_world(p, W, F, E)
            Ep, gp, Hp, _, _ = O.ipc_pair(int(p[0]), xw, dhat, kappa)
            if Ep == 0.0:
                continue
            verts = ([(p[1], V[p[1]][p[3]])] + [(p[2], V[p[2]][k]) for k in F[p[2]][p[4]]]) if p[0] == 0 else \
                [(p[1], V[p[1]][k]) for k in E[p[1]][p[3]]] + [(p[2], V[p[2]][k]) for k in E[p[2]][p[4]]]
            J = np.zeros((12, 12))
            for v, (b, r) in enumerate(verts):
                col = 0 if b == lo else 6
                J[3 * v:3 * v + 3, col:col + 3] = np.eye(3)
                for k in range(3):
                    e = np.zeros(6)
                    e[3 + k] = 1e-7
                    xb = x[6 * b:6 * b + 6]
                    J[3 * v:3 * v + 3, col + 3 + k] = (_xw_ref(xb + e, sc.q[b], r) - _xw_ref(xb - e, sc.q[b], r)) / 2e-7
            Hsum += J.T @ Hp @ J
        assert np.allclose(sH[s], Hsum, rtol=1e-6, atol=1e-7 * np.abs(Hsum).max()), s
    # with the projection: every pair's block is PSD, so each segment's is
    _, _, _, sHp = o.ipc_barrier_eval(pairs, x, dhat, kappa, psd=True)
    for Hb in sHp:
        assert np.linalg.eigvalsh(Hb).min() >= -1e-10 * np.abs(Hb).max()


def test_segments_and_bad_pairs(three_cubes):
    o, sc, x, V, F = three_cubes
    dhat, kappa = 3e-4, 1e4
    pairs, W, E = _pairs_near(sc, x, V, F, dhat)
    p01 = [p for p in pairs if sorted(p[1:3]) == [0, 1]]
    p12 = [p for p in pairs if sorted(p[1:3]) == [1, 2]]
    order = np.array([p01[0], p01[1], p12[0], p01[2], p12[1]], dtype=np.int32)
    val, segb, sg, sH = o.ipc_barrier_eval(order, x, dhat, kappa, psd=True)
    assert segb.tolist() == [[0, 1], [1, 2], [0, 1], [1, 2]]
    parts = [o.ipc_barrier_eval(order[i:j], x, dhat, kappa, psd=True) for i, j in ((0, 2), (2, 3), (3, 4), (4, 5))]
    assert val == sum(p[0] for p in parts) or abs(val - sum(p[0] for p in parts)) <= 1e-15 * abs(val)
    for s, p in enumerate(parts):
        assert np.array_equal(sg[s], p[2][0]) and np.array_equal(sH[s], p[3][0])
    # a PT pair whose vertex is on the higher body: slot 0 is still the lower body, and
    # its gradient goes to the face's body there
    rev = [p for p in p01 if p[0] == 0 and p[1] == 1]
    assert rev
    v1, sb1, g1, _ = o.ipc_barrier_eval(np.array(rev[:1]), x, dhat, kappa)
    assert sb1.tolist() == [[0, 1]] and v1 > 0.0
    assert np.abs(g1[0, :3] + g1[0, 6:9]).max() <= 1e-9 * np.abs(g1[0, :3]).max()  # no net translation force
    for bad in ([0, 1, 1, 0, 0], [0, 0, 5, 0, 0], [0, 0, 1, 99, 0], [1, 0, 1, 0, 99], [0, -1, 1, 0, 0]):
        with pytest.raises(ValueError):
            o.ipc_barrier_eval(np.array([bad], dtype=np.int32), x, dhat, kappa)


# ------------------------------------------------------------------ ipc-mode CCD and maxStep
_ID = [1.0, 0.0, 0.0, 0.0]
_TRI = [0, 0, 0, 1, 0, 0, 0, 1, 0]


def test_face_vertex_ccd_triangle_domain():
    """A vertex falling through the triangle's interior meets it at t = 0.5: kind 3
    reports it, conservatively (toi <= 0.5, within the time tolerance); the same fall
    in the corner u + v > 1 of the parameter box misses; the reference's kind 2 never
    reports (its predicate drops the root box)."""
    tol = 1e-8
    hit, toi, pops = O.ccd_toi(3, [0.25, 0.25, 1.0] + _ID + [0.25, 0.25, -1.0] + _ID, [0] * 3 + _ID + [0] * 3 + _ID,
                               [0, 0, 0], _TRI, tol=tol)
    assert hit == 1 and 0.5 - 2 * tol <= toi <= 0.5 and pops < 1000
    for k in (2, 3):
        hit, toi, _ = O.ccd_toi(k, [0.8, 0.8, 1.0] + _ID + [0.8, 0.8, -1.0] + _ID, [0] * 3 + _ID + [0] * 3 + _ID,
                                [0, 0, 0], _TRI, tol=tol)
        assert hit == 0 and toi == math.inf
    hit, _, pops = O.ccd_toi(2, [0.25, 0.25, 1.0] + _ID + [0.25, 0.25, -1.0] + _ID, [0] * 3 + _ID + [0] * 3 + _ID,
                             [0, 0, 0], _TRI, tol=tol)
    assert hit == 0 and pops == 1
    # the face moves instead (its body translates up through the still vertex)
    hit, toi, _ = O.ccd_toi(3, [0.2, 0.3, 0.5] + _ID + [0.2, 0.3, 0.5] + _ID, [0] * 3 + _ID + [0, 0, 2.0] + _ID,
                            [0, 0, 0], _TRI, tol=tol)
    assert hit == 1 and 0.25 - 2 * tol <= toi <= 0.25


def test_edge_edge_ccd_linear_translation_and_rotation():
    """kind 4: (a) an edge translating onto a crossing edge meets it at the fraction of
    the step that closes the gap; the reference's kind 1 (hull over the whole step)
    reports a time near the end of the step instead."""
    tol = 1e-8
    a = [-1, 0, 0, 1, 0, 0]
    b = [0, -1, 0, 0, 1, 0]
    pa = [0, 0, 1.0] + _ID + [0, 0, -3.0] + _ID  # the crossing is reached at t = 1/4
    pb = [0, 0, 0.0] + _ID + [0, 0, 0.0] + _ID
    hit, toi, _ = O.ccd_toi(4, pa, pb, a, b, tol=tol)
    assert hit == 1 and 0.25 - 2 * tol <= toi <= 0.25
    hit1, toi1, pops1 = O.ccd_toi(1, pa, pb, a, b, tol=tol)
    assert hit1 == 1 and toi1 > 0.99 and pops1 == 1000
    # (b) an edge spinning 0.5 rad about a point on it sweeps a vertical edge at radius
    # 0.37 and angle 0.5 / 3: t = 1/3 (radius and heights off the dyadic split points)
    ar = [-1.1, 0, 0, 0.9, 0, 0]
    qz = scenes.quat_from_axis_angle([0, 0, 1], 0.5).tolist()
    c, s = math.cos(0.5 / 3), math.sin(0.5 / 3)
    hit, toi, pops = O.ccd_toi(4, [0, 0, 0] + _ID + [0, 0, 0] + qz, [0] * 3 + _ID + [0] * 3 + _ID, ar,
                               [0.37 * c, 0.37 * s, -0.7, 0.37 * c, 0.37 * s, 1.1], tol=tol)
    assert hit == 1 and 1 / 3 - 1e-7 <= toi <= 1 / 3 and pops < 1000
    # the same edge at radius 1.5 is never reached
    hit, _, _ = O.ccd_toi(4, [0, 0, 0] + _ID + [0, 0, 0] + qz, [0] * 3 + _ID + [0] * 3 + _ID, ar,
                          [1.5 * c, 1.5 * s, -0.7, 1.5 * c, 1.5 * s, 1.1], tol=tol)
    assert hit == 0
    # a search stopped by the pop cap still reports a lower bound: the quarter turn
    qq = scenes.quat_from_axis_angle([0, 0, 1], math.pi / 2).tolist()
    c, s = math.cos(math.pi / 6), math.sin(math.pi / 6)
    hit, toi, pops = O.ccd_toi(4, [0, 0, 0] + _ID + [0, 0, 0] + qq, [0] * 3 + _ID + [0] * 3 + _ID, ar,
                               [0.37 * c, 0.37 * s, -0.7, 0.37 * c, 0.37 * s, 1.1], tol=tol)
    assert hit == 1 and toi <= 1 / 3


def test_ipc_max_step(three_cubes):
    """maxStep by conservative advancement, keeping sep = 0.1 of each pair's distance:
    (a) the upper cube, level over the lower one at gap g, moved straight down by 3 g:
    every close pair reaches 0.1 g at t = 0.3, and the bound converges to it from below;
    (b) moved sideways by half the gap, nothing comes closer: alpha = 1; (c) a tumbling
    step: sampled densely along the path (independent map, textbook distances), every
    pair stays above 0.1 of its starting distance up to toi, and some pair is within 1%
    of that at toi; (d) the bound over a list is the least of its pairs' bounds; bad
    pairs are named."""
    o, sc, x, V, F = three_cubes
    size = 6e-4
    lvl = O.OracleScene(_level(sc))
    xs = x.copy()
    xs[6 + 3:6 + 6] = 0.0
    xs[12 + 3:12 + 6] = 0.0
    gap = 1e-4
    xs[12:15] = [xs[6], xs[7] + size + gap, xs[8]]
    pairs, _, _ = _pairs_near(lvl.sc, xs, V, F, 2 * gap)
    pairs = pairs[(pairs[:, 1] + pairs[:, 2]) == 3]  # bodies 1 and 2
    assert len(pairs) > 8
    dx = np.zeros_like(xs)
    dx[13] = -3 * gap
    alpha, toi = lvl.ipc_max_step(pairs, xs, dx)
    assert 0.3 * (1 - 1e-5) <= toi <= 0.3 and alpha == toi * 0.8
    dx2 = np.zeros_like(xs)
    dx2[12] = 0.5 * gap
    assert lvl.ipc_max_step(pairs, xs, dx2) == (1.0, math.inf)
    # (c) tumbling: translation and rotation of the upper cube
    dx3 = np.zeros_like(xs)
    dx3[12:18] = [0.3 * gap, -2.5 * gap, 0.2 * gap, 0.4, -0.3, 0.5]
    alpha, toi = lvl.ipc_max_step(pairs, xs, dx3)
    assert 0.0 < toi < 1.0
    q = lvl.sc.q
    E = [_unique_edges(F[b]) for b in range(3)]

    def dists(t):
        xt = xs + t * dx3
        W = [np.array([_xw_ref(xt[6 * b:6 * b + 6], q[b], r) for r in V[b]]) for b in range(3)]
        return np.array([(_pt_ref if p[0] == 0 else _ee_ref)(_pair_world(p, W, F, E)) for p in pairs])

    d0 = dists(0.0)
    for t in np.linspace(0.0, toi, 41):
        assert (dists(t) >= 0.1 * d0 * (1 - 1e-9)).all(), t
    assert ((dists(toi) - 0.1 * d0) / d0).min() < 0.01
    per = [lvl.ipc_max_step(p[None, :], xs, dx3)[1] for p in pairs]
    assert toi == min(per)
    bad = pairs[:4].copy()
    bad[2, 3] = 999
    with pytest.raises(ValueError, match="bad pair 2"):
        lvl.ipc_max_step(bad, xs, dx)


def _level(sc):
    import copy
    s2 = copy.deepcopy(sc)
    s2.q[1:] = [1.0, 0.0, 0.0, 0.0]
    return s2


# ------------------------------------------------------------------ ipc-mode time step
def _drop_scene(n_cubes=1, seed=0):
    rng = np.random.default_rng(seed)
    body = scenes._box_body
    bodies = [body("g", 4, 1, 4, [0, -0.5, 0], static=True)]
    for k in range(n_cubes):
        bodies.append(body("c", 1, 1, 1, [0.1 + 0.05 * k, 0.6 + 1.05 * k, -0.05],
                           q=scenes.quat_from_axis_angle(rng.normal(size=3), 0.1 + 0.05 * k)))
    return scenes.pack([scenes.SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.0)])


def _lowest_vertex(sc, p, q, b):
    V = sc.verts.reshape(-1, 3)[sc.vert_off[b]:sc.vert_off[b] + sc.n_verts[b]]
    return min((p[b] + _qmat(q[b]) @ v)[1] for v in V)


def test_ipc_step_cube_lands_and_rests_without_penetration():
    """A tilted unit cube dropped 0.1 m onto a static box (its top face at y = 0): with the
    ipc mode's primitive-distance barrier (dhat 1 cm) it lands, never crosses the face,
    and comes to rest within dhat of it (there is no friction, so it may still slide and
    spin about the vertical);
    the reference's centre-distance barrier (ref mode, the same scene) lets it sink into
    the box."""
    sc = _drop_scene()
    o = O.OracleScene(sc)
    lows = []
    for k in range(80):
        st = o.ipc_step(1e-2, dhat=1e-2, kappa=1e5, inflation=0.05)
        assert st[2] > 0 and st[0] <= 50
        p, q, v, w = o.get_state()[:4]
        lows.append(_lowest_vertex(sc, p, q, 1))
    assert min(lows) > 0.0
    # at rest vertically and not tipping (no friction: it may still slide and spin about y)
    assert 0.0 < lows[-1] < 1e-2 and abs(v[1][1]) < 1e-4 and max(abs(w[1][0]), abs(w[1][2])) < 1e-3
    ref = O.OracleScene(sc)
    for k in range(80):
        ref.step(1e-2)
    p, q = ref.get_state()[:2]
    assert _lowest_vertex(sc, p, q, 1) < -0.1


def test_ipc_step_rejects_other_joints():
    sc = scenes.hinge_chain_scene(n_links=3)
    with pytest.raises(ValueError, match="single-body fixed"):
        O.OracleScene(sc).ipc_step(1e-2)


def test_step_pcg_solves_block_systems():
    """The ipc step's PCG (the device solver's order of operations) against numpy's
    dense solve on random SPD block systems, off-diagonal blocks listed once (some pairs
    twice, as repeated body-pair segments are), with relative tolerance 1e-14."""
    rng = np.random.default_rng(31)
    for nb in (1, 3, 9, 40):
        pairs = [(i, j) for i in range(nb) for j in range(i + 1, nb) if rng.random() < 0.3]
        pairs += pairs[:2]
        offij = np.array(pairs, dtype=np.int32).reshape(-1, 2)
        offb = rng.normal(size=(len(pairs), 6, 6)) * 0.3
        A = np.zeros((6 * nb, 6 * nb))
        for (i, j), B in zip(pairs, offb):
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] += B
            A[6 * j:6 * j + 6, 6 * i:6 * i + 6] += B.T
        diag = np.zeros((nb, 6, 6))
        for i in range(nb):
            M = rng.normal(size=(6, 6))
            diag[i] = M @ M.T + 6.0 * np.eye(6) + np.abs(A[6 * i:6 * i + 6]).sum() * np.eye(6)
            A[6 * i:6 * i + 6, 6 * i:6 * i + 6] += diag[i]
        b = rng.normal(size=6 * nb)
        x, it = O.pcg_ordered(diag.reshape(-1), offij, offb, b, tol=1e-14, max_it=2000)
        ref = np.linalg.solve(A, b)
        assert np.abs(x - ref).max() <= 1e-11 * np.abs(ref).max(), nb
        assert 0 < it < 2000
