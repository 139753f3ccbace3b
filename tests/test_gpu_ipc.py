"""ipc mode on the GPU against the oracle's evaluation of the same inputs:
ndipc_ipc_barrier_eval (total energy, segments, per-segment body gradient and Hessian,
bitwise, with and without the PSD projection; the segment structure of interleaved
pair lists; the error for a bad pair; a size-independent check at 100k pairs, the
segment sums against the per-pair evaluations summed on the host), the CCD query kinds
3 and 4 (bitwise on random moving primitives) and ndipc_ipc_max_step (bitwise)."""
import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scenes
from test_oracle_ipc import _pairs_near

pytestmark = pytest.mark.gpu


def _three_cubes():
    size = 6e-4
    body = scenes._box_body
    spec = scenes.SceneSpec([body("g", 0.01, size, 0.01, [0, -0.5 * size, 0], static=True),
                             body("c", size, size, size, [1.3e-4, 0.5 * size + 1e-4, -0.7e-4],
                                  q=scenes.quat_from_axis_angle([1.0, 0.2, -0.5], 0.02)),
                             body("c", size, size, size, [1e-4, 1.5 * size + 2.5e-4, 0],
                                  q=scenes.quat_from_axis_angle([0.3, 0.1, 1.0], 0.03))],
                            dt=1e-3, dhat=1e-3, mu=0.0)
    sc = scenes.pack([spec])
    n = sc.n_bodies
    x = np.zeros(6 * n)
    for b in range(n):
        x[6 * b:6 * b + 3] = sc.p[b]
    x[6 * 1 + 3:6 * 1 + 6] = [-1e-3, 5e-4, 2e-3]
    x[6 * 2 + 3:6 * 2 + 6] = [2e-3, -1e-3, 4e-3]
    V = [sc.verts.reshape(-1, 3)[sc.vert_off[b]:sc.vert_off[b] + sc.n_verts[b]] for b in range(n)]
    F = [sc.faces.reshape(-1, 3)[sc.face_off[b]:sc.face_off[b] + sc.n_faces[b]] for b in range(n)]
    return sc, x, V, F


def _same(dev, ora):
    vd, sbd, sgd, sHd = dev
    vo, sbo, sgo, sHo = ora
    assert vd == vo
    assert np.array_equal(sbd, sbo)
    assert np.array_equal(sgd, sgo)
    assert np.array_equal(sHd, sHo)


@pytest.mark.parametrize("psd", [True, False])
def test_ipc_eval_matches_oracle(engine_factory, psd):
    sc, x, V, F = _three_cubes()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    for dhat, kappa in ((3e-4, 1e4), (1e-3, 2.5e6)):
        pairs, _, _ = _pairs_near(sc, x, V, F, dhat)
        assert len(pairs) > 20
        # pairs beyond dhat in the list too: they open or continue a segment and add nothing
        far, _, _ = _pairs_near(sc, x, V, F, 2.5 * dhat)
        for pl in (pairs, far, far[::-1]):
            _same(e.ipc_barrier_eval(0, pl, x, dhat, kappa, psd), o.ipc_barrier_eval(pl, x, dhat, kappa, psd))
        # the pairs within dhat are the close list (textbook distances)
        assert e.ipc_barrier_eval(0, far, x, dhat, kappa, psd, with_active=True)[4] == len(pairs)


def test_ipc_eval_segments_and_errors(engine_factory):
    sc, x, V, F = _three_cubes()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    dhat, kappa = 3e-4, 1e4
    pairs, _, _ = _pairs_near(sc, x, V, F, dhat)
    rng = np.random.default_rng(7)
    for _ in range(3):
        pl = pairs[rng.permutation(len(pairs))]
        dev = e.ipc_barrier_eval(0, pl, x, dhat, kappa)
        _same(dev, o.ipc_barrier_eval(pl, x, dhat, kappa))
        assert len(dev[1]) > 2  # interleaved body pairs: many segments
    v, sb, sg, sH = e.ipc_barrier_eval(0, pairs[:0], x, dhat, kappa)
    assert v == 0.0 and len(sb) == 0
    bad = pairs[:6].copy()
    bad[4] = [0, 0, 1, 99, 0]
    with pytest.raises(RuntimeError, match="ipc pair 4"):
        e.ipc_barrier_eval(0, bad, x, dhat, kappa)
    bad[4] = [2, 0, 1, 0, 0]
    with pytest.raises(RuntimeError, match="ipc pair 4"):
        e.ipc_barrier_eval(0, bad, x, dhat, kappa)
    # the context still works after the errors
    _same(e.ipc_barrier_eval(0, pairs, x, dhat, kappa), o.ipc_barrier_eval(pairs, x, dhat, kappa))


def test_ipc_eval_large_list(engine_factory):
    """100k pairs of a stack of 3 cubes drawn at random from the close pairs, sorted by
    body pair: bitwise against the oracle, and the segment sums against the per-pair
    evaluations of the same list."""
    sc, x, V, F = _three_cubes()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    dhat, kappa = 1e-3, 1e5
    pairs, _, _ = _pairs_near(sc, x, V, F, dhat)
    rng = np.random.default_rng(11)
    pl = pairs[rng.integers(0, len(pairs), size=100_000)]
    key = np.minimum(pl[:, 1], pl[:, 2]) * 8 + np.maximum(pl[:, 1], pl[:, 2])
    pl = pl[np.argsort(key, kind="stable")]
    dev = e.ipc_barrier_eval(0, pl, x, dhat, kappa)
    _same(dev, o.ipc_barrier_eval(pl, x, dhat, kappa))
    assert dev[1].tolist() == [[0, 1], [1, 2]]
    # linearity: the list is a multiset of distinct pairs; summing each distinct pair's
    # own evaluation times its multiplicity gives the segment gradient to rounding
    uniq, cnt = np.unique(pl, axis=0, return_counts=True)
    g_ref = {(0, 1): np.zeros(12), (1, 2): np.zeros(12)}
    for p, c in zip(uniq, cnt):
        _, sb, sg, _ = e.ipc_barrier_eval(0, p[None, :], x, dhat, kappa)
        g_ref[tuple(sb[0])] += c * sg[0]
    for s, k in enumerate(((0, 1), (1, 2))):
        assert np.allclose(dev[2][s], g_ref[k], rtol=1e-9, atol=1e-12 * np.abs(g_ref[k]).max())


def _qrot(q, v):
    w, x, y, z = q
    u = np.array([x, y, z])
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def _ccd_queries(kind, n, rng):
    """Random queries of kind 3 or 4, most of whose motions pass near a point of B's
    primitive (a random convex combination of its corners at t = 0), with small
    rotations of both bodies: hits and misses both."""
    pa = np.zeros((n, 14))
    pb = np.zeros((n, 14))
    for P in (pa, pb):
        q = rng.normal(size=(n, 4))
        P[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
        dq = P[:, 3:7] + 0.05 * rng.normal(size=(n, 4))
        P[:, 10:14] = dq / np.linalg.norm(dq, axis=1, keepdims=True)
    pb[:, 0:3] = rng.normal(size=(n, 3)) * 0.1
    pb[:, 7:10] = pb[:, 0:3] + rng.normal(size=(n, 3)) * 0.02
    a = rng.normal(size=(n, 6)) * 0.05
    b = rng.normal(size=(n, 9)) * 0.5
    nbp = 3 if kind == 3 else 2
    for i in range(n):
        w = rng.dirichlet(np.ones(nbp))
        c = pb[i, 0:3] + sum(w[k] * _qrot(pb[i, 3:7], b[i, 3 * k:3 * k + 3]) for k in range(nbp))
        d = rng.normal(size=3) * 0.3
        pa[i, 0:3] = c + d
        pa[i, 7:10] = c - d + rng.normal(size=3) * 0.05
    return pa, pb, a, b


def test_ipc_ccd_kinds_match_oracle(engine_factory):
    """Query kinds 3 and 4 on the device against the oracle, bitwise, on random moving
    and spinning primitives (hits and misses)."""
    sc, x, V, F = _three_cubes()
    e = engine_factory(sc)
    rng = np.random.default_rng(21)
    n = 300
    for kind in (3, 4):
        pa, pb, a, b = _ccd_queries(kind, n, rng)
        na, nb = (3, 9) if kind == 3 else (6, 6)
        hit, toi, pops = e.ccd_toi_batch(kind, pa, pb, a[:, :na], b[:, :nb], 1.0)
        assert n // 10 < (hit == 1).sum() < n - n // 10
        for i in range(n):
            r, t, p = O.ccd_toi(kind, pa[i], pb[i], a[i, :na], b[i, :nb])
            assert (hit[i], pops[i]) == (r, p) and (t == toi[i] or (r == 0 and toi[i] == np.inf)), (kind, i)


def test_ipc_max_step_matches_oracle(engine_factory):
    sc, x, V, F = _three_cubes()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    pairs, _, _ = _pairs_near(sc, x, V, F, 1e-3)
    rng = np.random.default_rng(5)
    tois = []
    for k in range(6):
        dx = rng.normal(size=len(x)) * 3e-4
        for c in (3, 4, 5):
            dx[c::6] *= 0.05  # small rotations
        dx[:6] = 0.0  # the ground stays
        dev = e.ipc_max_step(0, pairs, x, dx)
        ora = o.ipc_max_step(pairs, x, dx)
        assert dev == ora, (k, dev, ora)
        tois.append(dev[1])
    assert min(tois) < 1.0
    bad = pairs[:5].copy()
    bad[3, 0] = 7
    with pytest.raises(RuntimeError, match="ipc pair 3"):
        e.ipc_max_step(0, bad, x, dx)
    assert e.ipc_max_step(0, pairs[:0], x, dx) == (1.0, np.inf)


def test_ipc_step_matches_oracle(engine_factory):
    """The ipc-mode time step on the device against the oracle's: three tilted cubes
    dropped in a stack onto a static box, 40 steps, every state bitwise, the same Newton
    iterations, line-search trials, pair records and PCG iterations."""
    from test_oracle_ipc import _drop_scene

    sc = _drop_scene(n_cubes=3, seed=4)
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    kw = dict(dhat=1e-2, kappa=1e5, inflation=0.05)
    for k in range(40):
        so = o.ipc_step(1e-2, **kw)
        sg = e.ipc_step(0, 1e-2, **kw)
        assert so == sg, (k, so, sg)
        for a, b in zip(o.get_state()[:4], e.get_state()[:4]):
            assert np.array_equal(a, b), k
    with pytest.raises(RuntimeError, match="single-body fixed"):
        e2 = engine_factory(scenes.hinge_chain_scene(n_links=3))
        e2.ipc_step(0, 1e-2)


def test_ipc_step_grid_of_cubes_matches_oracle(engine_factory):
    """The bench's ipc-step workload (32 cubes over a ground box, ~5e4 pair records): the
    first two steps bitwise against the oracle."""
    sc = scenes.ipc_drop_scene()
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    kw = dict(dhat=1e-2, kappa=1e5, inflation=0.05)
    for k in range(2):
        assert o.ipc_step(1e-2, **kw) == e.ipc_step(0, 1e-2, **kw), k
        for a, b in zip(o.get_state()[:4], e.get_state()[:4]):
            assert np.array_equal(a, b), k
