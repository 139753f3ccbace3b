"""ipc mode on the GPU (ndipc_ipc_barrier_eval) against the oracle's evaluation of the
same pairs: total energy, segments, per-segment body gradient and Hessian, bitwise,
with and without the PSD projection; the segment structure of interleaved pair lists;
the error for a bad pair; and a size-independent check at 100k pairs (the segment sums
against the per-pair evaluations summed on the host)."""
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
