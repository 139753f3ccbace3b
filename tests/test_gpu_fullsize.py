"""Parity of the HIP path with the CPU oracle at the benchmark configurations' full sizes:
C2 (512 icospheres + container, one scene), C3 (the 256-scene batch) over its whole run on
the scenes that reach the reference's stiff regime, C5 (128 scenes) over its whole run,
the in-step block-Jacobi PCG branch (a coupled component above 64 bodies), and the
2000-contact cap.

The oracle evaluates computeSceneMinDistance with exact box pruning for large scenes
(ndo_set_scene_min_pruned; pinned equal to the unpruned loop for the 1e-9 comparison in
tests/test_oracle_pins.py), since the unpruned loop is ~1e11 tests per C2 step.
"""
import collections

import numpy as np
import pytest

import _oracle as O
from nexdynipc_amd import scenes
from test_gpu_parity import quat_angle

pytestmark = pytest.mark.gpu


def _oracle(sc, s=0):
    o = O.OracleScene(sc, s)
    o.set_scene_min_pruned(True)
    return o


def _same_candidates(o, e, dhat):
    ro = o.candidates(dhat)
    rg = e.candidates(0, dhat)
    for a, b in zip(ro, rg):
        np.testing.assert_array_equal(a, b)
    return sum(len(x) for x in ro)


def _same_contacts(o, e, x=None):
    oi, od = o.narrow_phase(x)
    gi, gd = e.narrow_phase(0, x)
    np.testing.assert_array_equal(oi, gi)
    assert np.array_equal(od.view(np.int64), gd.view(np.int64))
    return len(oi)


def test_c2_full_size_candidates_narrow_phase_and_trajectory(engine_factory):
    """C2 at full size: body pairs, the six top-K lists of every close pair (as the
    emitted VV/EV/EE/FV tuples) and the narrow phase bitwise at the start state, then ten
    steps bitwise (poses, velocities, iteration counts, contact counts), then candidates
    and narrow phase again at the state reached and at a perturbed iterate."""
    sc = scenes.make_config("c2")
    e = engine_factory(sc)
    o = _oracle(sc)
    dhat = float(sc.dhat[0])
    for r in (dhat, 2 * dhat, 0.05):
        np.testing.assert_array_equal(o.body_pairs(r), e.body_pairs(0, r))
        assert o.min_body_pair_distance(r) == e.min_body_pair_distance(0, r)
    assert _same_candidates(o, e, dhat) > 100000
    _same_contacts(o, e)
    dt = float(sc.dt[0])
    for k in range(10):
        o.step(dt)
        e.step(1)
        so, sg = o.get_state(), e.get_state()
        for a, b in zip(so[:4], sg[:4]):
            assert np.array_equal(a, b), k
        assert np.array_equal(e.stats()[0][:3], np.asarray(o.stats())[:3]), k
    assert o.barrier()[0] == e.barrier()[0][0]
    assert _same_candidates(o, e, dhat) > 100000
    n = _same_contacts(o, e)
    p = e.get_state()[0]
    rng = np.random.default_rng(8)
    x = np.zeros(6 * sc.n_bodies)
    for k in range(3):
        x[k::6] = p[:, k]
    x += rng.normal(0, 2e-4, size=x.shape)
    n2 = _same_contacts(o, e, x)
    assert n + n2 > 0


def test_c3_stiff_scenes_whole_run(engine_factory):
    """The 256-scene batch for its whole 100 steps; the three scenes with the most Newton
    iterations in a step (the reference's stiff regime: 20+ iterations, hundreds of
    line-search trials) are compared with the oracle at every step, bitwise while their
    coupled components stay at or below 64 bodies (direct solves), else within 1e-8."""
    specs = scenes.config3()
    sc = scenes.pack(specs)
    e = engine_factory(sc)
    steps = 100
    states, iters, pcg = [], np.zeros((steps, sc.n_scenes)), np.zeros(sc.n_scenes)
    for k in range(steps):
        e.step(1)
        st = e.get_state()
        states.append((st[0].copy(), st[1].copy()))
        c = e.solver_counters()
        iters[k] = c[:, 0]
        pcg += c[:, 3]
    worst = np.argsort(-iters.max(axis=0), kind="stable")[:3]
    assert iters.max() >= 20
    for s in worst:
        s = int(s)
        sl = sc.scene_slice(s)
        o = _oracle(sc, s)
        dt = float(sc.dt[s])
        for k in range(steps):
            o.step(dt)
            po, qo = o.get_state()[:2]
            pg, qg = states[k][0][sl], states[k][1][sl]
            if pcg[s] == 0:
                assert np.array_equal(pg, po) and np.array_equal(qg, qo), (s, k)
            else:
                assert np.max(np.abs(pg - po)) <= 1e-8 and np.max(quat_angle(qg, qo)) <= 1e-8, (s, k)


def test_c5_sixteen_scenes_whole_run(engine_factory):
    """The 128-scene C5 batch for its whole 50 steps; 16 seeded scenes compared with the
    oracle at every step (bitwise)."""
    sc = scenes.pack(scenes.config5(128))
    e = engine_factory(sc)
    steps = 50
    states = []
    for _ in range(steps):
        e.step(1)
        st = e.get_state()
        states.append((st[0].copy(), st[1].copy()))
    pick = np.random.default_rng(5).choice(sc.n_scenes, size=16, replace=False)
    for s in pick:
        s = int(s)
        sl = sc.scene_slice(s)
        o = _oracle(sc, s)
        for k in range(steps):
            o.step(float(sc.dt[s]))
            po, qo = o.get_state()[:2]
            assert np.array_equal(states[k][0][sl], po) and np.array_equal(states[k][1][sl], qo), (s, k)


def test_in_step_pcg_branch_matches_oracle(engine_factory):
    """A free 70-link hinge chain is one coupled component of 70 bodies (> 64): the device
    Newton solve takes the block-Jacobi PCG branch (LinearSolver.cpp:11-38 replaced), the
    oracle its LDL^T. With the links at rest the hinges stay satisfied and every solve
    converges: every step's poses within 1e-8 m / 1e-8 rad. (Once the links turn, the
    chain's Newton matrix reaches condition numbers ~1e7, where block-Jacobi CG does not
    reach its tolerance within 2 x rows iterations; the device solve then fails and the
    Newton loop ends, as the reference's does when its solve fails. Components above 64
    bodies do not occur in the benchmark configurations.)"""
    sc = scenes.hinge_chain_scene(n_links=70)
    sc.v[:] = 0.0
    sc.w[:] = 0.0
    e = engine_factory(sc)
    o = _oracle(sc)
    dt = float(sc.dt[0])
    n_pcg = 0
    for k in range(5):
        o.step(dt)
        e.step(1)
        c = e.solver_counters()[0]
        n_pcg += c[3]
        assert c[5] == 70  # the largest coupled component
        po, qo = o.get_state()[:2]
        pg, qg = e.get_state()[:2]
        assert np.max(np.abs(pg - po)) <= 1e-8 and np.max(quat_angle(qg, qo)) <= 1e-8, k
    assert n_pcg > 0


def test_contact_cap_keeps_the_2000_closest(engine_factory):
    """A stack of coincident cubes gives 7420 raw contacts; device and oracle keep 2000:
    the same multiset of distances, and the same contacts below the cut distance (which
    of several contacts at exactly the cut distance is kept is the sort's tie order)."""
    rng = np.random.default_rng(1)
    bodies = [scenes._box_body("ground", 10, 1, 10, [0, -0.5, 0], static=True)]
    for _ in range(10):
        bodies.append(scenes._box_body("cube", 1, 1, 1, [rng.uniform(-1e-4, 1e-4), 0.5 + rng.uniform(-1e-4, 1e-4),
                                                          rng.uniform(-1e-4, 1e-4)]))
    sc = scenes.pack([scenes.SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.3)])
    e = engine_factory(sc)
    o = O.OracleScene(sc)
    _same_candidates(o, e, 1e-3)
    oi, od = o.narrow_phase(None)
    gi, gd = e.narrow_phase(0, None)
    assert len(oi) == len(gi) == 2000
    assert np.array_equal(np.sort(od[:, 6]), np.sort(gd[:, 6]))
    cut = np.max(od[:, 6])
    ko = collections.Counter(tuple(r) + tuple(d) for r, d in zip(oi.tolist(), od.tolist()) if d[6] < cut)
    kg = collections.Counter(tuple(r) + tuple(d) for r, d in zip(gi.tolist(), gd.tolist()) if d[6] < cut)
    assert ko == kg
