"""Device adjoint (ndipc_adjoint_forward / ndipc_adjoint_backward, and the
torch.autograd.Function nexdynipc_amd.autograd.Rollout) against two references:

  1. the oracle's FP64 CPU adjoint (ndo_adjoint): the same implicit-function-theorem sweep
     over the recorded ALM solves, with every derivative taken by central differences of
     the CPU restatement's own functions and the transposed systems solved by dense LU.
     Compared on every gradient (start poses, velocities, masses, mu): within 1e-6 of the
     largest entry of each gradient array;
  2. central finite differences of the device forward run (bitwise the oracle's
     trajectory, test_gpu_parity), with h = 1e-6, or the first of 1e-5, 1e-4 whose +h and
     -h runs keep the reference's discrete path (the ALM and Newton iteration counts of
     every step): where the path changes the trajectory is not differentiable at that
     scale. Components with no such step are reported, not compared. Bar: within 1e-5
     relative, normwise over the compared components.

Both adjoints hold the contact sets fixed (discrete) and follow each contact's point and
normal, and the lagged normal forces, to the state the narrow phase ran at (the previous
ALM solve's x*, or the step-start poses for the first solve). The friction-coefficient
gradient is checked against differences on millimetre cubes sliding on a plate with the
friction acting (scenes.friction_slide_scene).
"""
import numpy as np
import pytest
import torch

import _oracle as O
from nexdynipc_amd import scenes
from nexdynipc_amd.autograd import Rollout, Simulation

pytestmark = pytest.mark.gpu

FD_STEPS = (1e-6, 1e-5, 1e-4)
FD_RTOL = 1e-5
CPU_RTOL = 1e-6


@pytest.fixture(scope="module")
def c4():
    spec, targets = scenes.config4()
    sc = scenes.pack([spec])
    rng = np.random.default_rng(11)
    cq = rng.normal(0.0, 1e-2, size=(sc.n_bodies, 4))
    return sc, np.asarray(targets), cq


def _device_grads(sim, sc, t, cq, v0=None, every=None):
    v0 = sc.v if v0 is None else v0
    sim.reset(v0, sc.w, sc.mass, sc.mu)
    sim.e.adjoint_forward(sim.n_steps, every or sim.every)
    p, q = sim.e.get_state()[:2]
    g = sim.e.adjoint_backward(dl_dp=2.0 * (p - t), dl_dq=cq)
    return g, p, q


def _loss_and_path(sim, v0, w0, mass, mu, t, cq):
    p, q, path = sim.run(v0, w0, mass, mu, with_path=True)
    return float(((p - t) ** 2).sum() + (cq * q).sum()), path


def _fd_check(sim, sc, t, cq, g, comps):
    base = dict(v0=sc.v.copy(), w0=sc.w.copy(), mass=sc.mass.copy(), mu=sc.mu.copy())
    _, path0 = _loss_and_path(sim, base["v0"], base["w0"], base["mass"], base["mu"], t, cq)
    adj, fd, skipped = [], [], []
    for key, idx in comps:
        for hh in FD_STEPS:
            step = hh * base[key][idx] if key == "mass" else hh
            vals, same = [], True
            for sgn in (1.0, -1.0):
                a = {k: v.copy() for k, v in base.items()}
                a[key][idx] += sgn * step
                L, path = _loss_and_path(sim, a["v0"], a["w0"], a["mass"], a["mu"], t, cq)
                vals.append(L)
                same = same and np.array_equal(path, path0)
            if same:
                break
        if not same:
            skipped.append((key, idx))
            continue
        adj.append(float(g[key][idx]))
        fd.append((vals[0] - vals[1]) / (2.0 * step))
    adj, fd = np.array(adj), np.array(fd)
    return adj, fd, skipped


def test_adjoint_c4_matches_cpu_adjoint(engine_factory, c4):
    """Every gradient of a 3-step C4 rollout (hinge chain, pinned link, 48 cubes on a
    ramp; two checkpoint segments) against the oracle's CPU adjoint."""
    sc, t, cq = c4
    sim = Simulation(sc, 3, 2, engine=engine_factory(sc))
    g, p, q = _device_grads(sim, sc, t, cq)
    o = O.OracleScene(sc, 0)
    gc = o.adjoint(float(sc.dt[0]), 3, dl_dp=2.0 * (p - t), dl_dq=cq)
    po, qo = o.get_state()[:2]
    assert np.array_equal(po, p) and np.array_equal(qo, q)  # same trajectory, bit for bit
    for key in ("p0", "q0", "v0", "w0", "mass", "mu"):
        ref = gc[key]
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(g[key] - ref).max() / scale
        assert err <= CPU_RTOL, (key, err)


def test_adjoint_c4_matches_finite_differences(engine_factory, c4):
    """Central differences of a 3-step C4 rollout: free cubes' and chain links' initial
    velocities, angular velocities and masses."""
    sc, t, cq = c4
    sim = Simulation(sc, 3, 2, engine=engine_factory(sc))
    g, _, _ = _device_grads(sim, sc, t, cq)
    comps = [("v0", (b, c)) for b in (3, 9, 14, 20, 33, 47, 60) for c in range(3)]
    comps += [("w0", (b, c)) for b in (6, 25, 52) for c in range(3)]
    comps += [("mass", (b,)) for b in (4, 11, 15)]
    adj, fd, skipped = _fd_check(sim, sc, t, cq, g, comps)
    # the chain links' Newton counts change under any of the steps: they are pinned by
    # the CPU adjoint above and by the free chain below
    assert len(adj) >= 12, skipped
    err = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
    assert err <= FD_RTOL, (err, list(zip(adj, fd)))


def test_adjoint_hinge_chain_matches_finite_differences(engine_factory):
    """A free 3-link hinge chain (ALM stiffness 1e4, spinning links), 5 steps: all
    velocities, angular velocities and masses against central differences, and every
    gradient against the CPU adjoint."""
    sc = scenes.hinge_chain_scene()
    sim = Simulation(sc, 5, 2, engine=engine_factory(sc))
    rng = np.random.default_rng(14)
    t = sc.p + rng.normal(0.0, 0.1, size=sc.p.shape)
    cq = rng.normal(0.0, 0.1, size=(sc.n_bodies, 4))
    g, p, q = _device_grads(sim, sc, t, cq)
    comps = [(k, (b, c)) for k in ("v0", "w0") for b in range(sc.n_bodies) for c in range(3)]
    comps += [("mass", (b,)) for b in range(sc.n_bodies)]
    adj, fd, skipped = _fd_check(sim, sc, t, cq, g, comps)
    assert not skipped, skipped
    # Every solve here stops at the reference's Newton tolerance (|g|_inf < 1e-6) after
    # one or more steps, so the computed trajectory is the exact solution map only to
    # O(1e-6); finite differences see the derivative of that remainder, which the adjoint
    # (the derivative of the solution map) does not: 1.5e-5 normwise here, the same with
    # the CPU adjoint. Bar for this scene: 5e-5. Tightening the Newton tolerance does not
    # help: the ALM iteration counts then change under every perturbation.
    err = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
    assert err <= 5e-5, (err, list(zip(adj, fd)))
    o = O.OracleScene(sc, 0)
    gc = o.adjoint(float(sc.dt[0]), 5, dl_dp=2.0 * (p - t), dl_dq=cq)
    for key in ("p0", "q0", "v0", "w0", "mass"):
        ref = gc[key]
        err = np.abs(g[key] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err <= CPU_RTOL, (key, err)


def test_adjoint_c5_batch_matches_finite_differences(engine_factory):
    """Three C5 scenes (32 cubes thrown at a wall and a floor each) in one batch: per-scene
    gradients from one batched adjoint against central differences; mu gradient zero
    (no friction acts: the barrier is inactive)."""
    sc = scenes.pack(scenes.config5(3))
    sim = Simulation(sc, 8, 4, engine=engine_factory(sc))
    p_end = sim.run(sc.v, sc.w, sc.mass, sc.mu)[0]
    rng = np.random.default_rng(12)
    t = p_end + rng.normal(0.0, 0.05, size=p_end.shape)
    cq = rng.normal(0.0, 1e-2, size=(sc.n_bodies, 4))
    g, _, _ = _device_grads(sim, sc, t, cq)
    off = sc.scene_body_off
    comps = [("v0", (int(off[s]) + b, c)) for s in range(3) for b in (2, 19, 33) for c in range(3)]
    comps += [("w0", (int(off[s]) + 7, c)) for s in range(3) for c in range(3)]
    adj, fd, skipped = _fd_check(sim, sc, t, cq, g, comps)
    assert len(adj) >= len(comps) // 2, skipped
    err = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
    assert err <= FD_RTOL, err
    assert np.all(g["mu"] == 0.0)


def test_adjoint_checkpoint_interval_does_not_change_gradients(engine_factory, c4):
    """The backward recomputes each segment from its checkpoint; the recomputed forward is
    bitwise the recorded one, so the gradients do not depend on the interval."""
    sc, t, cq = c4
    sim = Simulation(sc, 5, 5, engine=engine_factory(sc))
    ref, _, _ = _device_grads(sim, sc, t, cq, every=5)
    for every in (1, 2, 3):
        g, _, _ = _device_grads(sim, sc, t, cq, every=every)
        for key in ref:
            assert np.array_equal(g[key], ref[key]), (every, key)


def test_rollout_autograd_function(engine_factory, c4):
    """torch.autograd through Rollout gives the engine's adjoint gradients."""
    sc, t, cq = c4
    sim = Simulation(sc, 2, 1, engine=engine_factory(sc))
    v0 = torch.tensor(sc.v, dtype=torch.float64, requires_grad=True)
    w0 = torch.tensor(sc.w, dtype=torch.float64, requires_grad=True)
    mass = torch.tensor(sc.mass, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(sc.mu, dtype=torch.float64, requires_grad=True)
    p, q = Rollout.apply(v0, w0, mass, mu, sim)
    loss = ((p - torch.from_numpy(t)) ** 2).sum() + (torch.from_numpy(cq) * q).sum()
    loss.backward()
    g, _, _ = _device_grads(sim, sc, t, cq)
    assert np.array_equal(v0.grad.numpy(), g["v0"])
    assert np.array_equal(w0.grad.numpy(), g["w0"])
    assert np.array_equal(mass.grad.numpy(), g["mass"])
    assert np.array_equal(mu.grad.numpy(), g["mu"])


def test_adjoint_contact_scene_matches_cpu_adjoint(engine_factory):
    """Millimetre cubes with active barriers (some distances at the 1e-9 clamp, stiffness
    ~5e11): the device adjoint against the CPU adjoint on the velocities, masses and mu.
    The step-start pose gradients also run through the first solve's contact geometry,
    where the adjoint vector's components along the contact normals are small differences
    of large cancelling terms weighted by b'' ~ 3e8; the CPU adjoint's Jacobian by
    differences resolves those components to a few tens of percent there, so the pose
    gradients are compared on the friction slide scene instead (next test)."""
    spec = scenes.micro_contact_scene(seed=0, n_cubes=4)
    sc = scenes.pack([spec])
    sim = Simulation(sc, 6, 3, engine=engine_factory(sc))
    rng = np.random.default_rng(13)
    p_end = sim.run(sc.v, sc.w, sc.mass, sc.mu)[0]
    t = p_end + rng.normal(0.0, 1e-5, size=p_end.shape)
    cq = rng.normal(0.0, 1e-3, size=(sc.n_bodies, 4))
    g, p, q = _device_grads(sim, sc, t, cq)
    o = O.OracleScene(sc, 0)
    gc = o.adjoint(float(sc.dt[0]), 6, dl_dp=2.0 * (p - t), dl_dq=cq)
    for key in ("v0", "w0", "mass", "mu"):
        ref = gc[key]
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(g[key] - ref).max() / scale
        assert err <= CPU_RTOL, (key, err)


def test_adjoint_friction_coefficient_matches_cpu_adjoint_and_differences(engine_factory):
    """Millimetre cubes sliding on a static plate with the lagged friction acting (the
    ALM loop iterates, normal forces and contact points come from the previous iterate),
    10 steps at Newton tolerance 1e-11 (every solve converges or accepts no step): every
    device gradient, the friction coefficient's included, against the CPU adjoint, and the
    friction coefficient's and the velocities' against central differences of the device
    forward within 1e-5. The CPU bar is 1e-5 here: with this target dL/dmu is a small
    difference of cancelling terms, which the CPU adjoint's differences resolve to ~3e-6
    (the device adjoint is within 1e-7 of the central differences)."""
    sc = scenes.friction_slide_scene()
    sim = Simulation(sc, 10, 5, engine=engine_factory(sc), newton_tol=1e-11, newton_max_iters=50)
    sim.barrier0 = (scenes.FRICTION_SLIDE_KAPPA, 1e-3, 1)
    p_end = sim.run(sc.v, sc.w, sc.mass, sc.mu)[0]
    rng = np.random.default_rng(31)
    t = p_end + rng.normal(0.0, 1e-5, size=p_end.shape)
    cq = np.zeros((sc.n_bodies, 4))
    g, p, q = _device_grads(sim, sc, t, cq)
    assert np.array_equal(p, p_end)
    assert g["mu"][0] != 0.0
    o = O.OracleScene(sc, 0)
    o.set_newton(1e-11, 50)
    o.set_barrier(scenes.FRICTION_SLIDE_KAPPA, 1e-3, 1)
    gc = o.adjoint(float(sc.dt[0]), 10, dl_dp=2.0 * (p - t))
    for key in ("p0", "q0", "v0", "w0", "mass", "mu"):
        ref = gc[key]
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(g[key] - ref).max() / scale
        assert err <= 1e-5, (key, err)
    adj, fd, skipped = _fd_check(sim, sc, t, cq, g, [("mu", (0,))])
    assert not skipped and abs(adj[0] - fd[0]) <= FD_RTOL * abs(fd[0]), (adj, fd)
    comps = [("v0", (b, c)) for b in (1, 2) for c in range(3)]
    adj, fd, skipped = _fd_check(sim, sc, t, cq, g, comps)
    assert not skipped
    assert np.linalg.norm(adj - fd) <= FD_RTOL * np.linalg.norm(fd), list(zip(adj, fd))
