"""Differentiable rollouts: a torch.autograd.Function over the device simulation.

`Rollout.apply(v0, w0, mass, mu, sim)` runs `sim.n_steps` steps of every scene from the
scene's initial poses with the given initial linear / angular velocities ([n, 3] each),
body masses ([n]) and per-scene friction coefficients ([S]), and returns the final body
positions [n, 3] and orientations [n, 4]. Its backward is the device adjoint
(ndipc_adjoint_backward): the implicit-function theorem at every converged Newton step,
with the step's Hessian solved on the device, checkpoints every `sim.every` steps and
forward recompute of each segment before it is swept back.

The reference has no differentiation; this is the interface of the adjoint built on its
contact stage. Tensors are float64 on any device; the engine copies them to and from the
host at the call boundary (the simulation itself stays on the GPU).
"""
import numpy as np
import torch

from ._lib import Engine, make_params


class Simulation:
    """An engine with a scene batch uploaded, reset to the scene's initial state (and the
    joint and barrier state of that moment) before every rollout."""

    def __init__(self, sc, n_steps, every=10, engine=None, newton_tol=None, newton_max_iters=None):
        self.sc = sc
        self.n_steps = int(n_steps)
        self.every = int(every)
        self.e = engine if engine is not None else Engine(0)
        params = None
        if newton_tol is not None or newton_max_iters is not None:
            # a tighter Newton solve than the reference's (1e-6, 20 iterations) makes the
            # computed step the exact solution map that the adjoint differentiates
            params = make_params(dhat=float(sc.dhat[0]), mu=float(sc.mu[0]), friction=bool(sc.enable_friction[0]),
                                 newton_tol=newton_tol or 1e-6, newton_max_iters=newton_max_iters or 20)
        self.e.upload(sc, params)
        self.lam0, self.jmu0 = self.e.joint_state()
        self.barrier0 = self.e.barrier()

    def reset(self, v0, w0, mass, mu):
        n = self.sc.n_bodies
        self.e.set_state(self.sc.p, self.sc.q, v0, w0, np.zeros((n, 3)), np.zeros((n, 3)))
        self.e.set_mass(mass)
        self.e.set_mu(mu)
        if self.sc.n_joints:
            self.e.set_joint_state(self.lam0, self.jmu0)
        self.e.set_barrier(*self.barrier0)

    def run(self, v0, w0, mass, mu, with_path=False):
        """Forward only (numpy in, numpy out): final positions and orientations. With
        `with_path`, also the discrete path of the run: the ALM and Newton iteration
        counts of every scene in every step ([n_steps, S, 2])."""
        self.reset(v0, w0, mass, mu)
        if not with_path:
            self.e.step(self.n_steps)
            p, q = self.e.get_state()[:2]
            return p, q
        path = np.zeros((self.n_steps, self.sc.n_scenes, 2))
        for k in range(self.n_steps):
            self.e.step(1)
            path[k] = self.e.stats()[:, :2]
        p, q = self.e.get_state()[:2]
        return p, q, path


def _np(t):
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, w0, mass, mu, sim):
        sim.reset(_np(v0), _np(w0), _np(mass), _np(mu))
        sim.e.adjoint_forward(sim.n_steps, sim.every)
        p, q = sim.e.get_state()[:2]
        ctx.sim = sim
        ctx.dev = v0.device
        return (torch.from_numpy(p).to(v0.device), torch.from_numpy(q).to(v0.device))

    @staticmethod
    def backward(ctx, gp, gq):
        sim = ctx.sim
        n = sim.sc.n_bodies
        dl_dp = _np(gp) if gp is not None else np.zeros((n, 3))
        dl_dq = _np(gq) if gq is not None else np.zeros((n, 4))
        g = sim.e.adjoint_backward(dl_dp=dl_dp, dl_dq=dl_dq)
        to = lambda a: torch.from_numpy(a).to(ctx.dev)
        return to(g["v0"]), to(g["w0"]), to(g["mass"]), to(g["mu"]), None
