"""The reference's solver interface for this path, backed by the device engine.

Mirrors NexDynIPC's World / RigidBody / FixedJoint / HingeJoint / IPCSolver
(nexdynipc/include/Dynamics/World.h:11-15, RigidBody.h:11-31, Joints/FixedJoint.h,
Joints/HingeJoint.h, IPCSolver.h:21-91) with the same names, argument meaning and
defaults: `IPCSolver.step(world, dt)` runs one IPCSolver::step of the world on the GPU
(nexdynipc_amd._lib.Engine over the C ABI) and writes the new state back into the
world's bodies. The world's structure (bodies, meshes, joints, allowed contact pairs) is
uploaded once and again only when it changes; the joints' ALM multipliers and stiffness
and the adaptive barrier's state are carried over an upload, as the reference keeps them
in its Joint objects and its solver.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

from . import scenes as _scenes
from ._lib import Engine, make_params


@dataclasses.dataclass(eq=False)
class RigidBody:
    """RigidBody.h:11-31; shape: (V[n, 3], F[m, 3]) body-frame triangle mesh or None."""
    id: int = -1
    name: str = ""
    position: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    orientation: np.ndarray = dataclasses.field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))  # w x y z
    velocity: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    angular_velocity: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    linear_acceleration: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    angular_acceleration: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    mass: float = 1.0
    inertia_body: np.ndarray = dataclasses.field(default_factory=lambda: np.eye(3))
    friction_coeff: float = 0.0
    restitution_coeff: float = 0.0
    shape: Optional[tuple] = None


@dataclasses.dataclass(eq=False)
class FixedJoint:
    """FixedJoint(body_id, target_position, target_orientation) (Joints/FixedJoint.h)."""
    body_id: int
    position: np.ndarray
    orientation: np.ndarray = dataclasses.field(default_factory=lambda: np.array([1.0, 0.0, 0.0, 0.0]))
    stiffness: float = 1e8  # Joint.h:44
    lambda_: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(6))

    def setStiffness(self, k):
        self.stiffness = float(k)

    def getStiffness(self):
        return self.stiffness


@dataclasses.dataclass(eq=False)
class HingeJoint:
    """HingeJoint(body_a_id, body_b_id, anchor_a, anchor_b, axis_a, axis_b) (Joints/HingeJoint.h)."""
    body_a_id: int
    body_b_id: int
    anchor_a: np.ndarray
    anchor_b: np.ndarray
    axis_a: np.ndarray
    axis_b: np.ndarray
    stiffness: float = 1e8
    lambda_: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(6))

    def setStiffness(self, k):
        self.stiffness = float(k)

    def getStiffness(self):
        return self.stiffness


@dataclasses.dataclass
class World:
    """World.h:11-15 (forms other than the solver's own are not part of this path)."""
    bodies: List[RigidBody] = dataclasses.field(default_factory=list)
    joints: list = dataclasses.field(default_factory=list)

    def addBody(self, body: RigidBody):
        self.bodies.append(body)

    def addJoint(self, joint):
        self.joints.append(joint)


def world_to_scene(world: World, dt: float, dhat: float, mu: float, friction: bool, allowed: dict) -> _scenes.Scenes:
    """The world as one scene of the upload layout. Joint bodies are world indices by
    body id (IPCSolver.cpp:966-1066: an unknown id maps to index 0)."""
    n = len(world.bodies)
    idx = {}
    for i, b in enumerate(world.bodies):
        idx[b.id] = i
    verts, faces, vert_off, n_verts, face_off, n_faces = [], [], [], [], [], []
    nv = nf = 0
    for b in world.bodies:
        vert_off.append(nv)
        face_off.append(nf)
        if b.shape is None:
            n_verts.append(0)
            n_faces.append(0)
            continue
        V = np.asarray(b.shape[0], dtype=np.float64).reshape(-1, 3)
        F = np.asarray(b.shape[1], dtype=np.int32).reshape(-1, 3)
        verts.append(V)
        faces.append(F)
        n_verts.append(len(V))
        n_faces.append(len(F))
        nv += len(V)
        nf += len(F)
    jt, jb, jd, jm = [], [], [], []
    for j in world.joints:
        d = np.zeros(16)
        if isinstance(j, FixedJoint):
            jt.append(0)
            jb.append([idx.get(j.body_id, 0), -1])
            d[0:3] = j.position
            d[3:7] = j.orientation
        elif isinstance(j, HingeJoint):
            jt.append(1)
            jb.append([idx.get(j.body_a_id, 0), idx.get(j.body_b_id, 0)])
            d[0:3], d[3:6], d[6:9], d[9:12] = j.anchor_a, j.anchor_b, j.axis_a, j.axis_b
        else:
            raise NotImplementedError(f"{type(j).__name__}: only FixedJoint and HingeJoint run on the device path")
        jd.append(d)
        jm.append(j.stiffness)
    keys = list(allowed.keys())
    return _scenes.Scenes(
        n_scenes=1, scene_body_off=np.array([0, n], dtype=np.int32),
        body_id=np.array([b.id for b in world.bodies], dtype=np.int32),
        vert_off=np.array(vert_off, dtype=np.int32), n_verts=np.array(n_verts, dtype=np.int32),
        face_off=np.array(face_off, dtype=np.int32), n_faces=np.array(n_faces, dtype=np.int32),
        verts=np.ascontiguousarray(np.concatenate(verts).reshape(-1)) if verts else np.zeros(0),
        faces=np.ascontiguousarray(np.concatenate(faces).reshape(-1)) if faces else np.zeros(0, dtype=np.int32),
        mass=np.array([b.mass for b in world.bodies], dtype=np.float64),
        inertia=np.ascontiguousarray(np.stack([np.asarray(b.inertia_body, dtype=np.float64).reshape(9)
                                               for b in world.bodies]).reshape(-1)) if n else np.zeros(0),
        scene_joint_off=np.array([0, len(jt)], dtype=np.int32),
        joint_type=np.array(jt, dtype=np.int32), joint_body=np.array(jb, dtype=np.int32).reshape(-1),
        joint_data=np.array(jd, dtype=np.float64).reshape(-1), joint_mu=np.array(jm, dtype=np.float64),
        scene_allowed_off=np.array([0, len(keys)], dtype=np.int32),
        allowed_pairs=np.array(keys, dtype=np.int32).reshape(-1),
        allowed_mu=np.array([allowed[k] for k in keys], dtype=np.float64),
        p=np.array([b.position for b in world.bodies], dtype=np.float64).reshape(n, 3),
        q=np.array([b.orientation for b in world.bodies], dtype=np.float64).reshape(n, 4),
        v=np.array([b.velocity for b in world.bodies], dtype=np.float64).reshape(n, 3),
        w=np.array([b.angular_velocity for b in world.bodies], dtype=np.float64).reshape(n, 3),
        dt=np.array([dt]), dhat=np.array([dhat]), mu=np.array([mu]),
        enable_friction=np.array([1 if friction else 0], dtype=np.int32), names=["world"])


class IPCSolver:
    """IPCSolver (IPCSolver.h:21-91) on the GPU. Defaults as the reference's: CCD on
    (safety 0.8), adaptive barrier on, friction off (coefficient 0), dhat 1e-3, ALM 10
    iterations with tolerances 1e-4 / 1e-3 / 1e-3, hardening trigger 2 and ratio 2."""

    def __init__(self, device: int = 0):
        self._device = device
        self._engine: Optional[Engine] = None
        self._sig = None
        self._dirty_params = True
        self.enable_ccd_ = True
        self.ccd_safety_factor_ = 0.8
        self.enable_adaptive_barrier_ = True
        self.enable_friction_ = False
        self.friction_coeff_ = 0.0
        self.dhat_ = 1e-3
        self.kappa_ = 0.0
        self.alm_max_iters_ = 10
        self.alm_constraint_tolerance_ = 1e-4
        self.alm_dual_tolerance_ = 1e-3
        self.alm_contact_tolerance_ = 1e-3
        self.alm_hardening_trigger_ = 2.0
        self.alm_hardening_ratio_ = 2.0
        self.allowed_ = {}  # (id_a, id_b) -> friction override (< 0: none)
        self._resid = np.zeros(3)

    # ---- settings (IPCSolver.h:38-79)
    def _set(self, name, value):
        setattr(self, name, value)
        self._dirty_params = True

    def enableCCD(self, enable):
        self._set("enable_ccd_", bool(enable))

    def isCCDEnabled(self):
        return self.enable_ccd_

    def setCCDSafetyFactor(self, factor):
        self._set("ccd_safety_factor_", float(factor))

    def ccdSafetyFactor(self):
        return self.ccd_safety_factor_

    def enableAdaptiveBarrier(self, enable):
        self._set("enable_adaptive_barrier_", bool(enable))

    def isAdaptiveBarrierEnabled(self):
        return self.enable_adaptive_barrier_

    def enableFriction(self, enable):
        self._set("enable_friction_", bool(enable))

    def isFrictionEnabled(self):
        return self.enable_friction_

    def setFrictionCoefficient(self, mu):
        self._set("friction_coeff_", float(mu))

    def frictionCoefficient(self):
        return self.friction_coeff_

    def setContactDistanceThreshold(self, dhat):
        self._set("dhat_", max(1e-9, float(dhat)))

    def setContactStiffness(self, stiffness):
        if stiffness > 0.0:
            self._set("kappa_", float(stiffness))

    def clearAllowedContactPairs(self):
        self.allowed_ = {}

    def addAllowedContactPair(self, body_id_a, body_id_b, friction=-1.0):
        """IPCSolver.cpp:275-281: an unordered pair; a friction >= 0 sets the pair's
        override, a negative one leaves an earlier override in place."""
        a, b = sorted((int(body_id_a), int(body_id_b)))
        if friction >= 0.0 or (a, b) not in self.allowed_:
            self.allowed_[(a, b)] = float(friction) if friction >= 0.0 else -1.0

    def setALMMaxIters(self, n):
        self._set("alm_max_iters_", max(1, int(n)))

    def setALMConstraintTolerance(self, tol):
        self._set("alm_constraint_tolerance_", max(1e-12, float(tol)))

    def setALMDualTolerance(self, tol):
        self._set("alm_dual_tolerance_", max(1e-12, float(tol)))

    def setALMContactTolerance(self, tol):
        self._set("alm_contact_tolerance_", max(1e-12, float(tol)))

    def setALMHardeningTrigger(self, trigger):
        self._set("alm_hardening_trigger_", max(1.0, float(trigger)))

    def setALMHardeningRatio(self, ratio):
        self._set("alm_hardening_ratio_", max(1.0, float(ratio)))

    # ---- last-step diagnostics (IPCSolver.h:74-79); drives are not part of this path
    def lastAngularVelocityError(self):
        return 0.0

    def lastTorqueSaturationRatio(self):
        return 0.0

    def lastMaxConstraintViolation(self):
        return float(self._resid[0])

    def lastDualResidual(self):
        return float(self._resid[1])

    def lastContactResidual(self):
        return float(self._resid[2])

    # ---- stepping
    def _params(self):
        p = make_params(dhat=self.dhat_, mu=self.friction_coeff_, friction=self.enable_friction_,
                        ccd=self.enable_ccd_, adaptive=self.enable_adaptive_barrier_,
                        alm_max_iters=self.alm_max_iters_, kappa=self.kappa_)
        p.alm_constraint_tol = self.alm_constraint_tolerance_
        p.alm_dual_tol = self.alm_dual_tolerance_
        p.alm_contact_tol = self.alm_contact_tolerance_
        p.alm_hardening_trigger = self.alm_hardening_trigger_
        p.alm_hardening_ratio = self.alm_hardening_ratio_
        p.ccd_safety_factor = self.ccd_safety_factor_
        return p

    def _signature(self, world):
        return (tuple((b.id, id(b.shape), b.shape is None or len(b.shape[0]), float(b.mass),
                       np.asarray(b.inertia_body, dtype=np.float64).tobytes()) for b in world.bodies),
                tuple(id(j) for j in world.joints), tuple(sorted(self.allowed_.items())))

    def step(self, world: World, dt: float):
        """IPCSolver::step(world, dt) (IPCSolver.cpp:808-1231)."""
        if not world.bodies:
            return
        sig = self._signature(world)
        if self._engine is None or sig != self._sig:
            carry = None
            if self._engine is not None:
                carry = self._engine.barrier()
                self._engine.close()
            sc = world_to_scene(world, dt, self.dhat_, self.friction_coeff_, self.enable_friction_, self.allowed_)
            self._engine = Engine(self._device)
            self._engine.upload(sc, self._params())
            if world.joints:
                self._engine.set_joint_state(np.array([j.lambda_ for j in world.joints], dtype=np.float64),
                                             np.array([j.stiffness for j in world.joints], dtype=np.float64))
            if carry is not None:
                self._engine.set_barrier(*carry)
            self._sig = sig
            self._dirty_params = False
        elif self._dirty_params:
            self._engine.configure(self._params())
            self._dirty_params = False
        B = world.bodies
        self._engine.set_state(np.array([b.position for b in B], dtype=np.float64),
                               np.array([b.orientation for b in B], dtype=np.float64),
                               np.array([b.velocity for b in B], dtype=np.float64),
                               np.array([b.angular_velocity for b in B], dtype=np.float64),
                               np.array([b.linear_acceleration for b in B], dtype=np.float64),
                               np.array([b.angular_acceleration for b in B], dtype=np.float64))
        self._engine.step(1, dt=dt)
        p, q, v, w, al, aa = self._engine.get_state()
        for i, b in enumerate(B):
            b.position, b.orientation, b.velocity = p[i].copy(), q[i].copy(), v[i].copy()
            b.angular_velocity, b.linear_acceleration, b.angular_acceleration = w[i].copy(), al[i].copy(), aa[i].copy()
        if world.joints:
            lam, mu = self._engine.joint_state()
            for j, l, m in zip(world.joints, lam, mu):
                j.lambda_ = l.copy()
                j.stiffness = float(m)
        self._resid = self._engine.last_residuals()[0]

    def close(self):
        if self._engine is not None:
            self._engine.close()
            self._engine = None
