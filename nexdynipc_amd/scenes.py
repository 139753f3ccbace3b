"""Synthetic rigid-body scenes for the NexDynIPC contact stage.

Scenes are flat numpy arrays in the layout of ``ndipc_scene_desc``
(include/ndipc.h): bodies of all scenes are concatenated, each body points at a
triangle mesh (body-frame vertices, faces), and static bodies are pinned with
FixedJoint constraints, the reference's idiom for static geometry
(nexdynipc/src/App/SceneLoader.cpp:593-596).

Gravity is hard-coded to -Y in the reference (nexdynipc/src/Dynamics/IPCSolver.cpp:142),
so every scene is Y-up. Draws use numpy's PCG64 generator seeded with
``20261203 + 1000 * config + scene``.

Configurations (sizes, seeds and draw order are part of the benchmark definition):
  c1  8 unit cubes in a 2x2x2 grid (1.1 m spacing, +-5 deg random tilt) 0.5 m above
      a static 10x1x10 m ground box; dt 1e-2, dhat 1e-3, mu 0.3.
  c2  512 icospheres (subdivision 3, 642 verts / 1280 faces), r ~ U[0.1, 0.3],
      non-overlapping in a 4 x 8 (y) x 4 m region (y in [2.5, 10.5]) above an open-top
      container of 5 static boxes (4.4 x 0.2 x 4.4 floor, 2 m walls, 1 cm gaps between
      parts); v ~ N(0, 0.5) per axis; dt 5e-3, dhat 1e-3, mu 0.4.
  c3  256 scenes, n ~ round(exp(U[ln 4, ln 128])) bodies each, shapes uniform over
      {unit cube, icosphere subdivision 2 (r 0.5), 2x1x0.5 box}, dropped from
      U[0.2, 2] m onto a static ground box (0.7 m thick, so no body's vertices sit on
      its faces when the centres align); dt 1e-2, mu ~ U[0.1, 0.6] per scene.
  c4  16-link hinge chain (0.4x0.1x0.1 links 5 cm apart, hinges mid-gap, link 0 pinned)
      + 48 unit cubes above a
      static 20 deg ramp box; dt 1e-2, mu 0.3.
  c5  128 scenes of 32 unit cubes (grid jitter +-0.05 m, v ~ N(0, 1)) thrown at a
      static wall box and floor box; dt 1e-2, mu ~ U[0.1, 0.8] per scene.
All bodies have density 1000 kg/m^3.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

DENSITY = 1000.0
SEED_BASE = 20261203


# --------------------------------------------------------------------------- meshes
def box_mesh(sx: float, sy: float, sz: float):
    """Axis-aligned box centred at the origin: 8 vertices, 12 outward triangles."""
    hx, hy, hz = 0.5 * sx, 0.5 * sy, 0.5 * sz
    V = np.array([[-hx, -hy, -hz], [hx, -hy, -hz], [hx, hy, -hz], [-hx, hy, -hz],
                  [-hx, -hy, hz], [hx, -hy, hz], [hx, hy, hz], [-hx, hy, hz]], dtype=np.float64)
    F = np.array([[0, 2, 1], [0, 3, 2],   # -z
                  [4, 5, 6], [4, 6, 7],   # +z
                  [0, 1, 5], [0, 5, 4],   # -y
                  [3, 7, 6], [3, 6, 2],   # +y
                  [0, 4, 7], [0, 7, 3],   # -x
                  [1, 2, 6], [1, 6, 5]],  # +x
                 dtype=np.int32)
    return V, F


def icosphere_mesh(subdiv: int, radius: float):
    """Icosahedron refined ``subdiv`` times by edge midpoints projected to the sphere."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
             [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
             [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    verts = [list(np.asarray(v, dtype=np.float64) / np.linalg.norm(v)) for v in verts]
    faces = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
             [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
             [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
             [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key in cache:
                return cache[key]
            m = np.asarray(verts[a]) + np.asarray(verts[b])
            m = m / np.linalg.norm(m)
            verts.append(list(m))
            cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        faces = nf
    V = np.asarray(verts, dtype=np.float64) * radius
    F = np.asarray(faces, dtype=np.int32)
    return V, F


def grid_box_mesh(sx: float, sy: float, sz: float, m: int):
    """Axis-aligned box centred at the origin whose faces are m x m grids of cells, two
    outward triangles each: 6 (m-1)^2 + 12 (m-1) + 8 vertices, 12 m^2 faces. Vertices
    sit at (L / m - 1/2) * size on the integer lattice L in [0, m]^3, so the six face
    planes are exactly +-size / 2."""
    size = np.array([sx, sy, sz], dtype=np.float64)
    index = {}
    verts = []

    def vid(L):
        key = tuple(L)
        if key not in index:
            index[key] = len(verts)
            verts.append((np.asarray(L, dtype=np.float64) / m - 0.5) * size)
        return index[key]

    faces = []
    for a in range(3):
        u, v = (a + 1) % 3, (a + 2) % 3
        for side in (0, m):
            for i in range(m):
                for j in range(m):
                    c = []
                    for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        L = [0, 0, 0]
                        L[a], L[u], L[v] = side, i + di, j + dj
                        c.append(vid(L))
                    # e_u x e_v = e_a: counter-clockwise seen from +a
                    if side == m:
                        faces += [[c[0], c[1], c[2]], [c[0], c[2], c[3]]]
                    else:
                        faces += [[c[0], c[2], c[1]], [c[0], c[3], c[2]]]
    return np.asarray(verts, dtype=np.float64), np.asarray(faces, dtype=np.int32)


def box_mass_properties(sx, sy, sz, density=DENSITY):
    m = density * sx * sy * sz
    I = np.diag([m * (sy * sy + sz * sz) / 12.0, m * (sx * sx + sz * sz) / 12.0, m * (sx * sx + sy * sy) / 12.0])
    return m, I


def polyhedron_mass_properties(V, F, density=DENSITY):
    """Mass and inertia about the body origin of a closed, outward-oriented mesh
    (signed tetrahedra from the origin, canonical covariance 1/120 [[2,1,1],[1,2,1],[1,1,2]])."""
    S = np.array([[2.0, 1.0, 1.0], [1.0, 2.0, 1.0], [1.0, 1.0, 2.0]]) / 120.0
    C = np.zeros((3, 3))
    vol = 0.0
    for a, b, c in F:
        A = np.stack([V[a], V[b], V[c]], axis=1)
        det = float(np.linalg.det(A))
        vol += det / 6.0
        C += det * (A @ S @ A.T)
    m = density * vol
    Cm = density * C
    I = np.trace(Cm) * np.eye(3) - Cm
    return m, I


# ------------------------------------------------------------------------ rotations
def quat_from_axis_angle(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    h = 0.5 * angle
    s = math.sin(h)
    return np.array([math.cos(h), s * axis[0], s * axis[1], s * axis[2]])


def random_tilt(rng, max_deg):
    axis = rng.normal(size=3)
    ang = math.radians(rng.uniform(-max_deg, max_deg))
    return quat_from_axis_angle(axis, ang)


def rotate(q, v):
    w, x, y, z = q
    qv = np.array([x, y, z])
    uv = np.cross(qv, v)
    uv = uv + uv
    return v + w * uv + np.cross(qv, uv)


# --------------------------------------------------------------------------- scenes
@dataclasses.dataclass
class BodySpec:
    mesh_key: str
    V: np.ndarray
    F: np.ndarray
    mass: float
    inertia: np.ndarray
    p: np.ndarray
    q: np.ndarray
    v: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    w: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    static: bool = False


@dataclasses.dataclass
class SceneSpec:
    bodies: List[BodySpec]
    dt: float
    dhat: float
    mu: float
    hinges: list = dataclasses.field(default_factory=list)  # (a, b, anchorA, anchorB, axisA, axisB)
    enable_friction: bool = True
    name: str = ""


@dataclasses.dataclass
class Scenes:
    """Flattened scene batch, field-for-field the ndipc_scene_desc arrays plus state."""
    n_scenes: int
    scene_body_off: np.ndarray
    body_id: np.ndarray
    vert_off: np.ndarray
    n_verts: np.ndarray
    face_off: np.ndarray
    n_faces: np.ndarray
    verts: np.ndarray
    faces: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    scene_joint_off: np.ndarray
    joint_type: np.ndarray
    joint_body: np.ndarray
    joint_data: np.ndarray
    joint_mu: np.ndarray
    scene_allowed_off: np.ndarray
    allowed_pairs: np.ndarray
    allowed_mu: np.ndarray
    p: np.ndarray
    q: np.ndarray
    v: np.ndarray
    w: np.ndarray
    dt: np.ndarray
    dhat: np.ndarray
    mu: np.ndarray
    enable_friction: np.ndarray
    names: list

    @property
    def n_bodies(self):
        return int(self.scene_body_off[-1])

    @property
    def n_joints(self):
        return int(self.scene_joint_off[-1])

    def n_joints_in(self, s):
        return int(self.scene_joint_off[s + 1] - self.scene_joint_off[s])

    def scene_slice(self, s):
        return slice(int(self.scene_body_off[s]), int(self.scene_body_off[s + 1]))

    def total_verts(self):
        """Vertices over all bodies (shared meshes counted once per body)."""
        return int(self.n_verts.sum())


def pack(specs: List[SceneSpec]) -> Scenes:
    mesh_pool = {}
    verts, faces = [], []
    nv_tot = nf_tot = 0
    body_off = [0]
    joint_off = [0]
    cols = {k: [] for k in ("body_id", "vert_off", "n_verts", "face_off", "n_faces", "mass", "inertia",
                            "p", "q", "v", "w")}
    jt, jb, jd, jm = [], [], [], []
    for sp in specs:
        for bi, b in enumerate(sp.bodies):
            if b.mesh_key not in mesh_pool:
                mesh_pool[b.mesh_key] = (nv_tot, len(b.V), nf_tot, len(b.F))
                verts.append(np.asarray(b.V, dtype=np.float64))
                faces.append(np.asarray(b.F, dtype=np.int32))
                nv_tot += len(b.V)
                nf_tot += len(b.F)
            vo, nv, fo, nf = mesh_pool[b.mesh_key]
            cols["body_id"].append(bi)
            cols["vert_off"].append(vo)
            cols["n_verts"].append(nv)
            cols["face_off"].append(fo)
            cols["n_faces"].append(nf)
            cols["mass"].append(b.mass)
            cols["inertia"].append(np.asarray(b.inertia, dtype=np.float64).reshape(9))
            cols["p"].append(b.p)
            cols["q"].append(b.q)
            cols["v"].append(b.v)
            cols["w"].append(b.w)
            if b.static:
                jt.append(0)
                jb.append([bi, -1])
                d = np.zeros(16)
                d[0:3] = b.p
                d[3:7] = b.q
                jd.append(d)
                jm.append(1e8)
        for (a, b2, anA, anB, axA, axB) in sp.hinges:
            jt.append(1)
            jb.append([a, b2])
            d = np.zeros(16)
            d[0:3], d[3:6], d[6:9], d[9:12] = anA, anB, axA, axB
            jd.append(d)
            jm.append(1e8)
        body_off.append(body_off[-1] + len(sp.bodies))
        joint_off.append(len(jt))
    n = body_off[-1]
    S = len(specs)
    return Scenes(
        n_scenes=S,
        scene_body_off=np.asarray(body_off, dtype=np.int32),
        body_id=np.asarray(cols["body_id"], dtype=np.int32),
        vert_off=np.asarray(cols["vert_off"], dtype=np.int32),
        n_verts=np.asarray(cols["n_verts"], dtype=np.int32),
        face_off=np.asarray(cols["face_off"], dtype=np.int32),
        n_faces=np.asarray(cols["n_faces"], dtype=np.int32),
        verts=np.ascontiguousarray(np.concatenate(verts).reshape(-1)),
        faces=np.ascontiguousarray(np.concatenate(faces).reshape(-1).astype(np.int32)),
        mass=np.asarray(cols["mass"], dtype=np.float64),
        inertia=np.ascontiguousarray(np.stack(cols["inertia"]).reshape(-1)),
        scene_joint_off=np.asarray(joint_off, dtype=np.int32),
        joint_type=np.asarray(jt, dtype=np.int32).reshape(-1),
        joint_body=np.asarray(jb, dtype=np.int32).reshape(-1),
        joint_data=np.asarray(jd, dtype=np.float64).reshape(-1) if jd else np.zeros(0),
        joint_mu=np.asarray(jm, dtype=np.float64),
        scene_allowed_off=np.zeros(S + 1, dtype=np.int32),
        allowed_pairs=np.zeros(0, dtype=np.int32),
        allowed_mu=np.zeros(0, dtype=np.float64),
        p=np.ascontiguousarray(np.stack(cols["p"]).reshape(n, 3)),
        q=np.ascontiguousarray(np.stack(cols["q"]).reshape(n, 4)),
        v=np.ascontiguousarray(np.stack(cols["v"]).reshape(n, 3)),
        w=np.ascontiguousarray(np.stack(cols["w"]).reshape(n, 3)),
        dt=np.asarray([sp.dt for sp in specs]),
        dhat=np.asarray([sp.dhat for sp in specs]),
        mu=np.asarray([sp.mu for sp in specs]),
        enable_friction=np.asarray([1 if sp.enable_friction else 0 for sp in specs], dtype=np.int32),
        names=[sp.name for sp in specs],
    )


_IDENT = np.array([1.0, 0.0, 0.0, 0.0])


def _box_body(key, sx, sy, sz, p, q=_IDENT, static=False, v=None, w=None):
    V, F = box_mesh(sx, sy, sz)
    m, I = box_mass_properties(sx, sy, sz)
    return BodySpec(key, V, F, m, I, np.asarray(p, dtype=np.float64), np.asarray(q, dtype=np.float64),
                    np.zeros(3) if v is None else np.asarray(v, dtype=np.float64),
                    np.zeros(3) if w is None else np.asarray(w, dtype=np.float64), static)


_ICO_UNIT = {}


def _ico_body(subdiv, radius, p, q=_IDENT, v=None):
    """Icosphere of the given radius: the unit mesh scaled by r (mass ~ r^3, inertia ~ r^5)."""
    if subdiv not in _ICO_UNIT:
        V, F = icosphere_mesh(subdiv, 1.0)
        m, I = polyhedron_mass_properties(V, F)
        _ICO_UNIT[subdiv] = (V, F, m, I)
    V1, F, m1, I1 = _ICO_UNIT[subdiv]
    r = float(radius)
    return BodySpec(f"ico{subdiv}_{r!r}", V1 * r, F, m1 * r ** 3, I1 * r ** 5, np.asarray(p, dtype=np.float64),
                    np.asarray(q, dtype=np.float64), np.zeros(3) if v is None else np.asarray(v, dtype=np.float64))


def config1(scene=0) -> SceneSpec:
    rng = np.random.default_rng(SEED_BASE + 1000 * 1 + scene)
    bodies = [_box_body("ground10x1x10", 10.0, 1.0, 10.0, [0.0, -0.5, 0.0], static=True)]
    for iy in range(2):
        for ix in range(2):
            for iz in range(2):
                q = random_tilt(rng, 5.0)
                p = [(ix - 0.5) * 1.1, 1.0 + 1.1 * iy, (iz - 0.5) * 1.1]
                bodies.append(_box_body("cube1", 1.0, 1.0, 1.0, p, q))
    return SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.3, name="c1")


def container_boxes(half=2.0, wall_h=2.0, t=0.2, gap=0.01):
    """Open-top container as 5 static boxes: a floor and four walls around [-half, half]^2.

    The parts are separated by `gap` (> 2 dhat) so no two static parts touch or overlap:
    touching or interpenetrating static boxes produce contacts whose normal can be
    orthogonal to the centre offset, which the reference's centre-distance barrier
    clamps to d = 1e-9 (DistanceBarrierForm.cpp:89-98)."""
    L = 2 * half + 2 * t
    return [
        _box_body(f"floor{L}", L, t, L, [0.0, -0.5 * t, 0.0], static=True),
        _box_body(f"wallx{wall_h}", t, wall_h, 2 * half - 2 * gap, [half + 0.5 * t, gap + 0.5 * wall_h, 0.0],
                  static=True),
        _box_body(f"wallx{wall_h}", t, wall_h, 2 * half - 2 * gap, [-half - 0.5 * t, gap + 0.5 * wall_h, 0.0],
                  static=True),
        _box_body(f"wallz{wall_h}", L, wall_h, t, [0.0, gap + 0.5 * wall_h, half + 0.5 * t + gap], static=True),
        _box_body(f"wallz{wall_h}", L, wall_h, t, [0.0, gap + 0.5 * wall_h, -half - 0.5 * t - gap], static=True),
    ]


def config2(scene=0, n_spheres=512, subdiv=3) -> SceneSpec:
    rng = np.random.default_rng(SEED_BASE + 1000 * 2 + scene)
    dhat = 1e-3
    bodies = container_boxes()
    centres, radii = [], []
    margin = 0.01
    tries = 0
    while len(centres) < n_spheres:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place spheres")
        r = rng.uniform(0.1, 0.3)
        c = np.array([rng.uniform(-2 + r + margin, 2 - r - margin),
                      rng.uniform(2.5 + r, 10.5 - r),
                      rng.uniform(-2 + r + margin, 2 - r - margin)])
        if centres:
            C = np.asarray(centres)
            R = np.asarray(radii)
            if np.any(np.linalg.norm(C - c, axis=1) <= R + r + 2 * dhat):
                continue
        centres.append(c)
        radii.append(r)
    for c, r in zip(centres, radii):
        v = rng.normal(0.0, 0.5, size=3)
        bodies.append(_ico_body(subdiv, r, c, v=v))
    return SceneSpec(bodies, dt=5e-3, dhat=dhat, mu=0.4, name="c2")


def config3(n_scenes=256, bodies_override: Optional[int] = None) -> List[SceneSpec]:
    out = []
    for s in range(n_scenes):
        rng = np.random.default_rng(SEED_BASE + 1000 * 3 + s)
        n = int(round(math.exp(rng.uniform(math.log(4), math.log(128)))))
        if bodies_override is not None:
            n = bodies_override
        mu = rng.uniform(0.1, 0.6)
        side = int(math.ceil(math.sqrt(n)))
        spacing = 2.5
        extent = side * spacing + 2.0
        bodies = [_box_body(f"ground{extent}", extent, 0.7, extent, [0.0, -0.35, 0.0], static=True)]
        for k in range(n):
            shape = int(rng.integers(0, 3))
            h = rng.uniform(0.2, 2.0)
            gx = (k % side - 0.5 * (side - 1)) * spacing
            gz = (k // side - 0.5 * (side - 1)) * spacing
            if shape == 0:
                bodies.append(_box_body("cube1", 1.0, 1.0, 1.0, [gx, h + 0.5, gz]))
            elif shape == 1:
                bodies.append(_ico_body(2, 0.5, [gx, h + 0.5, gz]))
            else:
                bodies.append(_box_body("box2x1x05", 2.0, 1.0, 0.5, [gx, h + 0.5, gz]))
        out.append(SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=mu, name=f"c3_{s}"))
    return out


def config4(scene=0):
    """Returns (SceneSpec, targets[n_bodies, 3])."""
    rng = np.random.default_rng(SEED_BASE + 1000 * 4 + scene)
    ang = math.radians(20.0)
    qr = quat_from_axis_angle([0, 0, 1], ang)
    bodies = [_box_body("ramp", 12.0, 0.5, 6.0, [0.0, -0.25, 0.0], qr, static=True)]
    base = len(bodies)
    for k in range(16):
        p = [-4.0 + 0.45 * k, 2.0, -1.5]  # 5 cm gaps: adjacent links never touch
        bodies.append(_box_body("link", 0.4, 0.1, 0.1, p, static=False))
    hinges = []
    for k in range(15):
        hinges.append((base + k, base + k + 1, np.array([0.225, 0, 0]), np.array([-0.225, 0, 0]),
                       np.array([0, 0, 1.0]), np.array([0, 0, 1.0])))
    # link 0 pinned with a FixedJoint (a world-anchored hinge is mis-indexed by the
    # reference: body_id_to_idx[-1] default-inserts 0, IPCSolver.cpp:972-973).
    bodies[base].static = True
    for k in range(48):
        gx = -4.0 + 1.3 * (k % 8)
        gz = 0.5 + 1.3 * (k // 8) - 3.0
        p = [gx, 4.0 + 0.2 * (k % 3), gz]
        bodies.append(_box_body("cube1", 1.0, 1.0, 1.0, p, random_tilt(rng, 5.0)))
    targets = np.stack([b.p for b in bodies]) + rng.normal(0.0, 0.1, size=(len(bodies), 3))
    return SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.3, hinges=hinges, name="c4"), targets


def config5(n_scenes=128):
    out = []
    for s in range(n_scenes):
        rng = np.random.default_rng(SEED_BASE + 1000 * 5 + s)
        mu = rng.uniform(0.1, 0.8)
        bodies = [_box_body("floor20", 20.0, 1.0, 20.0, [0.0, -0.5, 0.0], static=True),
                  _box_body("wall", 1.0, 6.0, 20.0, [8.0, 3.0, 0.0], static=True)]
        for k in range(32):
            gx = (k % 4) * 1.6 - 2.4 + rng.uniform(-0.05, 0.05)
            gy = 1.0 + (k // 16) * 1.6 + rng.uniform(-0.05, 0.05)
            gz = ((k // 4) % 4) * 1.6 - 2.4 + rng.uniform(-0.05, 0.05)
            v = rng.normal(0.0, 1.0, size=3)
            bodies.append(_box_body("cube1", 1.0, 1.0, 1.0, [gx, gy, gz], v=v))
        out.append(SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=mu, name=f"c5_{s}"))
    return out


def ipc_drop_scene(nx=4, nz=4, layers=2, size=0.2, gap=0.05, tilt_deg=6.0, seed=0) -> Scenes:
    """A grid of nx x nz x layers cubes (edge `size`, `gap` apart, small random tilts) held
    just above a static ground box: the ipc mode's step workload (the cubes fall onto the
    ground and onto each other)."""
    rng = np.random.default_rng(SEED_BASE + 7000 + seed)
    half = 0.5 * (max(nx, nz) * (size + gap) + 1.0)
    bodies = [_box_body("ipc_ground", 2 * half, 0.5, 2 * half, [0.0, -0.25, 0.0], static=True)]
    for k in range(layers):
        for i in range(nx):
            for j in range(nz):
                p = [(i - 0.5 * (nx - 1)) * (size + gap), 0.5 * size + 0.02 + k * (size + gap),
                     (j - 0.5 * (nz - 1)) * (size + gap)]
                bodies.append(_box_body(f"ipc_cube{size}", size, size, size, p, random_tilt(rng, tilt_deg)))
    return pack([SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.0, name="ipc_drop")])


def large_mesh_scene(seed=0, subdiv=4, m=24, n_spheres=2, scale=1e-3, gap=0.67, dt=1e-3, dhat=1e-3,
                     mu=0.3) -> Scenes:
    """Meshes above 2048 primitives: a plate meshed as an m x m grid per face (m = 24:
    3,458 vertices, 10,368 edges, 6,912 faces) above a static ground box, with icospheres
    of subdivision `subdiv` (4: 2,562 vertices, 7,680 edges, 5,120 faces) and a small
    box above the plate, each `gap` above the body below it. Lengths are in units of
    `scale`. At 1e-3 with the default gap the plate's centre and the spheres' are
    0.97 dhat from the centre below them, where the barrier on the centres' offset along
    the contact normal acts (as in micro_contact_scene); at scale 1 with a small gap the
    meshes start within dhat of each other but far beyond it centre to centre."""
    rng = np.random.default_rng(SEED_BASE + 8000 + seed)
    L = float(scale)
    Vp, Fp = grid_box_mesh(1.5 * L, 0.1 * L, 1.5 * L, m)
    mp, Ip = polyhedron_mass_properties(Vp, Fp)
    yp = 0.05 + gap
    bodies = [_box_body(f"lm_ground{L!r}", 4.0 * L, 0.5 * L, 4.0 * L, [0.0, -0.25 * L, 0.0], static=True),
              BodySpec(f"grid_box{m}_{L!r}", Vp, Fp, mp, Ip, np.array([0.0, yp * L, 0.0]), _IDENT.copy(),
                       np.zeros(3), np.zeros(3))]
    for k in range(n_spheres):
        c = [(k - 0.5 * (n_spheres - 1) + rng.uniform(-0.02, 0.02)) * L, (yp + 0.3 + gap) * L,
             rng.uniform(-0.1, 0.1) * L]
        bodies.append(_ico_body(subdiv, 0.25 * L, c, q=random_tilt(rng, 20.0), v=rng.normal(0.0, 0.01, size=3)))
    bodies.append(_box_body(f"lm_box{L!r}", 0.1 * L, 0.1 * L, 0.1 * L, [0.6 * L, (yp + 0.1 + gap) * L, 0.6 * L],
                            random_tilt(rng, 3.0), v=rng.normal(0.0, 0.01, size=3)))
    return pack([SceneSpec(bodies, dt=dt, dhat=dhat, mu=mu, name="large_mesh")])


def micro_contact_scene(seed=0, n_cubes=4, size=6e-4, gap=2e-4, dhat=1e-3, mu=0.5, tilt_deg=0.0,
                        v_scale=0.01):
    """Millimetre cubes stacked so that centre separations are below dhat.

    The reference barrier acts on centre distance along the contact normal
    (DistanceBarrierForm.cpp:89-98), so only bodies whose centres are closer than
    dhat activate it. This scene makes the barrier, its normal forces and the lagged
    friction all non-zero, so the contact kernels are exercised end to end.
    """
    rng = np.random.default_rng(SEED_BASE + 9000 + seed)
    bodies = [_box_body("mground", 0.01, size, 0.01, [0.0, -0.5 * size, 0.0], static=True)]
    for k in range(n_cubes):
        q = random_tilt(rng, tilt_deg) if tilt_deg > 0 else _IDENT
        p = [(k % 2) * (size + gap) - 0.5 * (size + gap), 0.5 * size + gap * 0.5 + (k // 2) * (size + gap * 0.5),
             rng.uniform(-1e-5, 1e-5)]
        v = rng.normal(0.0, v_scale, size=3)
        bodies.append(_box_body(f"mcube{size}", size, size, size, p, q, v=v))
    return SceneSpec(bodies, dt=1e-3, dhat=dhat, mu=mu, name=f"micro{seed}")


# Barrier stiffness of friction_slide_scene (set as initialised barrier state): with it the
# cubes' weight is carried at a centre distance inside dhat, so they slide in contact.
FRICTION_SLIDE_KAPPA = 20.0


def friction_slide_scene(n_cubes=2, size=6e-4, gap=5e-5, mass=1e-3, vx=0.1, mu=0.4, spacing=3e-3,
                         inertia_scale=1e6) -> Scenes:
    """Millimetre cubes sliding along +x (and a little +z) on a static plate, their centres
    closer than dhat to the plate's so that the reference's centre-distance barrier carries
    them (the 50 um gap is about where it balances their weight), 3 mm apart (no cube-cube
    contact) and away from the plate's face diagonals (no edge-edge contact whose closest
    points sit on a crossing, where the Newton solve's last rounding error has a large
    derivative). They slide at 0.1 and 0.2 m/s, fast enough that the friction stays far
    from its smoothed stick region over the test's 10 steps. Masses are raised to `mass` (1 mg) so that the
    contact residual stays above the ALM break tolerance (IPCSolver.cpp:1132-1165): the
    ALM loop runs its iterations and the lagged friction, zero in the first, acts in the
    later ones, with normal forces and contact points taken at the previous iterate. The
    cubes' rotational inertia is scaled by `inertia_scale` so that they slide without
    tipping. With the barrier stiffness FRICTION_SLIDE_KAPPA and a tight Newton tolerance
    every solve converges or accepts no step, which makes the run the solution map that
    the adjoint differentiates (a differentiation test scene for the friction
    coefficient)."""
    bodies = [_box_body("mplate", 0.01, size, 0.01, [0.0, -0.5 * size, 0.0], static=True)]
    for k in range(n_cubes):
        bodies.append(_box_body(f"mcube{size}", size, size, size,
                                [(k - 0.5 * (n_cubes - 1)) * spacing, 0.5 * size + gap, 0.0],
                                v=np.array([vx * (1 + k), 0.0, 0.3 * vx])))
    sc = pack([SceneSpec(bodies, dt=1e-3, dhat=1e-3, mu=mu, name="slide")])
    m0 = float(sc.mass[1])
    sc.mass[1:] = mass
    sc.inertia.reshape(-1, 3, 3)[1:] *= mass / m0 * inertia_scale
    return sc


def hinge_chain_scene(n_links=3, joint_mu=1e4, seed=5) -> Scenes:
    """A free chain of 0.4 x 0.1 x 0.1 m links joined by hinges about z, with seeded
    initial velocities N(0, 0.3) and angular velocities N(0, 1), no static body. The
    hinges' ALM stiffness is joint_mu (the reference's default, 1e8, makes the Newton
    loop fail at every ALM iteration once links spin; 1e4 keeps every solve converged)."""
    rng = np.random.default_rng(SEED_BASE + 8000 + seed)
    bodies = [_box_body("link", 0.4, 0.1, 0.1, [0.45 * k, 1.0, 0.0], v=rng.normal(0.0, 0.3, 3),
                        w=rng.normal(0.0, 1.0, 3)) for k in range(n_links)]
    hinges = [(k, k + 1, np.array([0.225, 0, 0]), np.array([-0.225, 0, 0]), np.array([0, 0, 1.0]),
               np.array([0, 0, 1.0])) for k in range(n_links - 1)]
    sc = pack([SceneSpec(bodies, dt=1e-2, dhat=1e-3, mu=0.3, hinges=hinges, name="chain")])
    sc.joint_mu[:] = joint_mu
    return sc


def make_config(name: str, **kw) -> Scenes:
    if name == "c1":
        return pack([config1()])
    if name == "c2":
        return pack([config2(**kw)])
    if name == "c3":
        return pack(config3(**kw))
    if name == "c4":
        return pack([config4()[0]])
    if name == "c5":
        return pack(config5(**kw))
    raise ValueError(name)
