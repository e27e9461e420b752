"""Seeded scene generators for the AVBD step workloads.

A Scene lists bodies and forces in creation order, the way a reference scene
function builds them through ``new Rigid`` / ``new Joint`` / ``new Spring``
(scenes.h:59-136 show the pattern). ``build_state`` turns a scene into the
resident state tables of include/avbd_hip.h:

* bodies are stored in body-list order: the reference prepends each new body
  (rigid.cpp:29-31), so creation index c becomes list index N-1-c;
* each force's serial is its creation index, which fixes the per-body force
  order (force.cpp:25-36 prepends).

Workloads (gravity -9.81 along -z, the reference's up axis is +z):

* ``stack``   10 unit boxes resting on a ground box, horizontal jitter +-0.01 m.
* ``pyramid_chain``  a 20-wide pyramid of unit boxes (210) plus a 1000-link
  chain of 0.2 m boxes joined by ball joints, first link anchored to the world,
  with springs (rest 0.2 m, stiffness log-uniform in [1e3, 1e5]) on every other
  adjacent link pair.
* ``pile``    boxes on a jittered 3D grid (half-extents U[0.25, 1] m, uniform
  random orientations, velocities N(0, 0.5) m/s) dropping onto a ground box
  inside four static walls. 50,000 boxes is the headline workload; the
  generator takes any count.
* ``projectile_pile``  the differentiable workload: a 1024-box block and 64
  projectiles, with randomised masses and friction (``projectile_state``).
* ``batch``   independent worlds (4096 by default) with body counts
  log-uniform in [8, 512]: box-stack fields, piles and chains, for the
  batched step (one handle, every launch covers all worlds).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .state import BODY, JOINT, SPRING, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CYLINDER, SHAPE_SPHERE, State

FORCE_JOINT = 1
FORCE_SPRING = 2
FORCE_IGNORE = 3   # IgnoreCollision (force.cpp:75-79)


@dataclasses.dataclass
class Scene:
    name: str
    size: np.ndarray      # (n,3) full widths
    density: np.ndarray   # (n,)
    friction: np.ndarray  # (n,)
    pos: np.ndarray       # (n,3)
    rot: np.ndarray       # (n,4) x,y,z,w
    vel: np.ndarray       # (n,3)
    angvel: np.ndarray    # (n,3)
    ftype: np.ndarray     # (m,) 1 joint, 2 spring
    fa: np.ndarray        # (m,) creation index of body A, -1 = world
    fb: np.ndarray        # (m,)
    ra: np.ndarray        # (m,3)
    rb: np.ndarray        # (m,3)
    p0: np.ndarray        # joint: stiffnessLin; spring: stiffness
    p1: np.ndarray        # joint: stiffnessAng; spring: rest (-1 = current distance)
    p2: np.ndarray        # joint: fracture
    gravity: float = -9.81
    steps: int = 300
    shape: np.ndarray | None = None         # (n,) RigidShapeType; None: all boxes
    shape_radius: np.ndarray | None = None  # (n,) sphere/capsule/cylinder radius
    half_length: np.ndarray | None = None   # (n,) capsule/cylinder core half length

    def shape_columns(self):
        """(shape, radius, half length) per body, boxes filled in when the scene has none."""
        n = self.num_bodies
        if self.shape is None:
            return np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32)
        return (np.asarray(self.shape, np.int32), np.asarray(self.shape_radius, np.float32),
                np.asarray(self.half_length, np.float32))

    @property
    def num_bodies(self) -> int:
        return int(self.size.shape[0])

    @property
    def num_dynamic(self) -> int:
        return int((self.density > 0).sum())


class _Builder:
    def __init__(self):
        self.b = {k: [] for k in ("size", "density", "friction", "pos", "rot", "vel", "angvel", "shape", "srad",
                                  "hlen")}
        self.f = {k: [] for k in ("ftype", "fa", "fb", "ra", "rb", "p0", "p1", "p2")}

    def body(self, size, density, friction, pos, rot=(0, 0, 0, 1), vel=(0, 0, 0), angvel=(0, 0, 0),
             shape=SHAPE_BOX, radius=0.0, half_length=0.0) -> int:
        self.b["size"].append(size)
        self.b["density"].append(density)
        self.b["friction"].append(friction)
        self.b["pos"].append(pos)
        self.b["rot"].append(rot)
        self.b["vel"].append(vel)
        self.b["angvel"].append(angvel)
        self.b["shape"].append(shape)
        self.b["srad"].append(radius)
        self.b["hlen"].append(half_length)
        return len(self.b["size"]) - 1

    # Rigid::makeSphere / makeCapsule / makeCylinder (rigid.cpp:44-86): the recorded size is the
    # bounding box the constructors pass on to Rigid::Rigid.
    def sphere(self, radius, density, friction, pos, rot=(0, 0, 0, 1), vel=(0, 0, 0), angvel=(0, 0, 0)) -> int:
        d = np.float32(radius) * np.float32(2.0)
        return self.body((d, d, d), density, friction, pos, rot, vel, angvel, SHAPE_SPHERE, radius, 0.0)

    def capsule(self, radius, half_length, density, friction, pos, rot=(0, 0, 0, 1), vel=(0, 0, 0),
                angvel=(0, 0, 0)) -> int:
        d = np.float32(radius) * np.float32(2.0)
        h = np.float32(half_length) * np.float32(2.0) + d
        return self.body((d, d, h), density, friction, pos, rot, vel, angvel, SHAPE_CAPSULE, radius, half_length)

    def cylinder(self, radius, half_length, density, friction, pos, rot=(0, 0, 0, 1), vel=(0, 0, 0),
                 angvel=(0, 0, 0)) -> int:
        d = np.float32(radius) * np.float32(2.0)
        h = np.float32(half_length) * np.float32(2.0)
        return self.body((d, d, h), density, friction, pos, rot, vel, angvel, SHAPE_CYLINDER, radius, half_length)

    def joint(self, a, b, ra, rb, k_lin=math.inf, k_ang=0.0, fracture=math.inf):
        for k, v in zip(self.f, (FORCE_JOINT, a, b, ra, rb, k_lin, k_ang, fracture)):
            self.f[k].append(v)

    def spring(self, a, b, ra, rb, stiffness, rest=-1.0):
        for k, v in zip(self.f, (FORCE_SPRING, a, b, ra, rb, stiffness, rest, 0.0)):
            self.f[k].append(v)

    def ignore(self, a, b):
        """IgnoreCollision between bodies a and b: the broadphase never pairs them."""
        for k, v in zip(self.f, (FORCE_IGNORE, a, b, (0, 0, 0), (0, 0, 0), 0.0, 0.0, 0.0)):
            self.f[k].append(v)

    def scene(self, name, gravity, steps) -> Scene:
        f32 = np.float32
        n = len(self.b["size"])
        m = len(self.f["ftype"])
        return Scene(
            name=name,
            size=np.asarray(self.b["size"], f32).reshape(n, 3),
            density=np.asarray(self.b["density"], f32).reshape(n),
            friction=np.asarray(self.b["friction"], f32).reshape(n),
            pos=np.asarray(self.b["pos"], f32).reshape(n, 3),
            rot=np.asarray(self.b["rot"], f32).reshape(n, 4),
            vel=np.asarray(self.b["vel"], f32).reshape(n, 3),
            angvel=np.asarray(self.b["angvel"], f32).reshape(n, 3),
            ftype=np.asarray(self.f["ftype"], np.int32).reshape(m),
            fa=np.asarray(self.f["fa"], np.int32).reshape(m),
            fb=np.asarray(self.f["fb"], np.int32).reshape(m),
            ra=np.asarray(self.f["ra"], f32).reshape(m, 3),
            rb=np.asarray(self.f["rb"], f32).reshape(m, 3),
            p0=np.asarray(self.f["p0"], f32).reshape(m),
            p1=np.asarray(self.f["p1"], f32).reshape(m),
            p2=np.asarray(self.f["p2"], f32).reshape(m),
            gravity=gravity, steps=steps,
            shape=np.asarray(self.b["shape"], np.int32).reshape(n),
            shape_radius=np.asarray(self.b["srad"], f32).reshape(n),
            half_length=np.asarray(self.b["hlen"], f32).reshape(n))


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake's subgroup algorithm), x,y,z,w."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32)


def stack(seed: int = 1, count: int = 10, jitter: float = 0.01, steps: int = 300) -> Scene:
    """Stack of unit boxes resting on a static ground box (ground top at z = 0.5)."""
    rng = _rng(seed)
    sb = _Builder()
    sb.body((100, 100, 1), 0.0, 0.5, (0, 0, 0))
    for i in range(count):
        jx, jy = rng.uniform(-jitter, jitter, 2)
        sb.body((1, 1, 1), 1.0, 0.5, (jx, jy, 1.0 + i))
    return sb.scene(f"stack{count}", -9.81, steps)


def pyramid_chain(seed: int = 2, base: int = 20, links: int = 1000, steps: int = 600) -> Scene:
    """Pyramid of unit boxes plus a world-anchored chain with springs."""
    rng = _rng(seed)
    sb = _Builder()
    sb.body((300, 300, 1), 0.0, 0.5, (0, 0, -0.5))
    for row in range(base):
        for col in range(base - row):
            sb.body((1, 1, 1), 1.0, 0.5, (col * 1.01 + row * 0.505 - base / 2.0, 0.0, row * 1.0 + 0.5))
    link = 0.2
    x0, y0, z0 = -100.0, 10.0, 20.0
    ids = [sb.body((link, link, link), 1.0, 0.5, (x0 + link * 0.5 + k * link, y0, z0)) for k in range(links)]
    if links:
        sb.joint(-1, ids[0], (x0, y0, z0), (-link * 0.5, 0, 0))
    stiff = np.exp(rng.uniform(math.log(1e3), math.log(1e5), max(links - 1, 0)))
    for k in range(links - 1):
        sb.joint(ids[k], ids[k + 1], (link * 0.5, 0, 0), (-link * 0.5, 0, 0))
        if k % 2 == 0:
            sb.spring(ids[k], ids[k + 1], (0, 0, 0), (0, 0, 0), float(stiff[k]), link)
    return sb.scene(f"pyramid{base}_chain{links}", -9.81, steps)


def pile(count: int = 50000, seed: int = 3, grid=None, pitch: float = 3.6, steps: int = 1000,
         jitter: float = 0.05) -> Scene:
    """Random box pile on a jittered grid inside four static walls.

    The default 50,000-box grid is 50 x 40 x 25 (x, y, z); other counts use a
    grid with the same 5:4 footprint aspect that holds ``count`` boxes.
    """
    rng = _rng(seed)
    if grid is None:
        if count == 50000:
            grid = (50, 40, 25)
        else:
            nz = max(1, int(round((count / 20.0) ** (1.0 / 3.0))))
            per_layer = math.ceil(count / nz)
            ny = max(1, int(math.floor(math.sqrt(per_layer * 4.0 / 5.0))))
            nx = math.ceil(per_layer / ny)
            grid = (nx, ny, nz)
    nx, ny, nz = grid
    assert nx * ny * nz >= count
    half_x, half_y = nx * pitch / 2.0, ny * pitch / 2.0
    sb = _Builder()
    wall_h = 20.0
    sb.body((2 * half_x + 20, 2 * half_y + 20, 2), 0.0, 0.5, (0, 0, -1.0))
    sb.body((2, 2 * half_y + 4, wall_h), 0.0, 0.5, (-half_x - 1.0 - 1.0, 0, wall_h / 2))
    sb.body((2, 2 * half_y + 4, wall_h), 0.0, 0.5, (half_x + 1.0 + 1.0, 0, wall_h / 2))
    sb.body((2 * half_x + 4, 2, wall_h), 0.0, 0.5, (0, -half_y - 1.0 - 1.0, wall_h / 2))
    sb.body((2 * half_x + 4, 2, wall_h), 0.0, 0.5, (0, half_y + 1.0 + 1.0, wall_h / 2))
    half = rng.uniform(0.25, 1.0, (count, 3))
    quats = random_quaternions(rng, count)
    vel = rng.normal(0.0, 0.5, (count, 3))
    jit = rng.uniform(-jitter, jitter, (count, 3))
    k = np.arange(count)
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    cx = -half_x + pitch * (ix + 0.5) + jit[:, 0]
    cy = -half_y + pitch * (iy + 0.5) + jit[:, 1]
    cz = 2.0 + pitch * (iz + 0.5) + jit[:, 2]
    for i in range(count):
        sb.body(tuple(2.0 * half[i]), 1.0, 0.5, (cx[i], cy[i], cz[i]), tuple(quats[i]), tuple(vel[i]))
    return sb.scene(f"pile{count}", -9.81, steps)


def spread(count: int = 2000, seed: int = 9, extent: float = 400.0, steps: int = 120) -> Scene:
    """Small boxes (edges U[0.05, 0.2] m) scattered over an ``extent`` x ``extent`` m ground slab at
    heights U[0.3, 12] m. Sparse enough that the uniform-grid broadphase's 2^22-cell limit grows
    its cells (solver.cpp:336-344)."""
    rng = _rng(seed)
    sb = _Builder()
    sb.body((extent + 20.0, extent + 20.0, 2.0), 0.0, 0.5, (0.0, 0.0, -1.0))
    size = rng.uniform(0.05, 0.2, (count, 3))
    quats = random_quaternions(rng, count)
    xy = rng.uniform(-extent / 2.0, extent / 2.0, (count, 2))
    z = rng.uniform(0.3, 12.0, count)
    vel = rng.normal(0.0, 0.5, (count, 3))
    for i in range(count):
        sb.body(tuple(size[i]), 1.0, 0.5, (xy[i, 0], xy[i, 1], z[i]), tuple(quats[i]), tuple(vel[i]))
    return sb.scene(f"spread{count}", -9.81, steps)


def stacks(count: int, seed: int = 1, height: int = 10, jitter: float = 0.01, spacing: float = 1.5,
           steps: int = 300) -> Scene:
    """Columns of up to ``height`` unit boxes on a static ground box, horizontal jitter +-0.01 m."""
    rng = _rng(seed)
    sb = _Builder()
    cols = max(1, math.ceil(count / height))
    side = math.ceil(math.sqrt(cols))
    sb.body((side * spacing + 10, side * spacing + 10, 1), 0.0, 0.5, (0, 0, 0))
    for i in range(count):
        c, h = divmod(i, height)
        cx = (c % side - (side - 1) / 2.0) * spacing
        cy = (c // side - (side - 1) / 2.0) * spacing
        jx, jy = rng.uniform(-jitter, jitter, 2)
        sb.body((1, 1, 1), 1.0, 0.5, (cx + jx, cy + jy, 1.0 + h))
    return sb.scene(f"stacks{count}", -9.81, steps)


BATCH_KINDS = ("stacks", "pile", "chain")


# ---- round shapes (Rigid::makeSphere / makeCapsule / makeCylinder) ----

def _axis_quat(axis, angle):
    """Unit quaternion (x, y, z, w) of a rotation by ``angle`` about a unit axis."""
    s, c = math.sin(angle * 0.5), math.cos(angle * 0.5)
    return (axis[0] * s, axis[1] * s, axis[2] * s, c)


def _rotate(q, v):
    x, y, z, w = q
    u = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def round_stack(kind: str = "sphere", count: int = 10, steps: int = 300) -> Scene:
    """A column of spheres, crosswise-lying capsules or standing cylinders on a static ground slab
    (the slab is a thin static box, so the shapes meet its top through the thin-top paths)."""
    sb = _Builder()
    sb.body((1000, 1000, 1), 0.0, 0.5, (0, 0, 0))
    for i in range(count):
        if kind == "sphere":
            r = 0.5
            sb.sphere(r, 1.0, 0.5, (0, 0, 0.5 + r + i * r * 2.05))
        elif kind == "capsule":
            r, hl = 0.35, 0.9
            q = _axis_quat((0, 1, 0) if i % 2 == 0 else (1, 0, 0), math.pi / 2)
            sb.capsule(r, hl, 1.0, 0.5, (0, 0, 0.5 + r + i * r * 2.4), q)
        elif kind == "cylinder":
            r, hl = 0.48, 0.35
            sb.cylinder(r, hl, 1.0, 0.6, (0, 0, 0.5 + hl + i * (hl * 2 + 0.025)))
        else:
            raise KeyError(kind)
    return sb.scene(f"{kind}_stack{count}", -9.81, steps)


def round_ramp(kind: str = "sphere", count: int = 6, angle_deg: float = 20.0, steps: int = 240) -> Scene:
    """Round shapes released on a tilted static slab with frictions from 0.1 up (spheres roll and
    slide, capsules lie along the slope, cylinders lie across it)."""
    sb = _Builder()
    sb.body((1000, 1000, 1), 0.0, 0.5, (0, 0, 0))
    q = _axis_quat((0, 1, 0), math.radians(angle_deg))
    sb.body((36, 16, 1), 0.0, 0.8, (0, 0, 3), q)
    tangent, normal = _rotate(q, (1, 0, 0)), _rotate(q, (0, 0, 1))
    for i in range(count):
        fr = 0.1 + 0.15 * i
        if kind == "sphere":
            r, rot = 0.55, (0, 0, 0, 1)
        elif kind == "capsule":
            r, rot = 0.35, _axis_quat((0, 1, 0), math.pi / 2)
        else:
            r, rot = 0.48, _axis_quat((1, 0, 0), -math.pi / 2)
        pos = np.array([0, 0, 3.0]) + tangent * (-12.0 + i * 2.0) + np.array([0, -3.0 + i, 0]) + normal * (0.5 + r + 0.02)
        if kind == "sphere":
            sb.sphere(r, 1.0, fr, tuple(pos), rot)
        elif kind == "capsule":
            sb.capsule(r, 0.9, 1.0, fr, tuple(pos), rot)
        else:
            sb.cylinder(r, 0.65, 1.0, fr, tuple(pos), rot)
    return sb.scene(f"{kind}_ramp{count}", -9.81, steps)


def sphere_pour(count: int = 1000, width: int = 8, depth: int = 8, start_z: float = 10.0, seed: int = 7,
                steps: int = 300, layer_scale: float = 1.0, forward: float = 1.1, down: float = -1.5,
                taper: bool = True) -> Scene:
    """Small spheres (r = 0.12) poured in a jittered column of width x depth per layer, moving
    forward and down, onto seven standing cylinders on a static ground slab (the layout of the
    reference's sphere-pour scenes, scenes.h:270-342, with this module's own random jitter)."""
    rng = _rng(seed)
    sb = _Builder()
    sb.body((1000, 1000, 1), 0.0, 0.7, (0, 0, 0))
    cr, chl, sp = 0.45, 0.55, 1.7
    for x, y in ((0, 0), (-sp, 0), (sp, 0), (-sp / 2, sp * 0.9), (sp / 2, sp * 0.9), (-sp / 2, -sp * 0.9),
                 (sp / 2, -sp * 0.9)):
        sb.cylinder(cr, chl, 1.0, 0.85, (x, y, 0.5 + chl))
    r = 0.12
    spacing = r * 2.35
    jit = rng.uniform(-1.0, 1.0, (count, 3)) * np.array([0.35, 0.35, 0.2]) * r
    vj = rng.uniform(-1.0, 1.0, (count, 3)) * np.array([0.18, 0.35, 0.18])
    per_layer = width * depth
    for k in range(count):
        layer, cell = divmod(k, per_layer)
        y, x = divmod(cell, width)
        tf = 1.0 - min(layer / 18.0, 0.45) if taper else 1.0
        px = (x - (width - 1) * 0.5) * spacing * tf + jit[k, 0] - 1.2
        py = (y - (depth - 1) * 0.5) * spacing * tf + jit[k, 1]
        pz = start_z + layer * spacing * layer_scale + jit[k, 2]
        sb.sphere(r, 1.0, 0.45, (px, py, pz), vel=(forward + vj[k, 0], vj[k, 1], down + vj[k, 2]))
    return sb.scene(f"sphere_pour{count}", -9.81, steps)


# The gates of the reference's benchmark pours (scenes.h:332-342 built by avbd/named_scenes.py;
# viewer/package.json:21-22): 600 frames; average physics step under 25 ms (5000 spheres) or
# 30 ms (20000 spheres); max sphere-ground penetration 0.05 / 0.08; final max speed <= 25;
# final dynamic max height <= 35.
POUR_GATES = {
    5000: dict(avg_ms=25.0, penetration=0.05, max_speed=25.0, max_height=35.0),
    20000: dict(avg_ms=30.0, penetration=0.08, max_speed=25.0, max_height=35.0),
}


def mixed_pile(count: int = 2000, seed: int = 8, pitch: float = 2.8, steps: int = 300) -> Scene:
    """Boxes, spheres, capsules and cylinders in random sizes and orientations on a jittered grid
    inside four static walls: every pair kind of the narrowphase meets every other."""
    rng = _rng(seed)
    side = max(1, int(math.ceil((count / 4.0) ** 0.5)))
    half = side * pitch / 2.0
    sb = _Builder()
    sb.body((2 * half + 20, 2 * half + 20, 2), 0.0, 0.5, (0, 0, -1.0))
    for sx, sy, px, py in ((2, 2 * half + 4, -half - 2, 0), (2, 2 * half + 4, half + 2, 0),
                           (2 * half + 4, 2, 0, -half - 2), (2 * half + 4, 2, 0, half + 2)):
        sb.body((sx, sy, 20), 0.0, 0.5, (px, py, 10.0))
    kind = rng.integers(0, 4, count)
    dims = rng.uniform(0.25, 0.9, (count, 3))
    quats = random_quaternions(rng, count)
    vel = rng.normal(0.0, 0.5, (count, 3))
    fric = rng.uniform(0.3, 0.8, count)
    jit = rng.uniform(-0.05, 0.05, (count, 3))
    for k in range(count):
        ix, iy, iz = k % side, (k // side) % side, k // (side * side)
        pos = (-half + pitch * (ix + 0.5) + jit[k, 0], -half + pitch * (iy + 0.5) + jit[k, 1],
               1.5 + pitch * (iz + 0.5) + jit[k, 2])
        rot, v = tuple(quats[k]), tuple(vel[k])
        if kind[k] == 0:
            sb.body(tuple(2.0 * dims[k]), 1.0, fric[k], pos, rot, v)
        elif kind[k] == 1:
            sb.sphere(dims[k, 0], 1.0, fric[k], pos, rot, v)
        elif kind[k] == 2:
            sb.capsule(0.6 * dims[k, 0], 0.5 * dims[k, 1], 1.0, fric[k], pos, rot, v)
        else:
            sb.cylinder(dims[k, 0], 0.5 * dims[k, 1], 1.0, fric[k], pos, rot, v)
    return sb.scene(f"mixed_pile{count}", -9.81, steps)


def round_cases(steps: int = 90) -> list:
    """Small scenes that each drive one branch of the round-shape narrowphase (collide.cpp:435-1181)
    and of the contact offsets (manifold.cpp:29-90)."""
    out = []

    def case(name, build):
        sb = _Builder()
        build(sb)
        out.append(sb.scene(name, -9.81, steps))

    def ground(sb):
        sb.body((100, 100, 1), 0.0, 0.6, (0, 0, 0))

    # sphere resting on a dynamic box: outside, within the radius
    case("sphere_on_box", lambda sb: (ground(sb), sb.body((2, 2, 1), 1.0, 0.5, (0, 0, 1.0)),
                                      sb.sphere(0.4, 1.0, 0.5, (0.3, 0.1, 1.9))))
    # sphere inside a static cube (not a thin slab): pushed out of the top face
    case("sphere_in_static_cube", lambda sb: (sb.body((2, 2, 2), 0.0, 0.5, (0, 0, 0)),
                                              sb.sphere(0.3, 1.0, 0.5, (0.2, -0.1, 0.7))))
    # sphere inside a moving dynamic box: the face most against the relative motion
    case("sphere_in_moving_box", lambda sb: (ground(sb), sb.body((2, 2, 2), 1.0, 0.5, (0, 0, 1.6), vel=(0.5, 0, 0)),
                                             sb.sphere(0.3, 1.0, 0.5, (0.2, 0.1, 1.8), vel=(-1.0, 0.2, 0.0))))
    # sphere inside a resting dynamic box with no relative motion: the nearest face
    case("sphere_in_resting_box", lambda sb: (ground(sb), sb.body((2, 2, 2), 1.0, 0.5, (0, 0, 1.6)),
                                              sb.sphere(0.3, 1.0, 0.5, (0.6, 0.1, 1.7))))
    # fast sphere crossing a dynamic plate within one step: the swept test
    case("sphere_swept_plate", lambda sb: (ground(sb), sb.body((3, 3, 0.2), 1.0, 0.5, (0, 0, 3.0)),
                                           sb.sphere(0.1, 1.0, 0.5, (0, 0, 3.35), vel=(0, 0, -30.0))))
    # fast sphere that ends a step inside a static cube: the swept test from inside
    case("sphere_fast_into_cube", lambda sb: (sb.body((2, 2, 2), 0.0, 0.5, (0, 0, 0)),
                                              sb.sphere(0.15, 1.0, 0.5, (0.1, 0.2, 1.5), vel=(0, 0, -40.0))))
    # sphere that crosses a whole plate within one step: the swept test from outside
    case("sphere_through_plate", lambda sb: (ground(sb), sb.body((3, 3, 0.2), 1.0, 0.5, (0, 0, 3.0)),
                                             sb.sphere(0.1, 1.0, 0.5, (0, 0, 3.6), vel=(0, 0, -60.0))))
    # capsule lying across a static slab narrower than its core: both end points miss the top
    # face, the point under the centre does not
    case("capsule_across_slab", lambda sb: (sb.body((1, 1, 0.1), 0.0, 0.5, (0, 0, 2.0)),
                                            sb.capsule(0.2, 1.0, 1.0, 0.5, (0, 0, 2.24), _axis_quat((0, 1, 0), math.pi / 2))))
    # capsule on a dynamic box, tilted so one end is nearer
    case("capsule_on_box", lambda sb: (ground(sb), sb.body((3, 2, 1), 1.0, 0.5, (0, 0, 1.0)),
                                       sb.capsule(0.3, 0.6, 1.0, 0.5, (0, 0, 2.1), _axis_quat((0, 1, 0), 1.2))))
    # capsule lying flat on the ground slab (both ends and the near-parallel branch)
    case("capsule_flat", lambda sb: (ground(sb), sb.capsule(0.3, 0.8, 1.0, 0.5, (0, 0, 0.85),
                                                            _axis_quat((1, 0, 0), math.pi / 2))))
    # capsule inside a static cube
    case("capsule_in_static_cube", lambda sb: (sb.body((2, 2, 2), 0.0, 0.5, (0, 0, 0)),
                                               sb.capsule(0.2, 0.3, 1.0, 0.5, (0.1, 0, 0.6))))
    # cylinder on a dynamic box: collided as the box of its bounding size
    case("cylinder_on_box", lambda sb: (ground(sb), sb.body((2, 2, 1), 1.0, 0.5, (0, 0, 1.0)),
                                        sb.cylinder(0.4, 0.3, 1.0, 0.5, (0.2, 0, 1.9), _axis_quat((0, 1, 0), 0.3))))
    # cylinder standing (cap rim) and lying (side lines) on the ground slab
    case("cylinder_standing", lambda sb: (ground(sb), sb.cylinder(0.5, 0.4, 1.0, 0.6, (0, 0, 0.95),
                                                                  _axis_quat((1, 0, 0), 0.2))))
    case("cylinder_lying", lambda sb: (ground(sb), sb.cylinder(0.4, 0.6, 1.0, 0.6, (0, 0, 0.95),
                                                               _axis_quat((1, 0, 0), math.pi / 2), angvel=(0, 0, 1.0))))
    # two coaxial cylinders: caps facing
    case("cylinders_caps", lambda sb: (ground(sb), sb.cylinder(0.5, 0.4, 1.0, 0.6, (0, 0, 0.95)),
                                       sb.cylinder(0.4, 0.3, 1.0, 0.6, (0.05, 0, 1.7))))
    # two parallel cylinders side by side
    case("cylinders_side", lambda sb: (ground(sb),
                                       sb.cylinder(0.4, 0.6, 1.0, 0.6, (0, 0, 0.95), _axis_quat((1, 0, 0), math.pi / 2)),
                                       sb.cylinder(0.4, 0.5, 1.0, 0.6, (0.2, 0.1, 1.8), _axis_quat((1, 0, 0), math.pi / 2))))
    # crossed cylinders: closest points of the two axes
    case("cylinders_crossed", lambda sb: (ground(sb),
                                          sb.cylinder(0.3, 0.8, 1.0, 0.6, (0, 0, 0.85), _axis_quat((1, 0, 0), math.pi / 2)),
                                          sb.cylinder(0.3, 0.8, 1.0, 0.6, (0, 0.1, 1.5), _axis_quat((0, 1, 0), math.pi / 2))))
    # cylinder against a sphere and against a capsule
    case("cylinder_sphere", lambda sb: (ground(sb), sb.cylinder(0.5, 0.5, 1.0, 0.8, (0, 0, 1.05)),
                                        sb.sphere(0.3, 1.0, 0.5, (0.1, 0.05, 1.95))))
    case("capsule_cylinder", lambda sb: (ground(sb),
                                         sb.capsule(0.3, 0.5, 1.0, 0.8, (0, 0, 1.9), _axis_quat((1, 0, 0), 1.3)),
                                         sb.cylinder(0.5, 0.5, 1.0, 0.8, (0, 0, 1.05))))
    # sphere-sphere, sphere-capsule, capsule-capsule
    case("sphere_sphere", lambda sb: (ground(sb), sb.sphere(0.5, 1.0, 0.5, (0, 0, 1.0)),
                                      sb.sphere(0.4, 1.0, 0.5, (0.3, 0.1, 1.85))))
    case("sphere_capsule", lambda sb: (ground(sb), sb.capsule(0.3, 0.7, 1.0, 0.5, (0, 0, 0.8),
                                                              _axis_quat((0, 1, 0), math.pi / 2)),
                                       sb.sphere(0.3, 1.0, 0.5, (0.4, 0.05, 1.4))))
    case("capsules_crossed", lambda sb: (ground(sb),
                                         sb.capsule(0.3, 0.7, 1.0, 0.5, (0, 0, 0.8), _axis_quat((0, 1, 0), math.pi / 2)),
                                         sb.capsule(0.3, 0.7, 1.0, 0.5, (0.05, 0, 1.45), _axis_quat((1, 0, 0), math.pi / 2))))
    # static sphere under a dynamic sphere: no rolling pass for the dynamic one (the static A is chosen)
    case("spheres_static_a", lambda sb: (sb.sphere(0.6, 0.0, 0.5, (0, 0, 0)),
                                         sb.sphere(0.3, 1.0, 0.5, (0.05, 0, 0.95))))
    return out


def batch(num_worlds: int = 4096, seed: int = 4, lo: int = 8, hi: int = 512, steps: int = 300,
          kinds=BATCH_KINDS) -> list:
    """Independent worlds for the batched step: body counts log-uniform in [lo, hi]
    (statics included), each world a box-stack field, a jittered pile (ground + 4
    walls) or a world-anchored chain with springs over a ground box, drawn with the
    same distributions as the single-world workloads."""
    rng = _rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), num_worlds))).astype(np.int64)
    which = rng.integers(0, len(kinds), num_worlds)
    seeds = rng.integers(1, 2 ** 31 - 1, num_worlds)
    worlds = []
    for n, k, sd in zip(sizes, which, seeds):
        n, kind, sd = int(n), kinds[int(k)], int(sd)
        if kind == "pile" and n >= 6:
            sc = pile(n - 5, seed=sd, steps=steps)
        elif kind == "chain" and n >= 3:
            sc = pyramid_chain(seed=sd, base=0, links=n - 1, steps=steps)
        else:
            sc = stacks(n - 1, seed=sd, steps=steps)
        worlds.append(sc)
    return worlds


def spring_chain(links: int = 6, seed: int = 12, ground: bool = False, steps: int = 60) -> Scene:
    """Unit boxes in a row 1.2 m apart, each tied to the next by a spring (rest 1.0 m, stiffness
    log-uniform in [1e2, 1e3], anchors on the facing faces), with seeded velocities and spins;
    free flight, or dropped onto a ground slab. Without the ground the motion is smooth: no
    contact or clamp decision changes under small perturbations."""
    rng = _rng(seed)
    sb = _Builder()
    if ground:
        sb.body((40.0, 40.0, 1.0), 0.0, 0.5, (0.0, 0.0, -0.5))
    first = len(sb.b["size"])
    for i in range(links):
        sb.body((1.0, 1.0, 1.0), 1.0, 0.5, (1.2 * i, 0.0, 2.0 + 0.1 * i), (0, 0, 0, 1),
                tuple(rng.normal(0.0, 0.5, 3)), tuple(rng.normal(0.0, 0.3, 3)))
    for i in range(links - 1):
        sb.spring(first + i, first + i + 1, (0.5, 0.0, 0.0), (-0.5, 0.0, 0.0),
                  float(np.exp(rng.uniform(np.log(1e2), np.log(1e3)))), 1.0)
    return sb.scene(f"spring_chain{links}", -9.81, steps)


def joint_chain(links: int = 6, seed: int = 14, k_ang: float = 0.0, pile_count: int = 0, steps: int = 60,
                k_lin: float = math.inf) -> Scene:
    """Boxes of 0.4 m in a row, each joined to the next at their facing faces (linear stiffness
    ``k_lin``, infinite by default; angular stiffness ``k_ang``, 0 for ball joints), with seeded
    velocities and spins;
    free flight, or (pile_count > 0) dropped onto a jittered pile of that many boxes."""
    rng = _rng(seed)
    sb = _Builder()
    top = 3.0
    if pile_count > 0:
        base = pile(pile_count, seed=seed, pitch=2.2)
        for i in range(base.num_bodies):
            sb.body(tuple(base.size[i]), float(base.density[i]), float(base.friction[i]), tuple(base.pos[i]),
                    tuple(base.rot[i]), tuple(base.vel[i]), tuple(base.angvel[i]))
        top = float(base.pos[:, 2].max()) + 3.0
    first = len(sb.b["size"])
    for i in range(links):
        sb.body((0.4, 0.4, 0.4), 1.0, 0.5, (-0.5 * (links - 1) * 0.5 + 0.5 * i, 0.0, top + 0.05 * i), (0, 0, 0, 1),
                tuple(rng.normal(0.0, 0.4, 3)), tuple(rng.normal(0.0, 0.4, 3)))
    for i in range(links - 1):
        sb.joint(first + i, first + i + 1, (0.25, 0.0, 0.0), (-0.25, 0.0, 0.0), k_lin=k_lin, k_ang=k_ang)
    return sb.scene(f"joint_chain{links}" + (f"+pile{pile_count}" if pile_count else ""), -9.81, steps)


def spring_drop(count: int = 60, links: int = 8, seed: int = 13, steps: int = 90) -> Scene:
    """A jittered pile of ``count`` boxes with a chain of ``links`` boxes tied by springs (rest
    0.3 m, stiffness log-uniform in [1e3, 1e4]) dropped onto it: contacts and springs together."""
    rng = _rng(seed)
    base = pile(count, seed=seed, pitch=2.2)
    sb = _Builder()
    for i in range(base.num_bodies):
        sb.body(tuple(base.size[i]), float(base.density[i]), float(base.friction[i]), tuple(base.pos[i]),
                tuple(base.rot[i]), tuple(base.vel[i]), tuple(base.angvel[i]))
    top = float(base.pos[:, 2].max()) + 3.0
    first = len(sb.b["size"])
    for i in range(links):
        sb.body((0.4, 0.4, 0.4), 1.0, 0.5, (-1.4 + 0.4 * i, 0.0, top), (0, 0, 0, 1), tuple(rng.normal(0.0, 0.3, 3)))
    for i in range(links - 1):
        sb.spring(first + i, first + i + 1, (0.2, 0.0, 0.0), (-0.2, 0.0, 0.0),
                  float(np.exp(rng.uniform(np.log(1e3), np.log(1e4)))), 0.3)
    return sb.scene(f"spring_drop{count}+{links}", -9.81, steps)


def projectile_pile(count: int = 1024, projectiles: int = 64, seed: int = 5, steps: int = 120) -> Scene:
    """The differentiable workload: a jittered block of unit boxes resting on a ground box
    (16 x 16 x 4 for 1024) and 0.5 m projectile boxes launched from a ring of radius 20 m so
    that they reach the block after about one second. Creation order: ground, the block, then
    the projectiles (masses and friction are randomised on the state, see ``projectile_state``)."""
    rng = _rng(seed)
    sb = _Builder()
    side = max(1, int(math.ceil(math.sqrt(count / 4.0))))
    layers = max(1, int(math.ceil(count / float(side * side))))
    pitch = 1.02
    sb.body((side * pitch + 60, side * pitch + 60, 1), 0.0, 0.5, (0, 0, 0))
    jit = rng.uniform(-0.01, 0.01, (count, 2))
    for k in range(count):
        ix, iy, iz = k % side, (k // side) % side, k // (side * side)
        x = (ix - (side - 1) / 2.0) * pitch + jit[k, 0]
        y = (iy - (side - 1) / 2.0) * pitch + jit[k, 1]
        sb.body((1, 1, 1), 1.0, 0.5, (x, y, 1.0 + iz))
    ang = rng.uniform(0.0, 2.0 * math.pi, projectiles)
    height = rng.uniform(1.0, layers + 1.0, projectiles)
    t_hit = 1.0
    for k in range(projectiles):
        r = 20.0
        px, py = r * math.cos(ang[k]), r * math.sin(ang[k])
        pz = 3.0 + 0.5 * layers
        tz = height[k]
        vx, vy = -px / t_hit, -py / t_hit
        vz = (tz - pz + 0.5 * 9.81 * t_hit * t_hit) / t_hit
        w = rng.normal(0.0, 1.0, 3)
        sb.body((0.5, 0.5, 0.5), 1.0, 0.5, (px, py, pz), tuple(random_quaternions(rng, 1)[0]), (vx, vy, vz),
                tuple(w))
    return sb.scene(f"pile{count}_proj{projectiles}", -9.81, steps)


def projectile_state(scene: Scene, maker, seed: int = 5):
    """State of ``projectile_pile`` with per-body masses U[0.5, 2] (moments scaled with the mass)
    and friction coefficients U[0.3, 0.8] for the dynamic bodies. Returns (state, pile, proj,
    target): body-list indices of the block and of the projectiles, and a target point for the
    block's centre of mass (its start COM plus N(0, 0.5) m per axis)."""
    st = build_state(scene, maker)
    rng = _rng(seed + 1000)
    n = len(st.bodies)
    created = n - 1 - np.arange(n)                 # list index -> creation index
    dyn = st.bodies["mass"] > 0
    m_new = rng.uniform(0.5, 2.0, n).astype(np.float32)
    mu = rng.uniform(0.3, 0.8, n).astype(np.float32)
    scale = np.where(dyn, m_new / np.where(dyn, st.bodies["mass"], 1.0), 1.0).astype(np.float32)
    st.bodies["moment"] = (st.bodies["moment"] * scale[:, None]).astype(np.float32)
    st.bodies["mass"] = np.where(dyn, m_new, st.bodies["mass"]).astype(np.float32)
    st.bodies["friction"] = np.where(dyn, mu, st.bodies["friction"]).astype(np.float32)
    n_block = int(scene.name.split("_")[0][4:])
    pile = np.where((created >= 1) & (created <= n_block))[0]
    proj = np.where(created > n_block)[0]
    w = st.bodies["mass"][pile].astype(np.float64)
    com = (st.bodies["pos"][pile].astype(np.float64) * w[:, None]).sum(0) / w.sum()
    target = com + rng.normal(0.0, 0.5, 3)
    return st, pile, proj, target


def by_name(name: str) -> Scene:
    if name == "stack":
        return stack()
    if name == "pyramid_chain":
        return pyramid_chain()
    if name == "pile50k":
        return pile(50000)
    raise KeyError(name)


def build_state(scene: Scene, maker) -> State:
    """Scene (creation order) -> State (list order, serials = creation index).

    ``maker`` supplies ``make_box_body(size, density, friction, pos, rot, vel, angvel)`` and
    ``make_round_body(shape, radius, half_length, density, friction, pos, rot, vel, angvel)``
    returning a BODY record, and ``joint_spring_derived(rec_a_or_None, rec_b, ra, rb)``
    returning (torque_arm, rest) -- the constructor-derived fields.
    """
    n = scene.num_bodies
    shape, srad, hlen = scene.shape_columns()
    created = np.zeros(n, BODY)
    for c in range(n):
        if shape[c] == SHAPE_BOX:
            created[c] = maker.make_box_body(scene.size[c], scene.density[c], scene.friction[c], scene.pos[c],
                                             scene.rot[c], scene.vel[c], scene.angvel[c])
        else:
            created[c] = maker.make_round_body(int(shape[c]), srad[c], hlen[c], scene.density[c], scene.friction[c],
                                               scene.pos[c], scene.rot[c], scene.vel[c], scene.angvel[c])
        created[c]["dense_id"] = c
    bodies = created[::-1].copy()
    to_list = lambda c: -1 if c < 0 else n - 1 - int(c)
    joints, springs, ignores = [], [], []
    for k in range(scene.ftype.shape[0]):
        a, b = int(scene.fa[k]), int(scene.fb[k])
        if scene.ftype[k] == FORCE_IGNORE:
            ignores.append((to_list(a), to_list(b)))
            continue
        rec_a = created[a] if a >= 0 else None
        arm, rest = maker.joint_spring_derived(rec_a, created[b], scene.ra[k], scene.rb[k])
        if scene.ftype[k] == FORCE_JOINT:
            j = np.zeros((), JOINT)
            j["body_a"], j["body_b"], j["serial"] = to_list(a), to_list(b), k
            j["r_a"], j["r_b"] = scene.ra[k], scene.rb[k]
            j["stiffness_lin"], j["stiffness_ang"], j["fracture"] = scene.p0[k], scene.p1[k], scene.p2[k]
            j["torque_arm"] = arm
            joints.append(j)
        else:
            s = np.zeros((), SPRING)
            s["body_a"], s["body_b"], s["serial"] = to_list(a), to_list(b), k
            s["r_a"], s["r_b"] = scene.ra[k], scene.rb[k]
            s["stiffness"] = scene.p0[k]
            s["rest"] = scene.p1[k] if scene.p1[k] >= 0 else rest
            springs.append(s)
    return State(bodies, np.array(joints, JOINT) if joints else None,
                 np.array(springs, SPRING) if springs else None, None, scene.ftype.shape[0],
                 np.array(ignores, np.int32) if ignores else None)
