"""The reference's named scenes (scenes.h:18-1245), as Scene descriptions for this build.

Each builder lays the scene out the way the reference's scene function does — same bodies in
the same creation order, same forces in the same order — with every position computed in FP32
in the reference's operation order (numpy float32 scalars round like C floats), so a scene built
here starts from the reference's exact state. sinf/cosf are taken as the correctly rounded FP32
value of the FP64 sine/cosine. Scenes are looked up by name the way the headless server does
(SimulationHost::loadSceneByName: lower-case letters and digits only).

The mesh torus scene uses the particle-lattice importer in mesh_lattice.py.
"""
from __future__ import annotations

import math

import numpy as np

from .scenes import Scene, _Builder

F = np.float32
INF = math.inf


def f3(x, y, z):
    return (F(x), F(y), F(z))


def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def mul(a, s):
    s = F(s)
    return (a[0] * s, a[1] * s, a[2] * s)


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def length(a):
    return F(np.sqrt(dot(a, a)))


def normalize(a):
    n = length(a)
    return (a[0] / n, a[1] / n, a[2] / n)


def rotate(q, v):
    """maths.h:441-446: v + 2(u x v) w + u x (2 u x v), u = q.xyz."""
    u = (q[0], q[1], q[2])
    t = mul(cross(u, v), 2.0)
    return add(add(v, mul(t, q[3])), cross(u, t))


def sinf(x):
    return F(math.sin(float(x)))


def cosf(x):
    return F(math.cos(float(x)))


def rad(deg):
    return F(deg) * F(0.01745329251994329577)


def half_angle_quat(axis: int, angle):
    """Rotation by angle about a coordinate axis, as {.., sinf(a/2), .., cosf(a/2)}."""
    h = F(angle) * F(0.5)
    q = [F(0), F(0), F(0), cosf(h)]
    q[axis] = sinf(h)
    return tuple(q)


class SceneBuilder(_Builder):
    """_Builder that takes FP32 tuples and remembers the created index of each body."""

    def box(self, size, density, friction, pos, vel=(0, 0, 0), rot=(0, 0, 0, 1)):
        return self.body(tuple(F(s) for s in size), F(density), F(friction), pos, rot, vel)


# ---- scenes.h:18-155 ----

def _ground(sb, size=(100, 100, 1), friction=0.5, pos=(0, 0, 0)):
    return sb.box(size, 0.0, friction, f3(*pos))


def scene_empty(sb):
    pass


def scene_ground(sb):
    _ground(sb)
    sb.box((1, 1, 1), 1.0, 0.5, f3(0, 0, 4))


def scene_dynamic_friction(sb):
    _ground(sb)
    for x in range(11):
        fr = F(5.0) - (F(x) / F(10.0) * F(5.0))
        sb.box((1, 1, 0.5), 1.0, fr, (F(0), F(-30.0) + F(x) * F(2.0), F(0.75)), vel=f3(10, 0, 0))


def _ramp(sb, size, friction, angle_deg):
    q = (F(0), sinf(rad(angle_deg) * F(0.5)), F(0), cosf(rad(angle_deg) * F(0.5)))
    pos = f3(0, 0, 3)
    sb.box(size, 0.0, friction, pos, rot=q)
    return pos, normalize(rotate(q, f3(1, 0, 0))), normalize(rotate(q, f3(0, 0, 1)))


def scene_static_friction(sb):
    _ground(sb)
    rp, tangent, normal = _ramp(sb, (40, 24, 1), 1.0, 30.0)
    for i in range(11):
        fr = F(i) / F(10.0) * F(0.25) + F(0.25)
        y = F(-10.0) + F(i) * F(2.0)
        pos = add(add(add(rp, mul(tangent, -12.0)), (F(0), y, F(0))), mul(normal, 1.05))
        sb.box((1, 1, 1), 1.0, fr, pos)


def scene_pyramid(sb):
    size = 16
    _ground(sb, pos=(0, 0, -0.5))
    for y in range(size):
        for x in range(size - y):
            px = F(x) * F(1.01) + F(y) * F(0.5) - F(size / 2.0)
            sb.box((1, 0.5, 0.5), 1.0, 0.5, (px, F(0), F(y) * F(0.85) + F(0.5)))


def scene_rope(sb):
    _ground(sb, pos=(0, 0, -20))
    prev = None
    for i in range(20):
        curr = sb.box((1, 0.5, 0.5), 0.0 if i == 0 else 1.0, 0.5, f3(i, 0, 10))
        if prev is not None:
            sb.joint(prev, curr, f3(0.5, 0, 0), f3(-0.5, 0, 0))
        prev = curr


def scene_heavy_rope(sb):
    n, size = 20, F(5)
    _ground(sb, pos=(0, 0, -20))
    prev = None
    for i in range(n):
        last = i == n - 1
        dims = (size, size, size) if last else (1, 0.5, 0.5)
        curr = sb.box(dims, 0.0 if i == 0 else 1.0, 0.5, (F(i) + (size / F(2) if last else F(0)), F(0), F(10)))
        if prev is not None:
            sb.joint(prev, curr, f3(0.5, 0, 0), (-size / F(2), F(0), F(0)) if last else f3(-0.5, 0, 0))
        prev = curr


def scene_spring(sb):
    _ground(sb)
    anchor = sb.box((1, 1, 1), 0.0, 0.5, f3(0, 0, 14))
    block = sb.box((2, 2, 2), 1.0, 0.5, f3(0, 0, 8))
    sb.spring(anchor, block, f3(0, 0, 0), f3(0, 0, 0), 100.0, 4.0)


def scene_spring_ratio(sb):
    n = 8
    _ground(sb, pos=(0, 0, -10))
    prev = None
    for i in range(n):
        x = (F(i) - F(n - 1) * F(0.5)) * F(3.0)
        curr = sb.box((1, 0.75, 0.75), 0.0 if i in (0, n - 1) else 1.0, 0.5, (x, F(0), F(12)))
        if prev is not None:
            sb.spring(prev, curr, f3(0.5, 0, 0), f3(-0.5, 0, 0), 10.0 if i % 2 == 0 else 10000.0, 3.0)
        prev = curr


def scene_stack(sb):
    _ground(sb)
    for i in range(10):
        sb.box((1, 1, 1), 1.0, 0.5, (F(0), F(0), F(i) * F(1.5) + F(1.0)))


def scene_stack_ratio(sb):
    _ground(sb)
    top, s = F(0.5), F(1.0)
    for _ in range(4):
        half = s * F(0.5)
        cz = top + half
        sb.box((s, s, s), 1.0, 0.5, (F(0), F(0), cz))
        top = cz + half
        s = s * F(2.0)


# ---- round shapes, scenes.h:156-342 ----

def scene_sphere_stack(sb):
    _ground(sb, (1000, 1000, 1))
    r = F(0.5)
    for i in range(10):
        sb.sphere(r, 1.0, 0.5, (F(0), F(0), F(0.5) + r + F(i) * r * F(2.05)))


def scene_sphere_ramp(sb):
    _ground(sb, (1000, 1000, 1))
    rp, tangent, normal = _ramp(sb, (36, 16, 1), 0.8, 20.0)
    r = F(0.55)
    for i in range(8):
        fr = F(0.1) + F(i) * F(0.2)
        pos = add(add(add(rp, mul(tangent, F(-12.0) + F(i) * F(1.4))), (F(0), F(-3.5) + F(i), F(0))),
                  mul(normal, F(0.5) + r + F(0.02)))
        sb.sphere(r, 1.0, fr, pos)


def scene_capsule_stack(sb):
    _ground(sb, (1000, 1000, 1))
    r, hl = F(0.35), F(0.9)
    for i in range(10):
        q = half_angle_quat(1 if i % 2 == 0 else 0, rad(90.0))
        sb.capsule(r, hl, 1.0, 0.5, (F(0), F(0), F(0.5) + r + F(i) * r * F(2.4)), rot=q)


def scene_capsule_ramp(sb):
    _ground(sb, (1000, 1000, 1))
    rp, tangent, normal = _ramp(sb, (36, 16, 1), 0.8, 20.0)
    r, hl = F(0.35), F(0.9)
    for i in range(6):
        fr = F(0.2) + F(i) * F(0.15)
        pos = add(add(add(rp, mul(tangent, F(-12.0) + F(i) * F(2.0))), (F(0), F(-2.5) + F(i), F(0))),
                  mul(normal, F(0.5) + r + F(0.03)))
        sb.capsule(r, hl, 1.0, fr, pos, rot=half_angle_quat(1, rad(90.0)))


def scene_cylinder_stack(sb):
    _ground(sb, (1000, 1000, 1), friction=0.6)
    r, hl = F(0.48), F(0.35)
    for i in range(10):
        z = F(0.5) + hl + F(i) * (hl * F(2.0) + F(0.025))
        sb.cylinder(r, hl, 1.0, 0.6, (F(0), F(0), z))


def scene_cylinder_ramp(sb):
    _ground(sb, (1000, 1000, 1))
    rp, tangent, normal = _ramp(sb, (36, 16, 1), 0.8, 20.0)
    r, hl = F(0.48), F(0.65)
    for i in range(7):
        fr = F(0.2) + F(i) * F(0.12)
        pos = add(add(add(rp, mul(tangent, F(-12.0) + F(i) * F(2.1))), (F(0), F(-3.0) + F(i), F(0))),
                  mul(normal, F(0.5) + r + F(0.03)))
        sb.cylinder(r, hl, 1.0, fr, pos, rot=half_angle_quat(0, rad(-90.0)))


def scene_jitter(value: int):
    """sceneJitter (scenes.h:262-268): an integer hash mapped to [-1, 1] (32-bit arithmetic)."""
    m = 0xFFFFFFFF
    x = (value * 747796405 + 2891336453) & m
    x = (((x >> ((x >> 28) + 4)) ^ x) * 277803737) & m
    x = (x >> 22) ^ x
    return F(x & 0xFFFF) / F(32767.5) - F(1.0)


def _pour(sb, count, width, depth, start_z, layer_scale, drift, forward=1.1, down=-1.5, taper_column=True):
    """sceneSpherePourOnCylindersVariant (scenes.h:270-327)."""
    _ground(sb, (1000, 1000, 1), friction=0.7)
    cr, chl, cz_, cf, s = F(0.45), F(0.55), None, F(0.85), F(1.7)
    cz_ = F(0.5) + chl
    for x, y in ((F(0), F(0)), (-s, F(0)), (s, F(0)), (-s * F(0.5), s * F(0.9)), (s * F(0.5), s * F(0.9)),
                 (-s * F(0.5), -s * F(0.9)), (s * F(0.5), -s * F(0.9))):
        sb.cylinder(cr, chl, 1.0, cf, (x, y, cz_))
    r = F(0.12)
    spacing = r * F(2.35)
    layer_spacing = spacing * F(layer_scale)
    n, layer = 0, 0
    while n < count:
        for y in range(depth):
            for x in range(width):
                if n >= count:
                    break
                jx = scene_jitter(n * 3 + 0) * r * F(0.35)
                jy = scene_jitter(n * 3 + 1) * r * F(0.35)
                jz = scene_jitter(n * 3 + 2) * r * F(0.20)
                taper = F(1.0) - min(F(layer) / F(18.0), F(0.45)) if taper_column else F(1.0)
                px = (F(x) - F(width - 1) * F(0.5)) * spacing * taper + jx - F(1.2) + F(layer) * F(drift)
                py = (F(y) - F(depth - 1) * F(0.5)) * spacing * taper + jy - F(layer) * F(drift) * F(0.35)
                pz = F(start_z) + F(layer) * layer_spacing + jz
                vel = (F(forward) + scene_jitter(n * 5 + 0) * F(0.18), scene_jitter(n * 5 + 1) * F(0.35),
                       F(down) + scene_jitter(n * 5 + 2) * F(0.18))
                sb.sphere(r, 1.0, 0.45, (px, py, pz), vel=vel)
                n += 1
        layer += 1


def scene_pour(sb):
    _pour(sb, 1000, 8, 8, 10.0, 1.0, 0.0)


def scene_pour5000(sb):
    _pour(sb, 5000, 14, 8, 9.0, 1.45, 0.0, 0.35, -1.0, False)


def scene_pour20000(sb):
    _pour(sb, 20000, 20, 14, 9.0, 1.45, 0.0, 0.35, -1.0, False)


# ---- contact probes, scenes.h:344-660 ----

def _probe_ground(sb, size=(24, 24, 1)):
    _ground(sb, size, friction=0.7)


def scene_probe_sphere_sphere(sb):
    _probe_ground(sb)
    r, cz = F(0.35), F(4.0)
    sb.sphere(r, 1.0, 0.5, (-r * F(0.45), F(0), cz))
    sb.sphere(r, 1.0, 0.5, (r * F(0.45), F(0), cz))


def scene_probe_sphere_cylinder(sb):
    _probe_ground(sb)
    cr, chl, sr, cz = F(0.45), F(0.55), F(0.2), F(4.0)
    sb.cylinder(cr, chl, 1.0, 0.8, (F(0), F(0), cz))
    sb.sphere(sr, 1.0, 0.5, (cr + sr * F(0.45), F(0), cz))


def scene_probe_sphere_box(sb):
    _probe_ground(sb)
    sr, top = F(0.2), F(4.0)
    sb.box((0.8, 0.8, 1.0), 1.0, 0.7, (F(0), F(0), top))
    sb.sphere(sr, 1.0, 0.5, (F(0.4) + sr * F(0.45), F(0), top))


def scene_probe_sphere_capsule(sb):
    _probe_ground(sb)
    cr, chl, sr, cz = F(0.35), F(0.55), F(0.2), F(4.0)
    sb.capsule(cr, chl, 1.0, 0.8, (F(0), F(0), cz))
    sb.sphere(sr, 1.0, 0.5, (cr + sr * F(0.45), F(0), cz))


def _rotated_probe(sb, kind, density, angle, axis):
    sr = F(0.2)
    pos = (F(0), F(0), F(2.0))
    q = half_angle_quat(axis, angle)
    if kind == "cylinder":
        cr = F(0.45)
        sb.cylinder(cr, F(0.55), density, 0.8, pos, rot=q)
        off = cr + sr * F(0.45)
    elif kind == "capsule":
        cr = F(0.35)
        sb.capsule(cr, F(0.55), density, 0.8, pos, rot=q)
        off = cr + sr * F(0.45)
    else:
        sb.box((0.8, 0.6, 1.0), density, 0.7, pos, rot=q)
        off = F(0.4) + sr * F(0.45)
    sb.sphere(sr, 1.0, 0.5, add(pos, rotate(q, (off, F(0), F(0)))))


def scene_probe_rot_cylinder(sb):
    _probe_ground(sb)
    _rotated_probe(sb, "cylinder", 0.0, 0.65, 1)


def scene_probe_rot_capsule(sb):
    _probe_ground(sb)
    _rotated_probe(sb, "capsule", 0.0, 0.65, 1)


def scene_probe_rot_box(sb):
    _probe_ground(sb)
    _rotated_probe(sb, "box", 0.0, 0.7, 2)


def scene_probe_dyn_rot_cylinder(sb):
    _probe_ground(sb)
    _rotated_probe(sb, "cylinder", 1.0, 0.65, 1)


def scene_probe_dyn_rot_capsule(sb):
    _probe_ground(sb)
    _rotated_probe(sb, "capsule", 1.0, 0.65, 1)


def scene_probe_dyn_rot_box(sb):
    _probe_ground(sb)
    _rotated_probe(sb, "box", 1.0, 0.7, 2)


def scene_probe_capsule_capsule(sb):
    _probe_ground(sb)
    r, hl, cz = F(0.35), F(0.55), F(4.0)
    sb.capsule(r, hl, 1.0, 0.8, (F(0), F(0), cz))
    sb.capsule(r, hl, 1.0, 0.8, (r * F(1.2), F(0), cz))


def scene_probe_cylinder_cylinder(sb):
    _probe_ground(sb)
    r, hl, cz = F(0.45), F(0.55), F(4.0)
    sb.cylinder(r, hl, 1.0, 0.8, (F(0), F(0), cz))
    sb.cylinder(r, hl, 1.0, 0.8, (r * F(1.2), F(0), cz))


def scene_probe_capsule_cylinder(sb):
    _probe_ground(sb)
    capr, cylr, hl, cz = F(0.35), F(0.45), F(0.55), F(4.0)
    sb.capsule(capr, hl, 1.0, 0.8, (F(0), F(0), cz))
    sb.cylinder(cylr, hl, 1.0, 0.8, (capr + cylr * F(0.45), F(0), cz))


def scene_probe_rot_capsule_cylinder(sb):
    _probe_ground(sb)
    capr, cylr, hl, cz = F(0.35), F(0.45), F(0.55), F(4.0)
    qc = half_angle_quat(1, 0.65)
    pos = (F(0), F(0), cz)
    sb.capsule(capr, hl, 1.0, 0.8, pos, rot=qc)
    off = rotate(qc, (capr + cylr * F(0.45), F(0), F(0)))
    sb.cylinder(cylr, hl, 1.0, 0.8, add(pos, off), rot=half_angle_quat(0, -0.55))


def scene_probe_mixed(sb):
    _probe_ground(sb, (32, 24, 1))
    z, sr, capr, cylr, hl = F(2.0), F(0.2), F(0.35), F(0.45), F(0.55)
    q = half_angle_quat(1, 0.5)
    p = (F(-6.0), F(0), z)
    sb.cylinder(cylr, hl, 1.0, 0.8, p, rot=q)
    sb.sphere(sr, 1.0, 0.5, add(p, rotate(q, (cylr + sr * F(0.45), F(0), F(0)))))
    q = half_angle_quat(1, 0.55)
    p = (F(-3.5), F(0), z)
    sb.capsule(capr, hl, 1.0, 0.8, p, rot=q)
    sb.sphere(sr, 1.0, 0.5, add(p, rotate(q, (capr + sr * F(0.45), F(0), F(0)))))
    q = half_angle_quat(2, 0.65)
    p = (F(-1.0), F(0), z)
    sb.box((0.8, 0.6, 1.0), 1.0, 0.7, p, rot=q)
    sb.sphere(sr, 1.0, 0.5, add(p, rotate(q, (F(0.4) + sr * F(0.45), F(0), F(0)))))
    sb.capsule(capr, hl, 1.0, 0.8, (F(1.5), F(0), z))
    sb.capsule(capr, hl, 1.0, 0.8, (F(1.5) + capr * F(1.2), F(0), z))
    sb.cylinder(cylr, hl, 1.0, 0.8, (F(4.0), F(0), z))
    sb.cylinder(cylr, hl, 1.0, 0.8, (F(4.0) + cylr * F(1.2), F(0), z))
    sb.capsule(capr, hl, 1.0, 0.8, (F(6.5), F(0), z), rot=half_angle_quat(1, 0.45))
    sb.cylinder(cylr, hl, 1.0, 0.8, (F(6.5) + capr + cylr * F(0.45), F(0), z), rot=half_angle_quat(0, -0.35))
    sb.box((0.8, 0.6, 1.0), 1.0, 0.7, (F(-2.0), F(3.0), z))
    sb.capsule(capr, hl, 1.0, 0.8, (F(-2.0) + F(0.4) + capr * F(0.45), F(3.0), z))
    sb.box((0.8, 0.6, 1.0), 1.0, 0.7, (F(2.0), F(3.0), z))
    sb.cylinder(cylr, hl, 1.0, 0.8, (F(2.0) + F(0.4) + cylr * F(0.45), F(3.0), z))


def scene_probe_capsule_box(sb):
    _probe_ground(sb)
    capr, hl, cz = F(0.35), F(0.55), F(4.0)
    sb.box((0.8, 0.6, 1.0), 1.0, 0.7, (F(0), F(0), cz))
    sb.capsule(capr, hl, 1.0, 0.8, (F(0.4) + capr * F(0.45), F(0), cz))


def scene_probe_cylinder_box(sb):
    _probe_ground(sb)
    cylr, hl, cz = F(0.45), F(0.55), F(2.0)
    sb.box((0.8, 0.6, 1.0), 1.0, 0.7, (F(0), F(0), cz))
    sb.cylinder(cylr, hl, 1.0, 0.8, (F(0.4) + cylr * F(0.45), F(0), cz))


def scene_probe_box_box(sb):
    _probe_ground(sb)
    sb.box((0.9, 0.65, 1.0), 1.0, 0.7, f3(-0.2, 0, 4))
    sb.box((0.9, 0.65, 1.0), 1.0, 0.7, f3(0.45, 0, 4))


def scene_probe_rot_box_box(sb):
    _probe_ground(sb)
    sb.box((1.1, 0.65, 1.0), 1.0, 0.7, f3(-0.18, 0, 4), rot=half_angle_quat(2, 0.55))
    sb.box((1.1, 0.65, 1.0), 1.0, 0.7, f3(0.42, 0.03, 4), rot=half_angle_quat(2, -0.35))


def scene_probe_box_cluster(sb):
    _probe_ground(sb)
    sp = F(0.78)
    for y in range(3):
        for x in range(3):
            pos = (F(x - 1) * sp + (F(0.10) if y & 1 else F(0)), F(y - 1) * sp * F(0.85), F(4.0))
            sb.box((0.9, 0.75, 1.0), 1.0, 0.7, pos, rot=half_angle_quat(2, F(0.10) * F(x - y)))


# ---- Newton's cradle, scenes.h:662-774 ----

def _align_z(d):
    """alignZToVector (scenes.h:662-675)."""
    d = normalize(d)
    c = dot(f3(0, 0, 1), d)
    if c > F(0.9999):
        return (F(0), F(0), F(0), F(1))
    if c < F(-0.9999):
        return (F(1), F(0), F(0), F(0))
    axis = cross(f3(0, 0, 1), d)
    s = F(np.sqrt((F(1.0) + c) * F(2.0)))
    inv = F(1.0) / s
    q = (axis[0] * inv, axis[1] * inv, axis[2] * inv, s * F(0.5))
    n = F(np.sqrt((q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) + q[3] * q[3]))
    return (q[0] / n, q[1] / n, q[2] / n, q[3] / n)


def _string(sb, ball, ball_pos, anchor, local, segments, radius):
    """addCapsuleString (scenes.h:688-714): capsule segments from a world anchor to the ball."""
    ball_anchor = add(ball_pos, local)   # the ball starts unrotated: rotate(identity, v) = v
    prev, prev_half = None, F(0)
    for i in range(segments):
        t0 = F(i) / F(segments)
        t1 = F(i + 1) / F(segments)
        a = add(anchor, mul(sub(ball_anchor, anchor), t0))
        b = add(anchor, mul(sub(ball_anchor, anchor), t1))
        span = sub(b, a)
        seg_len = max(length(span), radius * F(2.0))
        hl = seg_len * F(0.5)
        seg = sb.capsule(radius, hl, 0.08, 0.6, mul(add(a, b), 0.5), rot=_align_z(span))
        if prev is not None:
            sb.joint(prev, seg, (F(0), F(0), prev_half), (F(0), F(0), -hl), INF, 0.0)
        else:
            sb.joint(-1, seg, anchor, (F(0), F(0), -hl), INF, 0.0)
        prev, prev_half = seg, hl
    sb.joint(prev, ball, (F(0), F(0), prev_half), local, INF, 0.0)
    return prev


def scene_newtons_cradle(sb):
    sb.box((20, 16, 1), 0.0, 0.5, f3(0, 0, -1))
    for size, pos in (((7.0, 0.16, 0.16), (0, -0.65, 8.0)), ((7.0, 0.16, 0.16), (0, 0.65, 8.0)),
                      ((0.16, 0.16, 7.5), (-3.5, -0.65, 4.2)), ((0.16, 0.16, 7.5), (3.5, -0.65, 4.2)),
                      ((0.16, 0.16, 7.5), (-3.5, 0.65, 4.2)), ((0.16, 0.16, 7.5), (3.5, 0.65, 4.2))):
        sb.box(size, 0.0, 0.5, f3(*pos))
    count, r = 5, F(0.45)
    spacing = r * F(2.02)
    anchor_z, rest_z, sy, srad, pull = F(8.0), F(3.25), F(0.48), F(0.045), F(-1.65)
    last = []
    for i in range(count):
        rest_x = (F(i) - F(count - 1) * F(0.5)) * spacing
        x, z, vx = rest_x, rest_z, F(0)
        if i == 0:
            string_len = anchor_z - (rest_z + r * F(0.7))
            x = x + pull
            z = anchor_z - F(np.sqrt(max(string_len * string_len - pull * pull, F(0)))) - r * F(0.7)
            vx = F(1.0)
        pos = (x, F(0), z)
        ball = sb.sphere(r, 1.0, 0.65, pos, vel=(vx, F(0), F(0)))
        front = _string(sb, ball, pos, (rest_x, -sy, anchor_z), (F(0), -sy * F(0.55), r * F(0.7)), 5, srad)
        back = _string(sb, ball, pos, (rest_x, sy, anchor_z), (F(0), sy * F(0.55), r * F(0.7)), 5, srad)
        last.append((ball, front, back))
    # IgnoreCollision between each ball and the capsules jointed to it, visiting the force list
    # newest first (the back string's last joint, then the front one's)
    for ball, front, back in last:
        sb.ignore(ball, back)
        sb.ignore(ball, front)


# ---- soft bodies, scenes.h:776-988 ----

def _soft(sb, klin, kang, w, d, h, n, size):
    _ground(sb)
    size = F(size)
    half = size * F(0.5)
    base_z, gap = F(8.0), F(2.0)
    for i in range(n):
        sx = (F(i) - F(n - 1) * F(0.5)) * (F(w) * size + gap)
        g = {}
        for x in range(w):
            for y in range(d):
                for z in range(h):
                    px = sx + (F(x) - F(w - 1) * F(0.5)) * size
                    py = (F(y) - F(d - 1) * F(0.5)) * size
                    pz = base_z + F(0) + F(z) * size
                    g[x, y, z] = sb.box((size, size, size), 1.0, 0.5, (px, py, pz))
        for x in range(1, w):
            for y in range(d):
                for z in range(h):
                    sb.joint(g[x - 1, y, z], g[x, y, z], (half, F(0), F(0)), (-half, F(0), F(0)), klin, kang)
        for x in range(w):
            for y in range(1, d):
                for z in range(h):
                    sb.joint(g[x, y - 1, z], g[x, y, z], (F(0), half, F(0)), (F(0), -half, F(0)), klin, kang)
        for x in range(w):
            for y in range(d):
                for z in range(1, h):
                    sb.joint(g[x, y, z - 1], g[x, y, z], (F(0), F(0), half), (F(0), F(0), -half), klin, kang)
        for x in range(1, w):
            for y in range(d):
                for z in range(1, h):
                    sb.ignore(g[x - 1, y, z - 1], g[x, y, z])
                    sb.ignore(g[x, y, z - 1], g[x - 1, y, z])
        for x in range(w):
            for y in range(1, d):
                for z in range(1, h):
                    sb.ignore(g[x, y - 1, z - 1], g[x, y, z])
                    sb.ignore(g[x, y, z - 1], g[x, y - 1, z])
        for x in range(1, w):
            for y in range(1, d):
                for z in range(h):
                    sb.ignore(g[x - 1, y - 1, z], g[x, y, z])
                    sb.ignore(g[x, y - 1, z], g[x - 1, y, z])


def scene_soft_body(sb):
    _soft(sb, 5000.0, 1000.0, 4, 4, 4, 3, 0.8)


def scene_soft_body_fine(sb):
    _soft(sb, 4000.0, 800.0, 8, 8, 8, 3, 0.4)


# ---- cloth, mesh torus, bridge, breakable, scenes.h:990-1144 ----

def scene_cloth(sb):
    _ground(sb, friction=0.7)
    w = d = 30
    spacing, radius = F(0.25), F(0.12)
    half = spacing * F(0.5)
    klin, kshear = 1500.0, 450.0
    g = {}
    for x in range(w):
        for y in range(d):
            px = (F(x) - F(w - 1) * F(0.5)) * spacing
            py = (F(y) - F(d - 1) * F(0.5)) * spacing
            g[x, y] = sb.sphere(radius, 12.0, 0.6, (px, py, F(5.0)))
    for x in range(1, w):
        for y in range(d):
            sb.joint(g[x - 1, y], g[x, y], (half, F(0), F(0)), (-half, F(0), F(0)), klin, 0.0)
    for x in range(w):
        for y in range(1, d):
            sb.joint(g[x, y - 1], g[x, y], (F(0), half, F(0)), (F(0), -half, F(0)), klin, 0.0)
    for x in range(1, w):
        for y in range(1, d):
            sb.joint(g[x - 1, y - 1], g[x, y], (half, half, F(0)), (-half, -half, F(0)), kshear, 0.0)
            sb.joint(g[x - 1, y], g[x, y - 1], (half, -half, F(0)), (-half, half, F(0)), kshear, 0.0)
    for cx, cy in ((0, 0), (w - 1, 0), (0, d - 1), (w - 1, d - 1)):
        c = g[cx, cy]
        sb.joint(-1, c, tuple(sb.b["pos"][c]), f3(0, 0, 0), INF, 0.0)
    for p in ((0.0, 0.0, 6.2), (-1.6, -1.2, 6.8), (1.5, -1.5, 7.4), (-1.3, 1.6, 8.0), (1.2, 1.3, 8.6)):
        sb.sphere(F(0.45), 2.0, 0.5, f3(*p))


def scene_mesh_torus(sb):
    from . import mesh_lattice
    _ground(sb, friction=0.7)
    torus = mesh_lattice.torus_mesh(1.0, 0.45, 28, 14)
    soft = mesh_lattice.LatticeParams(spacing=0.26)
    mesh_lattice.build_lattice(sb, torus, f3(-2.0, 0, 3.0), 1.0, soft)
    hard = mesh_lattice.LatticeParams(spacing=0.26, stiffness_lin=INF, stiffness_ang=INF)
    mesh_lattice.build_lattice(sb, torus, f3(2.0, 0, 3.0), 1.0, hard)


def scene_bridge(sb):
    n = 40
    pl, pw, ph = F(1.0), F(4.0), F(0.5)
    hl, hw = pl * F(0.5), pw * F(0.5)
    _ground(sb)
    prev = None
    for i in range(n):
        curr = sb.box((pl, pw, ph), 0.0 if i in (0, n - 1) else 1.0, 0.5, (F(i) - F(n / 2.0), F(0), F(10.0)))
        if prev is not None:
            sb.joint(prev, curr, (hl, hw, F(0)), (-hl, hw, F(0)), INF, 0.0)
            sb.joint(prev, curr, (hl, -hw, F(0)), (-hl, -hw, F(0)), INF, 0.0)
        prev = curr
    for x in range(n // 4):
        for y in range(n // 8):
            sb.box((1, 1, 1), 1.0, 0.5, (F(x) - F(n / 8.0), F(0), F(y) + F(12.0)))


def scene_breakable(sb):
    n, m = 10, 5
    _ground(sb)
    prev = None
    for i in range(n + 1):
        curr = sb.box((1, 1, 0.5), 1.0, 0.5, (F(i) - F(n / 2.0), F(0), F(6.0)))
        if prev is not None:
            sb.joint(prev, curr, f3(0.5, 0, 0), f3(-0.5, 0, 0), INF, INF, 90.0)
        prev = curr
    sb.box((1, 1, 5), 0.0, 0.5, (F(-n / 2.0), F(0), F(2.5)))
    sb.box((1, 1, 5), 0.0, 0.5, (F(n / 2.0), F(0), F(2.5)))
    for i in range(m):
        sb.box((2, 1, 1), 1.0, 0.5, (F(0), F(0), F(i) * F(2.0) + F(8.0)))


# sceneNames / scenes (scenes.h:1146-1245), in index order
SCENES = [
    ("Empty", scene_empty), ("Ground", scene_ground), ("Dynamic Friction", scene_dynamic_friction),
    ("Static Friction", scene_static_friction), ("Pyramid", scene_pyramid), ("Rope", scene_rope),
    ("Heavy Rope", scene_heavy_rope), ("Spring", scene_spring), ("Spring Ratio", scene_spring_ratio),
    ("Stack", scene_stack), ("Stack Ratio", scene_stack_ratio), ("Sphere Stack", scene_sphere_stack),
    ("Sphere Ramp", scene_sphere_ramp), ("Capsule Stack", scene_capsule_stack), ("Capsule Ramp", scene_capsule_ramp),
    ("Cylinder Stack", scene_cylinder_stack), ("Cylinder Ramp", scene_cylinder_ramp),
    ("Sphere Pour on Cylinders", scene_pour), ("Sphere Pour 5000 on Cylinders", scene_pour5000),
    ("Sphere Pour 20000 on Cylinders", scene_pour20000),
    ("Sphere Sphere Contact Probe", scene_probe_sphere_sphere),
    ("Sphere Cylinder Contact Probe", scene_probe_sphere_cylinder),
    ("Sphere Capsule Contact Probe", scene_probe_sphere_capsule),
    ("Sphere Box Contact Probe", scene_probe_sphere_box),
    ("Sphere Rotated Cylinder Contact Probe", scene_probe_rot_cylinder),
    ("Sphere Rotated Capsule Contact Probe", scene_probe_rot_capsule),
    ("Sphere Rotated Box Contact Probe", scene_probe_rot_box),
    ("Sphere Dynamic Rotated Cylinder Contact Probe", scene_probe_dyn_rot_cylinder),
    ("Sphere Dynamic Rotated Capsule Contact Probe", scene_probe_dyn_rot_capsule),
    ("Sphere Dynamic Rotated Box Contact Probe", scene_probe_dyn_rot_box),
    ("Capsule Capsule Contact Probe", scene_probe_capsule_capsule),
    ("Cylinder Cylinder Contact Probe", scene_probe_cylinder_cylinder),
    ("Capsule Cylinder Contact Probe", scene_probe_capsule_cylinder),
    ("Rotated Capsule Cylinder Contact Probe", scene_probe_rot_capsule_cylinder),
    ("Mixed Primitive Contact Probe", scene_probe_mixed),
    ("Capsule Box Contact Probe", scene_probe_capsule_box),
    ("Cylinder Box Contact Probe", scene_probe_cylinder_box),
    ("Box Box Contact Probe", scene_probe_box_box),
    ("Rotated Box Box Contact Probe", scene_probe_rot_box_box),
    ("Box Cluster Contact Probe", scene_probe_box_cluster),
    ("Newton's Cradle", scene_newtons_cradle), ("Soft Body", scene_soft_body),
    ("Soft Body 8x8x8", scene_soft_body_fine), ("Cloth", scene_cloth),
    ("Mesh Torus Soft+Rigid", scene_mesh_torus), ("Bridge", scene_bridge), ("Breakable", scene_breakable),
]
NAMES = [n for n, _ in SCENES]


def normalize_name(text: str) -> str:
    """SimulationHost::normalizeSceneName: lower-case letters and digits only."""
    return "".join(c.lower() for c in text if c.isalnum())


def index_of(name: str) -> int:
    want = normalize_name(name)
    for i, n in enumerate(NAMES):
        if normalize_name(n) == want:
            return i
    raise KeyError(f"unknown scene '{name}'")


def build(name_or_index, steps: int = 600) -> Scene:
    """The named scene as a Scene (creation order; the reference's default gravity -10)."""
    i = name_or_index if isinstance(name_or_index, int) else index_of(name_or_index)
    name, fn = SCENES[i]
    sb = SceneBuilder()
    fn(sb)
    return sb.scene(name, -10.0, steps)
