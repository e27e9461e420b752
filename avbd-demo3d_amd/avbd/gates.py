"""The reference's headless benchmark metrics for its pour gates (headless_server.cpp:240-495).

Computed on the host from a downloaded state (numpy, FP32 inputs):

* ground top: the highest top face of the static, axis-up box slabs at least 8 x 8 wide and at
  most 2 thick (isBenchmarkGroundReceiver, :384-393);
* ground clearance of each dynamic body: centre z - its vertical contact extent - ground top,
  zeroed within 1e-5 (:405-495); spheres' penetrations give the sphere-ground gate;
* final max speed: the largest |v| over all bodies (:283-286);
* final dynamic max height: the top of the dynamic bodies' bounding spheres, max(z + radius)
  (:311-330).
"""
from __future__ import annotations

import numpy as np

from .state import SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CYLINDER, SHAPE_SPHERE


def _axis(rot, k):
    """Column k of the rotation of each quaternion (x, y, z, w): the body's k-th local axis."""
    x, y, z, w = (rot[:, i].astype(np.float64) for i in range(4))
    if k == 0:
        return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], 1)
    if k == 1:
        return np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], 1)
    return np.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], 1)


def contact_extent_z(b) -> np.ndarray:
    """Vertical half extent of each body (benchmarkGroundContactExtentZ, :405-433)."""
    shape = b["shape"]
    rot = b["rot"]
    ax, ay, az = _axis(rot, 0), _axis(rot, 1), _axis(rot, 2)
    size = b["size"].astype(np.float64)
    r = np.maximum(b["shape_radius"].astype(np.float64), 0.0)
    hl = np.maximum(b["half_length"].astype(np.float64), 0.0)
    box = (np.abs(ax[:, 2]) * np.maximum(size[:, 0] * 0.5, 0) + np.abs(ay[:, 2]) * np.maximum(size[:, 1] * 0.5, 0)
           + np.abs(az[:, 2]) * np.maximum(size[:, 2] * 0.5, 0))
    sph = np.where(r > 0, r, b["radius"].astype(np.float64))
    cap = np.abs(az[:, 2]) * hl + r
    cyl = np.abs(az[:, 2]) * hl + np.sqrt(np.maximum(0.0, 1.0 - az[:, 2] ** 2)) * r
    out = np.where(shape == SHAPE_SPHERE, sph, np.where(shape == SHAPE_CAPSULE, cap,
                                                         np.where(shape == SHAPE_CYLINDER, cyl, box)))
    return np.where(shape == SHAPE_BOX, box, out)


def pour_metrics(state) -> dict:
    b = state.bodies
    dyn = b["mass"] > 0
    up = _axis(b["rot"], 2)[:, 2]
    ground = (~dyn) & (b["shape"] == SHAPE_BOX) & (b["size"][:, 0] >= 8) & (b["size"][:, 1] >= 8) & \
             (b["size"][:, 2] <= 2) & (up > 0.98)
    out = {"dynamic_bodies": int(dyn.sum())}
    if ground.any():
        top = float((b["pos"][ground, 2].astype(np.float64) + b["size"][ground, 2] * 0.5).max())
        clear = b["pos"][:, 2].astype(np.float64) - contact_extent_z(b) - top
        clear = np.where(np.abs(clear) < 1e-5, 0.0, clear)
        sph = dyn & (b["shape"] == SHAPE_SPHERE)
        pen = np.where(clear < -1e-5, -clear, 0.0)
        out.update(ground_top=top,
                   max_sphere_ground_penetration=float(pen[sph].max()) if sph.any() else 0.0,
                   below_ground_spheres=int((sph & (clear < -1e-5)).sum()),
                   max_ground_body_penetration=float(pen[dyn].max()) if dyn.any() else 0.0)
    speed = np.linalg.norm(b["vel"].astype(np.float64), axis=1)
    out["final_max_speed"] = float(speed.max()) if len(speed) else 0.0
    if dyn.any():
        out["final_dynamic_max_height"] = float((b["pos"][dyn, 2].astype(np.float64) + b["radius"][dyn]).max())
    return out
