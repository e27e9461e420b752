"""NumPy mirrors of the POD records in include/avbd_hip.h.

The state tables (bodies, joints, springs, manifolds) are the exchange format
of the C ABI. Each dtype below is laid out exactly like its C struct
(``align=True`` reproduces the C alignment rules); ``check_layouts`` compares
the sizes against what a compiled library reports.
"""
from __future__ import annotations

import numpy as np

MAX_CONTACTS = 8

PARAMS = np.dtype(
    [("dt", "<f4"), ("gravity", "<f4"), ("iterations", "<i4"), ("alpha", "<f4"),
     ("beta_lin", "<f4"), ("beta_ang", "<f4"), ("gamma", "<f4")], align=True)

BODY = np.dtype(
    [("size", "<f4", 3), ("mass", "<f4"), ("moment", "<f4", 3), ("friction", "<f4"),
     ("radius", "<f4"), ("pos", "<f4", 3), ("rot", "<f4", 4), ("initial_lin", "<f4", 3),
     ("initial_ang", "<f4", 4), ("inertial_lin", "<f4", 3), ("inertial_ang", "<f4", 4),
     ("vel", "<f4", 3), ("ang_vel", "<f4", 3), ("prev_vel", "<f4", 3), ("shape_radius", "<f4"),
     ("half_length", "<f4"), ("shape", "<i4"), ("dense_id", "<i4")], align=True)

# RigidShapeType (solver.h:47-53)
SHAPE_BOX, SHAPE_SPHERE, SHAPE_CAPSULE, SHAPE_CYLINDER = 0, 1, 2, 3

CONTACT = np.dtype(
    [("feature", "<i4"), ("r_a", "<f4", 3), ("r_b", "<f4", 3), ("c0", "<f4", 3),
     ("penalty", "<f4", 3), ("lambda", "<f4", 3), ("stick", "<i4")], align=True)

MANIFOLD = np.dtype(
    [("body_a", "<i4"), ("body_b", "<i4"), ("serial", "<i8"), ("num_contacts", "<i4"),
     ("friction", "<f4"), ("basis", "<f4", 9), ("contacts", CONTACT, MAX_CONTACTS)], align=True)

EXTERNAL_CONTACT = np.dtype(
    [("body_a", "<i4"), ("body_b", "<i4"), ("num_contacts", "<i4"), ("pad_", "<i4"), ("basis", "<f4", 9),
     ("contacts", CONTACT, MAX_CONTACTS)], align=True)

JOINT = np.dtype(
    [("body_a", "<i4"), ("body_b", "<i4"), ("serial", "<i8"), ("r_a", "<f4", 3), ("r_b", "<f4", 3),
     ("c0_lin", "<f4", 3), ("c0_ang", "<f4", 3), ("penalty_lin", "<f4", 3), ("penalty_ang", "<f4", 3),
     ("lambda_lin", "<f4", 3), ("lambda_ang", "<f4", 3), ("stiffness_lin", "<f4"),
     ("stiffness_ang", "<f4"), ("fracture", "<f4"), ("torque_arm", "<f4"), ("broken", "<i4"),
     ("pad_", "<i4")], align=True)

SPRING = np.dtype(
    [("body_a", "<i4"), ("body_b", "<i4"), ("serial", "<i8"), ("r_a", "<f4", 3), ("r_b", "<f4", 3),
     ("rest", "<f4"), ("stiffness", "<f4")], align=True)

STATS = np.dtype(
    [("body_count", "<i4"), ("active_body_count", "<i4"), ("solved_body_count", "<i4"),
     ("pair_checks", "<i4"), ("sphere_hits", "<i4"), ("constrained_hits", "<i4"),
     ("manifolds_created", "<i4"), ("manifold_count", "<i4"), ("contact_count", "<i4"),
     ("joint_count", "<i4"), ("spring_count", "<i4"), ("color_count", "<i4"),
     ("coloring_rounds", "<i4"), ("sweep_axis", "<i4"), ("axis_fallbacks", "<i4"),
     ("cpu_fallbacks", "<i4"), ("steps_taken", "<i8"), ("broadphase_ms", "<f4"),
     ("narrowphase_ms", "<f4"), ("coloring_ms", "<f4"), ("iterate_ms", "<f4"),
     ("finish_ms", "<f4"), ("step_ms", "<f4"), ("iterate_bytes", "<f8"), ("step_bytes", "<f8"),
     ("iterate_bytes_total", "<f8"), ("step_bytes_total", "<f8"), ("iterate_ms_total", "<f8"),
     ("step_ms_total", "<f8"), ("counted_steps", "<i8"), ("timed_steps", "<i8"),
     ("primal_bytes_total", "<f8"), ("dual_bytes_total", "<f8"), ("primal_ms_total", "<f8"),
     ("dual_ms_total", "<f8"), ("primal_launches", "<i8"), ("dual_launches", "<i8"),
     ("primal_iterations", "<i8")], align=True)

CONFIG = np.dtype(
    [("params", PARAMS), ("max_bodies", "<i4"), ("max_joints", "<i4"), ("max_springs", "<i4"),
     ("max_manifolds", "<i4"), ("max_pairs", "<i4"), ("device", "<i4"), ("enable_timing", "<i4")],
    align=True)

EXPECTED_SIZES = {"body": 172, "contact": 68, "manifold": 608, "joint": 136, "spring": 48, "params": 28}


def default_params(gravity: float = -10.0, iterations: int = 10, dt: float = 1.0 / 60.0) -> np.ndarray:
    """Solver::defaultParams (solver.cpp:909-932) with overridable gravity/iterations/dt."""
    p = np.zeros((), PARAMS)
    p["dt"] = np.float32(dt)
    p["gravity"] = np.float32(gravity)
    p["iterations"] = iterations
    p["alpha"] = np.float32(0.99)
    p["beta_lin"] = np.float32(10000.0)
    p["beta_ang"] = np.float32(100.0)
    p["gamma"] = np.float32(0.999)
    return p


def as_bodies(arr) -> np.ndarray:
    """Body records in the BODY layout, copied field by field. Records written before the shape
    fields existed are boxes, whose RigidShape radius is their bounding radius."""
    a = np.asarray(arr)
    if a.dtype == BODY:
        return np.ascontiguousarray(a)
    out = np.zeros(a.shape, BODY)
    for name in a.dtype.names:
        if name in BODY.names:
            out[name] = a[name]
    if "shape_radius" not in a.dtype.names:
        out["shape_radius"] = np.where(out["shape"] == SHAPE_BOX, out["radius"], np.float32(0))
    return out


class State:
    """A full solver state: bodies in list order, forces in ascending serial."""

    def __init__(self, bodies, joints=None, springs=None, manifolds=None, next_serial=0, ignores=None):
        self.bodies = as_bodies(bodies)
        # IgnoreCollision pairs (list indices), shape (n, 2)
        self.ignores = np.ascontiguousarray(np.asarray(ignores if ignores is not None else np.zeros((0, 2)),
                                                       np.int32).reshape(-1, 2))
        self.joints = np.ascontiguousarray(joints if joints is not None else np.zeros(0, JOINT), dtype=JOINT)
        self.springs = np.ascontiguousarray(springs if springs is not None else np.zeros(0, SPRING), dtype=SPRING)
        self.manifolds = np.ascontiguousarray(
            manifolds if manifolds is not None else np.zeros(0, MANIFOLD), dtype=MANIFOLD)
        self.next_serial = int(next_serial)

    def copy(self) -> "State":
        return State(self.bodies.copy(), self.joints.copy(), self.springs.copy(), self.manifolds.copy(),
                     self.next_serial, self.ignores.copy())

    @property
    def num_dynamic(self) -> int:
        return int((self.bodies["mass"] > 0).sum())

    def save(self, path) -> None:
        np.savez_compressed(path, bodies=self.bodies, joints=self.joints, springs=self.springs,
                            manifolds=self.manifolds, next_serial=np.int64(self.next_serial), ignores=self.ignores)

    @staticmethod
    def load(path) -> "State":
        with np.load(path, allow_pickle=False) as z:
            return State(z["bodies"], z["joints"], z["springs"], z["manifolds"], int(z["next_serial"]),
                         z["ignores"] if "ignores" in z else None)


def batch_states(states) -> tuple[State, np.ndarray]:
    """Concatenate per-world states into one batched state (avbd_batch_upload).

    Body indices of forces are shifted by the world's body offset (-1, the world
    anchor, stays -1); serials are shifted by the preceding worlds' next_serial,
    so they stay unique over the batch and keep each world's order.
    Returns (state, world_body_offsets).
    """
    offsets = np.zeros(len(states) + 1, np.int32)
    bodies, joints, springs, manifolds, ignores = [], [], [], [], []
    serial0 = 0
    for w, st in enumerate(states):
        b0 = int(offsets[w])
        offsets[w + 1] = b0 + len(st.bodies)
        bodies.append(st.bodies)
        j = st.joints.copy()
        j["body_a"] = np.where(j["body_a"] >= 0, j["body_a"] + b0, -1)
        j["body_b"] += b0
        j["serial"] += serial0
        sp = st.springs.copy()
        sp["body_a"] += b0
        sp["body_b"] += b0
        sp["serial"] += serial0
        m = st.manifolds.copy()
        m["body_a"] += b0
        m["body_b"] += b0
        m["serial"] += serial0
        joints.append(j)
        springs.append(sp)
        manifolds.append(m)
        ignores.append(st.ignores + b0)
        serial0 += st.next_serial
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.zeros(0, dt)
    return (State(cat(bodies, BODY), cat(joints, JOINT), cat(springs, SPRING), cat(manifolds, MANIFOLD), serial0,
                  np.concatenate(ignores) if ignores else None), offsets)


def split_state(state: State, offsets) -> list:
    """Inverse of batch_states up to serial values: per-world states with local
    body indices. Serials keep their batch values (each world's order is kept)."""
    out = []
    for w in range(len(offsets) - 1):
        b0, b1 = int(offsets[w]), int(offsets[w + 1])
        inw = lambda idx: (idx >= b0) & (idx < b1)
        j = state.joints[inw(state.joints["body_b"])].copy()
        j["body_a"] = np.where(j["body_a"] >= 0, j["body_a"] - b0, -1)
        j["body_b"] -= b0
        sp = state.springs[inw(state.springs["body_b"])].copy()
        sp["body_a"] -= b0
        sp["body_b"] -= b0
        m = state.manifolds[inw(state.manifolds["body_a"])].copy()
        m["body_a"] -= b0
        m["body_b"] -= b0
        ig = state.ignores[inw(state.ignores[:, 0])] - b0
        out.append(State(state.bodies[b0:b1].copy(), j, sp, m, state.next_serial, ig))
    return out


LAUNCH = np.dtype([("primal_lanes", "<i4"), ("primal_threads", "<i4"), ("dual_threads", "<i4"),
                   ("use_graph", "<i4"), ("dataflow", "<i4"), ("lazy_dual", "<i4")], align=True)


def check_layouts(sizeof, records: int = 10) -> None:
    """Compare dtype sizes with a library's sizeof(which) for which in [0, records)
    (0 body, 1 contact, 2 manifold, 3 joint, 4 spring, 5 params, 6 launch config, 7 stats,
    8 config, 9 external contact)."""
    names = ["body", "contact", "manifold", "joint", "spring", "params", "launch config", "stats",
             "config", "external contact"][:records]
    dts = [BODY, CONTACT, MANIFOLD, JOINT, SPRING, PARAMS, LAUNCH, STATS, CONFIG, EXTERNAL_CONTACT][:records]
    for i, (n, dt) in enumerate(zip(names, dts)):
        got = sizeof(i)
        if got != dt.itemsize:
            raise RuntimeError(f"layout mismatch for {n}: library {got} B, numpy {dt.itemsize} B")
