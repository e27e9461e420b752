"""Headless benchmark front end of the GPU step (headless_server.cpp:605-945).

Same command line as the reference's avbd_headless_server benchmark mode and the same output:
one JSON line of type "headlessBenchmark" whose keys, key order and number format follow
runBenchmark, so scripts that parse the reference's line (viewer/scripts/headless-benchmark.mjs
and its gates) parse this one. The physics backend is "hip-avbd": the resident GPU step of
libavbd_hip.so; there is no CPU solver behind it.

    python avbd-demo3d_amd/avbd_headless.py --benchmark-scene "Sphere Pour 5000 on Cylinders" \\
        --benchmark-frames 600 --warmup-frames 5 --no-stream --physics-backend hip-avbd

Per frame, physicsMs is the wall time of one step including the wait for the GPU (the reference
times Solver::step, which returns with the state updated). Without --no-stream each frame is also
packed as an AVBS viewer snapshot (the viewer stream's binary frame) and counted; --avbs-out FILE
appends those frames to a file. A WebSocket bridge is not part of this library.

Backend comparison (runBackendComparison, :852-943) needs a reference solver next to the GPU one;
run_backend_comparison takes it as an argument (any object with step(n) and state()), so the
comparison and its JSON line are the reference's with whichever CPU solver the caller supplies.

Measurements on the final state (measureFinalStateSignature :225-333, measureSphereGroundClearance
:434-495, compareWorldStates :335-382) use the reference's arithmetic: FP64 sums of FP32 fields in
body dense-id order for the signature and the comparison, FP32 for the ground clearance.
"""
from __future__ import annotations

import math
import sys
import time

import numpy as np

from . import named_scenes
from .scenes import build_state
from .state import SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CYLINDER, SHAPE_SPHERE, State, default_params

F = np.float32
BACKEND_NAMES = ("hip-avbd", "HIP AVBD")
BACKEND_DISPLAY = "HIP AVBD"


# ---- measurements on a downloaded state ----

def dense_order(state: State) -> np.ndarray:
    """Bodies in SimWorld order (dense id ascending); list order is newest first."""
    return state.bodies[np.argsort(state.bodies["dense_id"], kind="stable")]


def _finite(x) -> np.ndarray:
    x = np.asarray(x, np.float64)
    return np.where(np.isfinite(x), x, 0.0)


def final_state_signature(state: State) -> dict:
    """measureFinalStateSignature (headless_server.cpp:225-333)."""
    b = dense_order(state)
    n = len(b)
    out = {"activeBodies": n, "dynamicBodies": 0, "staticBodies": 0, "boxes": 0, "spheres": 0, "capsules": 0,
           "cylinders": 0, "positionChecksum": 0.0, "positionAbsChecksum": 0.0, "rotationChecksum": 0.0,
           "linearVelocityChecksum": 0.0, "angularVelocityChecksum": 0.0, "maxSpeed": 0.0, "maxAngularSpeed": 0.0,
           "aabbMin": [0.0] * 3, "aabbMax": [0.0] * 3, "dynamicAabbMin": [0.0] * 3, "dynamicAabbMax": [0.0] * 3}
    if n == 0:
        return out
    dyn = b["mass"] > F(0)
    shape = b["shape"]
    out["dynamicBodies"], out["staticBodies"] = int(dyn.sum()), int((~dyn).sum())
    for key, s in (("boxes", SHAPE_BOX), ("spheres", SHAPE_SPHERE), ("capsules", SHAPE_CAPSULE),
                   ("cylinders", SHAPE_CYLINDER)):
        out[key] = int((shape == s).sum())
    p, q, v, w = _finite(b["pos"]), _finite(b["rot"]), _finite(b["vel"]), _finite(b["ang_vel"])
    idx = np.arange(1, n + 1, dtype=np.float64)
    sw = 1.0 + shape.astype(np.float64) * 0.125
    terms = {
        "positionChecksum": idx * sw * (p[:, 0] * 0.73 + p[:, 1] * 1.31 + p[:, 2] * 1.91),
        "positionAbsChecksum": idx * sw * (np.abs(p[:, 0]) + np.abs(p[:, 1]) + np.abs(p[:, 2])),
        "rotationChecksum": idx * (q[:, 0] * 0.59 + q[:, 1] * 0.83 + q[:, 2] * 1.17 + q[:, 3] * 1.43),
        "linearVelocityChecksum": idx * (v[:, 0] * 0.67 + v[:, 1] * 1.07 + v[:, 2] * 1.61),
        "angularVelocityChecksum": idx * (w[:, 0] * 0.71 + w[:, 1] * 1.19 + w[:, 2] * 1.73),
    }
    for key, t in terms.items():
        acc = 0.0
        for x in t.tolist():          # serial sum in dense order, as the reference accumulates
            acc += x
        out[key] = acc
    out["maxSpeed"] = float(np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]).max())
    out["maxAngularSpeed"] = float(np.sqrt(w[:, 0] * w[:, 0] + w[:, 1] * w[:, 1] + w[:, 2] * w[:, 2]).max())
    r = _finite(b["radius"])[:, None]
    out["aabbMin"], out["aabbMax"] = (p - r).min(0).tolist(), (p + r).max(0).tolist()
    if dyn.any():
        out["dynamicAabbMin"] = (p[dyn] - r[dyn]).min(0).tolist()
        out["dynamicAabbMax"] = (p[dyn] + r[dyn]).max(0).tolist()
    return out


def _rotate32(q: np.ndarray, v) -> np.ndarray:
    """rotate(quat, float3) in FP32 (maths.h:441-446), one quaternion per row."""
    u = q[:, :3].astype(F)
    qw = q[:, 3:4].astype(F)
    v = np.broadcast_to(np.asarray(v, F), u.shape)

    def cross(a, c):
        return np.stack([a[:, 1] * c[:, 2] - a[:, 2] * c[:, 1], a[:, 2] * c[:, 0] - a[:, 0] * c[:, 2],
                         a[:, 0] * c[:, 1] - a[:, 1] * c[:, 0]], 1)
    t = cross(u, v) * F(2.0)
    return v + t * qw + cross(u, t)


def ground_contact_extent_z(b) -> np.ndarray:
    """benchmarkGroundContactExtentZ (headless_server.cpp:405-432), FP32 per body."""
    rot = b["rot"].astype(F)
    shape = b["shape"]
    size = b["size"].astype(F)
    srad, hl = b["shape_radius"].astype(F), b["half_length"].astype(F)
    zero = F(0)
    az = _rotate32(rot, (0, 0, 1))[:, 2]
    ax = _rotate32(rot, (1, 0, 0))[:, 2]
    ay = _rotate32(rot, (0, 1, 0))[:, 2]
    box = (np.abs(ax) * np.maximum(zero, size[:, 0] * F(0.5)) + np.abs(ay) * np.maximum(zero, size[:, 1] * F(0.5))
           + np.abs(az) * np.maximum(zero, size[:, 2] * F(0.5)))
    sph = np.where(srad > zero, srad, b["radius"].astype(F))
    cap = np.abs(az) * np.maximum(zero, hl) + np.maximum(zero, srad)
    cyl = np.abs(az) * np.maximum(zero, hl) + np.sqrt(np.maximum(zero, F(1.0) - az * az)) * np.maximum(zero, srad)
    out = b["radius"].astype(F).copy()
    for s, val in ((SHAPE_SPHERE, sph), (SHAPE_BOX, box), (SHAPE_CAPSULE, cap), (SHAPE_CYLINDER, cyl)):
        out = np.where(shape == s, val, out)
    return out.astype(F)


def ground_clearance(state: State) -> dict:
    """measureSphereGroundClearance (headless_server.cpp:434-495), FP32."""
    b = dense_order(state)
    out = {"hasGround": False, "groundTop": 0.0, "dynamicGroundBodies": 0, "belowGroundBodies": 0,
           "dynamicSpheres": 0, "belowGroundSpheres": 0, "minGroundBodyClearance": 0.0, "minClearance": 0.0,
           "maxGroundBodyPenetration": 0.0, "maxPenetration": 0.0}
    if len(b) == 0:
        return out
    dyn = b["mass"] > F(0)
    size = b["size"].astype(F)
    up = _rotate32(b["rot"].astype(F), (0, 0, 1))[:, 2]
    recv = (~dyn) & (b["shape"] == SHAPE_BOX) & ~(size[:, 0] < F(8)) & ~(size[:, 1] < F(8)) & ~(size[:, 2] > F(2)) \
        & (up > F(0.98))
    body = dyn & np.isin(b["shape"], (SHAPE_BOX, SHAPE_SPHERE, SHAPE_CAPSULE, SHAPE_CYLINDER))
    out["dynamicGroundBodies"] = int(body.sum())
    if not recv.any():
        return out
    top = (b["pos"][recv, 2].astype(F) + size[recv, 2] * F(0.5)).max()
    out["hasGround"], out["groundTop"] = True, float(top)
    if not body.any():
        return out
    eps = F(1.0e-5)
    clear = b["pos"][:, 2].astype(F) - ground_contact_extent_z(b) - top
    clear = np.where(np.abs(clear) < eps, F(0), clear).astype(F)
    below = clear < -eps
    cb = clear[body]
    out["minGroundBodyClearance"] = float(cb.min())
    out["belowGroundBodies"] = int((below & body).sum())
    out["maxGroundBodyPenetration"] = float((-clear[below & body]).max()) if (below & body).any() else 0.0
    sph = body & (b["shape"] == SHAPE_SPHERE)
    out["dynamicSpheres"] = int(sph.sum())
    if sph.any():
        out["minClearance"] = float(clear[sph].min())
        out["belowGroundSpheres"] = int((below & sph).sum())
        out["maxPenetration"] = float((-clear[below & sph]).max()) if (below & sph).any() else 0.0
    return out


def compare_worlds(reference: State, candidate: State) -> dict:
    """compareWorldStates (headless_server.cpp:335-382): per-body errors in dense-id order."""
    a, c = dense_order(reference), dense_order(candidate)
    out = {"referenceBodies": len(a), "candidateBodies": len(c), "comparedBodies": 0,
           "bodyCountMismatch": int(len(a) != len(c))}
    n = min(len(a), len(c))
    errs = {}
    for key, f in (("Position", "pos"), ("Rotation", "rot"), ("LinearVelocity", "vel"), ("AngularVelocity", "ang_vel")):
        x, y = _finite(a[f][:n]), _finite(c[f][:n])
        e = np.sqrt(((x - y) ** 2).sum(1))
        if f == "rot":   # q and -q are the same rotation
            e = np.minimum(e, np.sqrt(((x + y) ** 2).sum(1)))
        errs[key] = e
    out["comparedBodies"] = n
    for key, e in errs.items():
        out[f"max{key}Error"] = float(e.max()) if n else 0.0
        sq = 0.0
        for x in (e * e).tolist():
            sq += x
        out[f"rms{key}Error"] = math.sqrt(sq / n) if n else 0.0
    return out


# ---- JSON in the reference's format (std::fixed, 6 decimals) ----

def _num(x) -> str:
    if isinstance(x, (bool, np.bool_)):
        return "true" if x else "false"
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if not math.isfinite(x):
        return "0.000000" if math.isnan(x) else ("1e308" if x > 0 else "-1e308")
    return f"{x:.6f}"


def _json(fields) -> str:
    parts = []
    for k, v in fields:
        if isinstance(v, str):
            val = '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
        elif isinstance(v, (list, tuple)):
            val = "[" + ",".join(_num(x) for x in v) + "]"
        else:
            val = _num(v)
        parts.append(f'"{k}":{val}')
    return "{" + ",".join(parts) + "}"


class _Stat:
    """RunningStat (headless_server.cpp:93-121)."""

    def __init__(self):
        self.total, self.count, self.min, self.max = 0.0, 0, 0.0, 0.0

    def add(self, v):
        if self.count == 0:
            self.min = self.max = v
        else:
            self.min, self.max = min(self.min, v), max(self.max, v)
        self.total += v
        self.count += 1

    def avg(self):
        return self.total / self.count if self.count else 0.0


# ---- the benchmark ----

class Host:
    """SimulationHost for the GPU step: a named scene resident on the GPU, reloadable."""

    def __init__(self, scene_index: int, iterations: int = 0, gravity=None, timing: bool = False,
                 broadphase: str | None = None):
        from . import capi
        self._capi = capi
        self.scene_index = scene_index
        self.scene_name = named_scenes.NAMES[scene_index]
        self.iterations, self.gravity, self.timing = iterations, gravity, timing
        self.broadphase = broadphase_mode(broadphase)
        self.world = None
        self.load()

    def params(self):
        p = default_params()
        if self.iterations > 0:
            p["iterations"] = self.iterations
        if self.gravity is not None:
            p["gravity"] = F(self.gravity)
        return p

    def load(self):
        state = build_state(named_scenes.build(self.scene_index), self._capi.Maker())
        if self.world is not None:
            self.world.close()
        self.initial = state
        self.world = self._capi.HipWorld.for_state(state, self.params(), timing=self.timing)
        self.world.set_broadphase(self.broadphase)
        self.snapshot_frame = 0
        self.constraints = state.next_serial

    def live_forces(self, state_counts) -> int:
        nb, nj, ns, nm, _ = state_counts
        n_ignore = 0 if self.initial.ignores is None else len(self.initial.ignores)
        return nj + ns + nm + n_ignore

    def step_frame(self) -> float:
        t0 = time.perf_counter()
        self.world.step(1)
        self.world.synchronize()
        return (time.perf_counter() - t0) * 1e3

    def close(self):
        if self.world is not None:
            self.world.close()
            self.world = None


def run_benchmark(scene: str | None, frames: int = 600, warmup_frames: int = 0, reset_after_warmup: bool = False,
                  no_stream: bool = True, iterations: int = 0, gravity=None, avbs_out: str | None = None,
                  phase_timing: bool = False, host: Host | None = None, broadphase: str | None = None) -> dict:
    """runBenchmark (headless_server.cpp:605-852) on the GPU step. Returns (fields, JSON line).
    broadphase: a --broadphase name (None: the reference Solver's default, the uniform grid)."""
    frames = frames if frames > 0 else 600
    warmup_frames = max(warmup_frames, 0)
    idx = named_scenes.index_of(scene) if scene else 1
    own = host is None
    host = host or Host(idx, iterations, gravity, phase_timing, broadphase)
    for _ in range(warmup_frames):
        host.step_frame()
    if reset_after_warmup and warmup_frames > 0:
        host.load()

    stat = {k: _Stat() for k in ("frame", "physics", "serialize", "iterate", "broadphase", "narrowphase",
                                 "coloring", "finish")}
    last_bytes, sent = 0, 0
    sink = open(avbs_out, "ab") if avbs_out and not no_stream else None
    total0 = time.perf_counter()
    try:
        for _ in range(frames):
            f0 = time.perf_counter()
            phys = host.step_frame()
            ser = 0.0
            if not no_stream:
                s0 = time.perf_counter()
                blob = host.world.snapshot_avbs(host.snapshot_frame, host.scene_name)
                host.snapshot_frame += 1
                ser = (time.perf_counter() - s0) * 1e3
                last_bytes, sent = len(blob), sent + 1
                if sink:
                    sink.write(blob)
            f1 = time.perf_counter()
            stat["frame"].add((f1 - f0) * 1e3)
            stat["physics"].add(phys)
            stat["serialize"].add(ser)
            if phase_timing:
                st = host.world.stats()
                for k in ("iterate", "broadphase", "narrowphase", "coloring", "finish"):
                    stat[k].add(float(st[f"{k}_ms"]))
            host.constraints = max(host.constraints, host.live_forces(host.world.counts()))
    finally:
        if sink:
            sink.close()
    total_ms = (time.perf_counter() - total0) * 1e3

    fields = benchmark_fields(host.scene_name, frames, warmup_frames, reset_after_warmup, no_stream, host.params(),
                              host.world.state(), host.world.stats(), host.constraints, stat, total_ms, last_bytes,
                              sent)
    if own:
        host.close()
    return dict(fields), _json(fields)


def benchmark_fields(scene_name, frames, warmup_frames, reset_after_warmup, no_stream, prm, state: State, st,
                     constraints, stat, total_ms, last_bytes, sent):
    """The headlessBenchmark line's fields in the reference's order (headless_server.cpp:741-849)."""
    sig = final_state_signature(state)
    gc = ground_clearance(state)
    nb, nj, ns, nm = len(state.bodies), len(state.joints), len(state.springs), len(state.manifolds)
    n_ignore = 0 if state.ignores is None else len(state.ignores)
    fields = [
        ("type", "headlessBenchmark"), ("scene", scene_name), ("physicsBackend", BACKEND_DISPLAY),
        ("collisionOnly", False), ("frames", frames), ("warmupFrames", warmup_frames),
        ("resetAfterWarmup", bool(reset_after_warmup)), ("iterations", int(prm["iterations"])),
        ("gravity", float(prm["gravity"])), ("bodies", nb), ("constraints", constraints),
        ("finalActiveBodies", sig["activeBodies"]), ("finalDynamicBodies", sig["dynamicBodies"]),
        ("finalStaticBodies", sig["staticBodies"]), ("finalBoxes", sig["boxes"]), ("finalSpheres", sig["spheres"]),
        ("finalCapsules", sig["capsules"]), ("finalCylinders", sig["cylinders"]),
        ("finalPositionChecksum", sig["positionChecksum"]), ("finalPositionAbsChecksum", sig["positionAbsChecksum"]),
        ("finalRotationChecksum", sig["rotationChecksum"]),
        ("finalLinearVelocityChecksum", sig["linearVelocityChecksum"]),
        ("finalAngularVelocityChecksum", sig["angularVelocityChecksum"]),
        ("finalMaxSpeed", sig["maxSpeed"]), ("finalMaxAngularSpeed", sig["maxAngularSpeed"]),
        ("finalAabbMin", sig["aabbMin"]), ("finalAabbMax", sig["aabbMax"]),
        ("finalDynamicAabbMin", sig["dynamicAabbMin"]), ("finalDynamicAabbMax", sig["dynamicAabbMax"]),
        ("finalDynamicMaxHeight", sig["dynamicAabbMax"][2]),
        ("forceCount", nj + ns + nm + n_ignore), ("manifoldCount", int(st["manifold_count"])),
        ("manifoldsCreated", int(st["manifolds_created"])),
        # CPU-solver bookkeeping with no counterpart in the GPU step
        ("jointInitializationSkipped", 0), ("ignoreInitializationSkipped", 0),
        ("pairChecks", int(st["pair_checks"])), ("sphereHits", int(st["sphere_hits"])),
        ("constrainedHits", int(st["constrained_hits"])),
        ("primalJointVisits", 0), ("dualJointVisits", 0), ("primalJointSkipped", 0), ("dualJointSkipped", 0),
        ("primalIgnoreVisits", 0), ("dualIgnoreVisits", 0), ("primalIgnoreSkipped", 0), ("dualIgnoreSkipped", 0),
        ("finalGroundFound", gc["hasGround"]), ("finalGroundTop", gc["groundTop"]),
        ("finalDynamicGroundBodies", gc["dynamicGroundBodies"]), ("finalBelowGroundBodies", gc["belowGroundBodies"]),
        ("finalMinGroundBodyClearance", gc["minGroundBodyClearance"]),
        ("finalMaxGroundBodyPenetration", gc["maxGroundBodyPenetration"]),
        ("finalDynamicSpheres", gc["dynamicSpheres"]), ("finalBelowGroundSpheres", gc["belowGroundSpheres"]),
        ("finalMinSphereGroundClearance", gc["minClearance"]),
        ("finalMaxSphereGroundPenetration", gc["maxPenetration"]),
        ("streaming", not no_stream), ("totalMs", total_ms),
        ("frameAvgMs", stat["frame"].avg()), ("frameMinMs", stat["frame"].min), ("frameMaxMs", stat["frame"].max),
        ("physicsAvgMs", stat["physics"].avg()), ("physicsMinMs", stat["physics"].min),
        ("physicsMaxMs", stat["physics"].max),
        ("simWorldSyncAvgMs", 0.0), ("broadphaseAvgMs", stat["broadphase"].avg()),
        ("forceInitAvgMs", stat["narrowphase"].avg()), ("forceInitGatherAvgMs", 0.0),
        ("forceInitParallelAvgMs", 0.0), ("forceInitCleanupAvgMs", 0.0), ("manifoldAllocAvgMs", 0.0),
        ("bodyInitAvgMs", stat["coloring"].avg()), ("primalAvgMs", 0.0), ("dualAvgMs", 0.0),
        ("velocityUpdateAvgMs", stat["finish"].avg()),
        ("webgpuStatus", "not used: physics runs on hip-avbd"), ("webgpuRuntimeFallbacks", int(st["cpu_fallbacks"])),
        ("avbdGpuActive", True), ("avbdGpuIterateMs", float(st["iterate_ms"])),
        ("avbdGpuIterateAvgMs", stat["iterate"].avg()), ("avbdGpuBuildAvgMs", 0.0),
        ("avbdGpuBuildBodiesAvgMs", 0.0), ("avbdGpuBuildRecordsAvgMs", 0.0), ("avbdGpuBuildSlotsAvgMs", 0.0),
        ("avbdGpuBuildAdjAvgMs", 0.0), ("avbdGpuBuildColorAvgMs", stat["coloring"].avg()),
        ("avbdGpuSubmitAvgMs", 0.0), ("avbdGpuWaitAvgMs", 0.0), ("avbdGpuApplyAvgMs", 0.0),
        ("avbdGpuBodies", nb), ("avbdGpuContacts", int(st["contact_count"])), ("avbdGpuSpherePairs", 0),
        ("avbdGpuJoints", int(st["joint_count"])), ("avbdGpuSprings", int(st["spring_count"])),
        ("avbdGpuColors", int(st["color_count"])),
        ("snapshotSerializeAvgMs", stat["serialize"].avg()), ("webSocketSendAvgMs", 0.0),
        ("snapshotMode", "binary"), ("snapshotBytes", last_bytes), ("snapshotsSent", sent),
    ]
    return fields


def run_backend_comparison(scene: str | None, frames: int, reference, iterations: int = 0,
                           reference_name: str = "CPU Reference", host: Host | None = None,
                           broadphase: str | None = None):
    """runBackendComparison (headless_server.cpp:852-943): the GPU step against `reference`, a solver
    on the same scene with step(n) and state(), both in the broadphase mode `broadphase` (None:
    the reference Solver's default, the uniform grid). Returns (fields, JSON line)."""
    frames = frames if frames > 0 else 120
    idx = named_scenes.index_of(scene) if scene else 1
    own = host is None
    host = host or Host(idx, iterations, broadphase=broadphase)
    rstat, cstat = _Stat(), _Stat()
    for _ in range(frames):
        t0 = time.perf_counter()
        reference.step(1)
        rstat.add((time.perf_counter() - t0) * 1e3)
        cstat.add(host.step_frame())
    fields = comparison_fields(host.scene_name, frames, int(host.params()["iterations"]), reference_name,
                               rstat.avg(), cstat.avg(), int(host.world.stats()["cpu_fallbacks"]),
                               reference.state(), host.world.state())
    if own:
        host.close()
    return dict(fields), _json(fields)


def comparison_fields(scene_name, frames, iterations, reference_name, reference_ms, candidate_ms, fallbacks,
                      rs: State, cs: State):
    """The backendComparison line's fields in the reference's order (headless_server.cpp:901-940)."""
    cmp = compare_worlds(rs, cs)
    rsig, csig = final_state_signature(rs), final_state_signature(cs)
    rg, cg = ground_clearance(rs), ground_clearance(cs)
    fields = [
        ("type", "backendComparison"), ("scene", scene_name), ("frames", frames),
        ("iterations", iterations), ("referenceBackend", reference_name),
        ("candidateBackend", BACKEND_DISPLAY), ("referencePhysicsAvgMs", reference_ms),
        ("candidatePhysicsAvgMs", candidate_ms),
        ("candidateWebgpuRuntimeFallbacks", fallbacks),
        ("candidateWebgpuStatus", "not used: physics runs on hip-avbd"),
        ("referenceBodies", cmp["referenceBodies"]), ("candidateBodies", cmp["candidateBodies"]),
        ("comparedBodies", cmp["comparedBodies"]), ("bodyCountMismatch", cmp["bodyCountMismatch"]),
    ]
    for key in ("Position", "Rotation", "LinearVelocity", "AngularVelocity"):
        fields += [(f"max{key}Error", cmp[f"max{key}Error"]), (f"rms{key}Error", cmp[f"rms{key}Error"])]
    for key in ("position", "rotation", "linearVelocity", "angularVelocity"):
        fields += [(f"reference{key[0].upper() + key[1:]}Checksum", rsig[f"{key}Checksum"]),
                   (f"candidate{key[0].upper() + key[1:]}Checksum", csig[f"{key}Checksum"])]
    fields += [("referenceBelowGroundSpheres", rg["belowGroundSpheres"]),
               ("candidateBelowGroundSpheres", cg["belowGroundSpheres"]),
               ("referenceMaxGroundPenetration", rg["maxPenetration"]),
               ("candidateMaxGroundPenetration", cg["maxPenetration"]),
               ("referenceBelowGroundBodies", rg["belowGroundBodies"]),
               ("candidateBelowGroundBodies", cg["belowGroundBodies"]),
               ("referenceMaxGroundBodyPenetration", rg["maxGroundBodyPenetration"]),
               ("candidateMaxGroundBodyPenetration", cg["maxGroundBodyPenetration"])]
    return fields


# ---- command line (headless_server.cpp:946-985) ----

# applyBroadphaseMode (headless_server.cpp:579-591): names of Solver::broadphaseMode; without
# --broadphase the Solver keeps its default, the uniform grid (solver.cpp:676-680).
BROADPHASE_NAMES = {"allpairs": 0, "grid": 1, "uniformgrid": 1, "spatialhash": 1, "sap": 2, "sweepandprune": 2}
DEFAULT_BROADPHASE = 1


def broadphase_mode(name: str | None) -> int:
    """The mode for a --broadphase name; None or an unknown name (reported on stderr, as the
    reference does) keeps the default."""
    if name is None:
        return DEFAULT_BROADPHASE
    if name not in BROADPHASE_NAMES:
        print(f"Unknown broadphase '{name}' (expected allpairs, grid, or sap)", file=sys.stderr)
        return DEFAULT_BROADPHASE
    return BROADPHASE_NAMES[name]


def _arg(argv, name):
    for i, a in enumerate(argv[:-1]):
        if a == name:
            return argv[i + 1]
    return None


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    scene = _arg(argv, "--scene")
    bench_scene = _arg(argv, "--benchmark-scene")
    compare_scene = _arg(argv, "--compare-scene")
    backend = _arg(argv, "--physics-backend")
    to_int = lambda s: int(float(s)) if s else 0
    frames = to_int(_arg(argv, "--benchmark-frames"))
    warmup = to_int(_arg(argv, "--warmup-frames"))
    compare_frames = to_int(_arg(argv, "--compare-frames"))
    iterations = to_int(_arg(argv, "--iterations"))
    g = _arg(argv, "--gravity")
    broadphase = _arg(argv, "--broadphase")
    if backend is not None and backend not in BACKEND_NAMES:
        print(f"Unknown physics backend '{backend}' (this library runs hip-avbd only)", file=sys.stderr)
        return 2
    if "--collision-only" in argv:
        print("--collision-only is not available on hip-avbd", file=sys.stderr)
        return 2
    if bench_scene or frames > 0:
        name = bench_scene or scene
        if name is not None:
            try:
                named_scenes.index_of(name)
            except KeyError:
                print(f"Unknown benchmark scene '{name}'", file=sys.stderr)
                return 2
        _, line = run_benchmark(name, frames, warmup, "--reset-after-warmup" in argv, "--no-stream" in argv,
                                iterations, float(np.float32(float(g))) if g is not None else None,
                                _arg(argv, "--avbs-out"), "--phase-timing" in argv, broadphase=broadphase)
        print(line, flush=True)
        return 0
    if compare_scene or compare_frames > 0:
        print("Backend comparison needs the CPU reference solver, which is not part of this library; "
              "call avbd.headless.run_backend_comparison with a reference solver", file=sys.stderr)
        return 2
    print("usage: avbd_headless.py --benchmark-scene NAME [--benchmark-frames N] [--warmup-frames W] "
          "[--reset-after-warmup] [--no-stream] [--iterations I] [--gravity G] [--broadphase allpairs|grid|sap] "
          "[--avbs-out FILE] [--phase-timing]", file=sys.stderr)
    return 2


if __name__ == "__main__":
    sys.exit(main())
