"""ctypes binding of libavbd_hip.so (the C ABI in include/avbd_hip.h).

``HipWorld`` is a thin owner of one ``avbd_handle``: upload a State, step it
on the GPU, download it. There is no CPU fallback here: if the library or a
GPU is missing, construction raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

from .state import (BODY, CONFIG, EXTERNAL_CONTACT, JOINT, LAUNCH, MANIFOLD, SPRING, STATS, State, batch_states,
                    check_layouts)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libavbd_hip.so")

_p = C.c_void_p
_LIB = None
# Every open HipWorld, closed by an exit hook while the HIP runtime is still whole: a handle
# left to the garbage collector during interpreter shutdown would otherwise free device memory
# and destroy its stream after the runtime (or a profiler attached to it) has begun to tear down.
_LIVE = weakref.WeakSet()


def close_all():
    """Close every open HipWorld (installed as an exit hook by load())."""
    for w in list(_LIVE):
        w.close()

AVBD_OK = 0
ERRORS = {-1: "invalid argument", -2: "capacity", -3: "hip", -4: "no device", -5: "unsupported", -6: "injected"}


class AvbdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"avbd error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _ptr(a):
    return None if a is None or a.size == 0 else a.ctypes.data_as(_p)


def load(path: str | None = None):
    """Load the HIP library (raises OSError if it has not been built)."""
    global _LIB
    if _LIB is None:
        lib = C.CDLL(path or LIB_PATH)
        lib.avbd_last_error.restype = C.c_char_p
        lib.avbd_create.argtypes = [_p, C.POINTER(_p)]
        lib.avbd_destroy.argtypes = [_p]
        lib.avbd_set_params.argtypes = [_p, _p]
        lib.avbd_upload_state.argtypes = [_p, _p, C.c_int32, _p, C.c_int32, _p, C.c_int32, _p, C.c_int32, C.c_int64]
        lib.avbd_set_ignore_pairs.argtypes = [_p, _p, C.c_int32]
        lib.avbd_state_counts.argtypes = [_p, _p, _p, _p, _p, _p]
        lib.avbd_download_state.argtypes = [_p, _p, _p, _p, _p]
        for n in ("avbd_synchronize", "avbd_prepare", "avbd_iterate", "avbd_finish"):
            getattr(lib, n).argtypes = [_p]
        lib.avbd_step.argtypes = [_p, C.c_int32]
        lib.avbd_get_stats.argtypes = [_p, _p]
        lib.avbd_reset_stats.argtypes = [_p]
        lib.avbd_get_colors.argtypes = [_p, _p, C.c_int32]
        lib.avbd_get_hits.argtypes = [_p, _p, C.c_int32, _p]
        lib.avbd_debug_inject_fault.argtypes = [_p, C.c_int32]
        lib.avbd_debug_narrowphase_capacity.argtypes = [_p, C.c_int32]
        lib.avbd_default_params.argtypes = [_p]
        lib.avbd_make_box_body.argtypes = [_p, C.c_float, C.c_float, _p, _p, _p, _p, _p]
        lib.avbd_make_round_body.argtypes = [C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float, _p, _p, _p, _p, _p]
        lib.avbd_snapshot_avbs_size.restype = C.c_int64
        lib.avbd_snapshot_avbs_size.argtypes = [_p, C.c_int32]
        lib.avbd_snapshot_avbs.restype = C.c_int64
        lib.avbd_snapshot_avbs.argtypes = [_p, C.c_uint64, C.c_char_p, _p, C.c_int64]
        lib.avbd_snapshot_avbs_static.restype = C.c_int64
        lib.avbd_snapshot_avbs_static.argtypes = [_p, C.c_uint64, C.c_char_p, _p, C.c_int64]
        lib.avbd_make_joint.argtypes = [_p, _p, C.c_int32, C.c_int32, C.c_int64, _p, _p, C.c_float, C.c_float,
                                        C.c_float, _p]
        lib.avbd_make_spring.argtypes = [_p, _p, C.c_int32, C.c_int32, C.c_int64, _p, _p, C.c_float, C.c_float, _p]
        lib.avbd_sizeof.argtypes = [C.c_int]
        lib.avbd_batch_upload.argtypes = [_p, _p, C.c_int32, _p, C.c_int32, _p, C.c_int32, _p, C.c_int32, _p,
                                          C.c_int32, C.c_int64]
        lib.avbd_batch_num_worlds.argtypes = [_p, _p]
        lib.avbd_set_broadphase.argtypes = [_p, C.c_int32]
        lib.avbd_get_broadphase.argtypes = [_p, _p]
        lib.avbd_set_external_pairs.argtypes = [_p, _p, C.c_int32]
        lib.avbd_clear_external_pairs.argtypes = [_p]
        lib.avbd_set_external_contacts.argtypes = [_p, _p, C.c_int32]
        lib.avbd_set_launch_config.argtypes = [_p, _p]
        lib.avbd_get_launch_config.argtypes = [_p, _p]
        lib.avbd_autotune.argtypes = [_p, C.c_int32, C.c_int32, _p]
        lib.avbd_adjoint_forward.argtypes = [_p, C.c_int32, C.c_int32]
        lib.avbd_adjoint_backward.argtypes = [_p] + [_p] * 10
        lib.avbd_set_body_inputs.argtypes = [_p] * 5
        lib.avbd_get_body_state_f64.argtypes = [_p] * 5
        lib.avbd_adjoint_rewind.argtypes = [_p]
        check_layouts(lib.avbd_sizeof)
        _LIB = lib
        atexit.register(close_all)
    return _LIB


def _check(code: int):
    if code != AVBD_OK:
        raise AvbdError(code, load().avbd_last_error().decode())


def _f32(x, n):
    return np.ascontiguousarray(np.asarray(x, np.float32).reshape(n))


class Maker:
    """Record builders of the library (Rigid/Joint/Spring constructor fields)."""

    def make_box_body(self, size, density, friction, pos, rot, vel, angvel):
        lib = load()
        out = np.zeros((), BODY)
        _check(lib.avbd_make_box_body(_ptr(_f32(size, 3)), float(density), float(friction), _ptr(_f32(pos, 3)),
                                      _ptr(_f32(rot, 4)), _ptr(_f32(vel, 3)), _ptr(_f32(angvel, 3)),
                                      out.ctypes.data_as(_p)))
        return out

    def make_round_body(self, shape, radius, half_length, density, friction, pos, rot, vel, angvel):
        lib = load()
        out = np.zeros((), BODY)
        _check(lib.avbd_make_round_body(int(shape), float(radius), float(half_length), float(density), float(friction),
                                        _ptr(_f32(pos, 3)), _ptr(_f32(rot, 4)), _ptr(_f32(vel, 3)),
                                        _ptr(_f32(angvel, 3)), out.ctypes.data_as(_p)))
        return out

    def joint_spring_derived(self, rec_a, rec_b, ra, rb):
        lib = load()
        j = np.zeros((), JOINT)
        a = None if rec_a is None else np.ascontiguousarray(rec_a)
        b = np.ascontiguousarray(rec_b)
        _check(lib.avbd_make_joint(_ptr(a), b.ctypes.data_as(_p), 0, 0, 0, _ptr(_f32(ra, 3)), _ptr(_f32(rb, 3)),
                                   float("inf"), 0.0, float("inf"), j.ctypes.data_as(_p)))
        rest = -1.0
        if rec_a is not None:
            s = np.zeros((), SPRING)
            _check(lib.avbd_make_spring(a.ctypes.data_as(_p), b.ctypes.data_as(_p), 0, 0, 0, _ptr(_f32(ra, 3)),
                                        _ptr(_f32(rb, 3)), 1.0, -1.0, s.ctypes.data_as(_p)))
            rest = float(s["rest"])
        return float(j["torque_arm"]), rest


class HipWorld:
    """One resident GPU solver instance."""

    def __init__(self, params, max_bodies: int, max_joints: int = 0, max_springs: int = 0,
                 max_manifolds: int = 0, max_pairs: int = 0, device: int = 0, timing: bool = False):
        lib = load()
        self._lib = lib
        cfg = np.zeros((), CONFIG)
        cfg["params"] = params
        cfg["max_bodies"] = max_bodies
        cfg["max_joints"] = max_joints
        cfg["max_springs"] = max_springs
        cfg["max_manifolds"] = max_manifolds
        cfg["max_pairs"] = max_pairs
        cfg["device"] = device
        cfg["enable_timing"] = int(timing)
        h = _p()
        _check(lib.avbd_create(cfg.ctypes.data_as(_p), C.byref(h)))
        self._h = h
        _LIVE.add(self)

    @classmethod
    def for_state(cls, state: State, params, **kw) -> "HipWorld":
        w = cls(params, max_bodies=max(len(state.bodies), 1), max_joints=len(state.joints),
                max_springs=len(state.springs),
                max_manifolds=kw.pop("max_manifolds", 0) or max(12 * len(state.bodies) + 1024,
                                                                2 * len(state.manifolds)), **kw)
        w.upload(state)
        return w

    @classmethod
    def for_batch(cls, states, params, **kw) -> "HipWorld":
        """One handle stepping many independent worlds (avbd_batch_upload)."""
        batch, offsets = batch_states(states)
        nb = len(batch.bodies)
        w = cls(params, max_bodies=max(nb, 1), max_joints=len(batch.joints), max_springs=len(batch.springs),
                max_manifolds=kw.pop("max_manifolds", 0) or max(12 * nb + 1024, 2 * len(batch.manifolds)), **kw)
        w.upload_batch(batch, offsets)
        return w

    def upload_batch(self, state: State, offsets):
        offsets = np.ascontiguousarray(offsets, np.int32)
        self._keep = (state, offsets)
        _check(self._lib.avbd_batch_upload(self._h, _ptr(offsets), len(offsets) - 1, _ptr(state.bodies),
                                           len(state.bodies), _ptr(state.joints), len(state.joints),
                                           _ptr(state.springs), len(state.springs), _ptr(state.manifolds),
                                           len(state.manifolds), C.c_int64(state.next_serial)))
        self.set_ignore_pairs(state.ignores)
        self.world_offsets = offsets

    @property
    def num_worlds(self) -> int:
        n = C.c_int32()
        _check(self._lib.avbd_batch_num_worlds(self._h, C.byref(n)))
        return n.value

    BROADPHASE_ALL_PAIRS, BROADPHASE_GRID, BROADPHASE_SAP = 0, 1, 2

    def set_broadphase(self, mode: int):
        """Solver::broadphaseMode: BROADPHASE_SAP (default), _GRID (the reference Solver's default)
        or _ALL_PAIRS."""
        _check(self._lib.avbd_set_broadphase(self._h, int(mode)))

    def set_external_pairs(self, pairs):
        """Following steps take their candidate pairs from `pairs` ((n, 2) body indices) instead of
        the broadphase (Solver::stepCpuReferenceWithExternalBroadphase), until
        clear_external_pairs."""
        p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        _check(self._lib.avbd_set_external_pairs(self._h, p.ctypes.data_as(_p) if len(p) else None, len(p)))

    def clear_external_pairs(self):
        _check(self._lib.avbd_clear_external_pairs(self._h))

    def set_external_contacts(self, contacts):
        """The external contact table (EXTERNAL_CONTACT records) manifold initialisation reads
        (Solver::setExternalManifoldContacts); an empty table clears it."""
        c = np.ascontiguousarray(np.zeros(0, EXTERNAL_CONTACT) if contacts is None else contacts, EXTERNAL_CONTACT)
        _check(self._lib.avbd_set_external_contacts(self._h, c.ctypes.data_as(_p) if len(c) else None, len(c)))

    @property
    def broadphase(self) -> int:
        m = C.c_int32()
        _check(self._lib.avbd_get_broadphase(self._h, C.byref(m)))
        return m.value

    def launch_config(self) -> dict:
        out = np.zeros((), LAUNCH)
        _check(self._lib.avbd_get_launch_config(self._h, out.ctypes.data_as(_p)))
        return {k: int(out[k]) for k in LAUNCH.names}

    def set_launch_config(self, primal_lanes=16, primal_threads=256, dual_threads=256, use_graph=0, dataflow=0,
                          lazy_dual=1):
        c = np.zeros((), LAUNCH)
        c["primal_lanes"], c["primal_threads"] = primal_lanes, primal_threads
        c["dual_threads"], c["use_graph"], c["dataflow"] = dual_threads, int(use_graph), int(dataflow)
        c["lazy_dual"] = int(lazy_dual)
        _check(self._lib.avbd_set_launch_config(self._h, c.ctypes.data_as(_p)))

    def autotune(self, trial_steps: int = 3, allow_graph: bool = True) -> dict:
        """Time every launch shape on the resident state and keep the fastest (state unchanged)."""
        out = np.zeros((), LAUNCH)
        _check(self._lib.avbd_autotune(self._h, trial_steps, 0 if allow_graph else 1, out.ctypes.data_as(_p)))
        return {k: int(out[k]) for k in LAUNCH.names}

    def adjoint_forward(self, steps: int, checkpoint_every: int = 10):
        """Step `steps` times keeping checkpoints for avbd_adjoint_backward (every
        `checkpoint_every` steps; 0: tape every step now, so the backward only sweeps)."""
        _check(self._lib.avbd_adjoint_forward(self._h, steps, checkpoint_every))

    def adjoint_backward(self, d_pos=None, d_rot=None, d_vel=None, d_ang_vel=None) -> dict:
        """Gradients w.r.t. the recorded forward's initial pose, velocities, masses (moments
        scale with the mass) and friction coefficients, given dL/d(final state) (FP64 arrays,
        None = 0). Returns a dict of FP64 arrays."""
        nb = self.counts()[0]
        def inp(x, k):
            return None if x is None else np.ascontiguousarray(np.asarray(x, np.float64).reshape(nb, k))
        ins = [inp(d_pos, 3), inp(d_rot, 4), inp(d_vel, 3), inp(d_ang_vel, 3)]
        out = {"pos": np.zeros((nb, 3)), "rot": np.zeros((nb, 4)), "vel": np.zeros((nb, 3)),
               "ang_vel": np.zeros((nb, 3)), "mass": np.zeros(nb), "friction": np.zeros(nb)}
        ptr = lambda a: None if a is None else a.ctypes.data_as(_p)
        _check(self._lib.avbd_adjoint_backward(self._h, *[ptr(a) for a in ins],
                                               *[out[k].ctypes.data_as(_p) for k in
                                                 ("pos", "rot", "vel", "ang_vel", "mass", "friction")]))
        return out

    # Raw-address forms (host or device memory, e.g. a torch tensor's data_ptr(); 0 or None: the
    # field is kept / skipped): the C ABI copies with hipMemcpyDefault and is done with the memory
    # when the call returns, so device inputs must be complete (their stream synchronized) first.
    def set_body_inputs_ptr(self, vel=None, ang_vel=None, mass=None, friction=None):
        """avbd_set_body_inputs: FP64 velocities, angular velocities (nb x 3), masses and friction (nb)."""
        _check(self._lib.avbd_set_body_inputs(self._h, vel or None, ang_vel or None, mass or None,
                                              friction or None))

    def get_body_state_f64_ptr(self, pos=None, rot=None, vel=None, ang_vel=None):
        """avbd_get_body_state_f64: FP64 pos, vel, ang_vel (nb x 3) and rot (nb x 4)."""
        _check(self._lib.avbd_get_body_state_f64(self._h, pos or None, rot or None, vel or None, ang_vel or None))

    def adjoint_rewind(self):
        """avbd_adjoint_rewind: back to the start state of the recorded adjoint forward."""
        _check(self._lib.avbd_adjoint_rewind(self._h))

    def adjoint_backward_ptr(self, d_pos=None, d_rot=None, d_vel=None, d_ang_vel=None, g_pos=None, g_rot=None,
                             g_vel=None, g_ang_vel=None, g_mass=None, g_friction=None):
        """avbd_adjoint_backward with raw addresses (host or device memory)."""
        _check(self._lib.avbd_adjoint_backward(self._h, *[a or None for a in (
            d_pos, d_rot, d_vel, d_ang_vel, g_pos, g_rot, g_vel, g_ang_vel, g_mass, g_friction)]))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.avbd_destroy(self._h)
            self._h = None
        _LIVE.discard(self)

    def __del__(self, _finalizing=sys.is_finalizing):
        # during interpreter shutdown the exit hook has already closed every handle; a handle
        # still open here is left to the process exit rather than torn down out of order
        if _finalizing():
            return
        self.close()

    def set_params(self, params):
        p = np.ascontiguousarray(params)
        _check(self._lib.avbd_set_params(self._h, p.ctypes.data_as(_p)))

    def upload(self, state: State):
        self._keep = state
        _check(self._lib.avbd_upload_state(self._h, _ptr(state.bodies), len(state.bodies), _ptr(state.joints),
                                           len(state.joints), _ptr(state.springs), len(state.springs),
                                           _ptr(state.manifolds), len(state.manifolds),
                                           C.c_int64(state.next_serial)))
        self.set_ignore_pairs(state.ignores)

    def set_ignore_pairs(self, pairs):
        """IgnoreCollision pairs (list indices) of the uploaded state."""
        pairs = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        self._ignores = pairs
        _check(self._lib.avbd_set_ignore_pairs(self._h, _ptr(pairs), len(pairs)))

    def step(self, n: int = 1):
        _check(self._lib.avbd_step(self._h, n))

    def prepare(self):
        _check(self._lib.avbd_prepare(self._h))

    def iterate(self):
        _check(self._lib.avbd_iterate(self._h))

    def finish(self):
        _check(self._lib.avbd_finish(self._h))

    def synchronize(self):
        _check(self._lib.avbd_synchronize(self._h))

    def counts(self):
        v = [C.c_int32() for _ in range(4)]
        s = C.c_int64()
        _check(self._lib.avbd_state_counts(self._h, *[C.byref(x) for x in v], C.byref(s)))
        return [x.value for x in v] + [s.value]

    def state(self) -> State:
        nb, nj, ns, nm, serial = self.counts()
        b, j, s, m = np.zeros(nb, BODY), np.zeros(nj, JOINT), np.zeros(ns, SPRING), np.zeros(nm, MANIFOLD)
        _check(self._lib.avbd_download_state(self._h, _ptr(b), _ptr(j), _ptr(s), _ptr(m)))
        ig = getattr(self, "_ignores", None)
        return State(b, j, s, m, serial, ig if ig is not None and len(ig) else None)

    def stats(self) -> np.ndarray:
        out = np.zeros((), STATS)
        _check(self._lib.avbd_get_stats(self._h, out.ctypes.data_as(_p)))
        return out

    def reset_stats(self):
        _check(self._lib.avbd_reset_stats(self._h))

    def colors(self) -> np.ndarray:
        nb = self.counts()[0]
        out = np.zeros(nb, np.int32)
        _check(self._lib.avbd_get_colors(self._h, _ptr(out), nb))
        return out

    def hits(self) -> np.ndarray:
        n = C.c_int32()
        _check(self._lib.avbd_get_hits(self._h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 3), np.int32)
        _check(self._lib.avbd_get_hits(self._h, _ptr(out), n.value, C.byref(n)))
        return out

    def snapshot_avbs(self, frame: int = 0, scene: str = "") -> bytes:
        """AVBS v1 binary snapshot of the resident state (the viewer stream format), packed on the GPU."""
        name = scene.encode("utf-8")
        size = int(self._lib.avbd_snapshot_avbs_size(self._h, len(name)))
        if size < 0:
            _check(size)
        buf = np.zeros(size, np.uint8)
        n = int(self._lib.avbd_snapshot_avbs(self._h, int(frame), name, buf.ctypes.data_as(_p), size))
        if n < 0:
            _check(n)
        self.synchronize()
        return buf[:n].tobytes()

    def stream_avbs(self, frames: int, every: int = 1, scene: str = "", first_frame: int = 0):
        """Step the resident state and yield (frame, AVBS snapshot) every `every` steps."""
        frame = first_frame
        for _ in range(frames):
            self.step(every)
            frame += every
            yield frame, self.snapshot_avbs(frame, scene)

    def inject_fault(self, code: int = 1):
        _check(self._lib.avbd_debug_inject_fault(self._h, code))

    def debug_narrowphase_capacity(self, vertices):
        """Clip-polygon vertices the box-only narrowphase keeps in LDS (1..8); results do not change."""
        _check(self._lib.avbd_debug_narrowphase_capacity(self._h, int(vertices)))
