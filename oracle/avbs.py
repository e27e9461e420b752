"""AVBS v1 binary snapshots, restated on the CPU (test infrastructure only: the checker of the
library's GPU-packed snapshots, never part of the product path).

encode_avbs follows makeSnapshotBinary (viewer_bridge.cpp:796-828) with the material table of
materialIdForBody (viewer_bridge.cpp:475-493): a numpy record array of the 64-byte body layout
in ascending dense id (SimWorld::bodies order, id = dense id). decode_avbs restates the viewer's
decodeBinarySnapshot (viewer/src/main.js:986-1050) as a sequential little-endian parse, so the
two are independent statements of the same layout.

Pinning: the layout by the decoder round trip and the fixed header bytes ("AVBS", version 1).
The material table (five entries) has one source, viewer_bridge.cpp:475-493, and the encoder
that uses it compiles only for Win32 with Winsock: parity unpinned for that table.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 0x53425641
HEADER = struct.Struct("<IIQII")
RECORD = np.dtype([("id", "<u4"), ("shape", "<u4"), ("material", "<u4"), ("dynamic", "<u4"),
                   ("pos", "<f4", 3), ("rot", "<f4", 4), ("size", "<f4", 3), ("radius", "<f4"),
                   ("half_length", "<f4")])
assert RECORD.itemsize == 64
# materialIdForBody for dynamic bodies, by RigidShapeType (box, sphere, capsule, cylinder)
_MATERIAL = np.array([1, 3, 5, 4], np.uint32)


def encode_avbs(bodies, frame: int, scene: str = "") -> bytes:
    """Snapshot bytes of body records (avbd_body layout, any order) the way the reference writes them."""
    order = np.argsort(bodies["dense_id"], kind="stable")
    b = bodies[order]
    rec = np.zeros(len(b), RECORD)
    dyn = b["mass"] > 0
    rec["id"] = b["dense_id"].astype(np.uint32)
    rec["shape"] = b["shape"].astype(np.uint32)
    rec["material"] = np.where(dyn, _MATERIAL[b["shape"]], 0)
    rec["dynamic"] = dyn.astype(np.uint32)
    rec["pos"], rec["rot"], rec["size"] = b["pos"], b["rot"], b["size"]
    rec["radius"], rec["half_length"] = b["shape_radius"], b["half_length"]
    name = scene.encode("utf-8")
    return HEADER.pack(MAGIC, 1, frame, len(name), len(b)) + name + rec.tobytes()


def decode_avbs(payload: bytes) -> dict:
    """The viewer's decoder: header fields, then each body field by field."""
    off = 0

    def take(fmt):
        nonlocal off
        v = struct.unpack_from("<" + fmt, payload, off)
        off += struct.calcsize("<" + fmt)
        return v

    magic, = take("I")
    if magic != MAGIC:
        raise ValueError("Invalid AVBD binary snapshot magic")
    version, = take("I")
    if version != 1:
        raise ValueError(f"Unsupported AVBD binary snapshot version {version}")
    frame, = take("Q")
    scene_bytes, = take("I")
    count, = take("I")
    scene = payload[off:off + scene_bytes].decode("utf-8")
    off += scene_bytes
    bodies = []
    for _ in range(count):
        ident, shape, material, dynamic = take("IIII")
        pos, rot, size = take("fff"), take("ffff"), take("fff")
        radius, half_length = take("ff")
        bodies.append(dict(id=ident, shape=shape, material=material, dynamic=dynamic != 0, position=pos,
                           rotation=rot, size=size, radius=radius, half_length=half_length))
    if off != len(payload):
        raise ValueError("trailing bytes after the last body")
    return dict(frame=frame, scene=scene, bodies=bodies)
