"""Writes headless_gates.json: the reference solver's final state after each benchmark gate's
frames (viewer/package.json:21-24: 5 warm-up frames, then 600 or 240), for the scenes those gates
run. Each entry holds a SHA-256 of the final body state (position, rotation, linear and angular
velocity as FP32, dense-id order) and the gate metrics of headless_server.cpp:225-495.

The reference solver is oracle/_ref (the reference's own sources, compiled), stepping in colour
order, the order the GPU step reproduces bit for bit."""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "avbd-demo3d_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle.bindings as ob  # noqa: E402
from avbd import headless, named_scenes  # noqa: E402
from avbd.state import default_params  # noqa: E402

GATES = [("Sphere Pour 5000 on Cylinders", 5, 600), ("Sphere Pour 20000 on Cylinders", 5, 600),
         ("Soft Body 8x8x8", 5, 240)]


def state_digest(state) -> str:
    b = headless.dense_order(state)
    h = hashlib.sha256()
    for f in ("pos", "rot", "vel", "ang_vel"):
        h.update(np.ascontiguousarray(b[f], np.float32).tobytes())
    return h.hexdigest()


def main():
    ob.load_ref()
    out = {}
    for name, warm, frames in GATES:
        t0 = time.time()
        w = ob.RefWorld(default_params(), ref_scene=named_scenes.index_of(name))
        w.step(warm + frames, colored=True)
        st = w.state()
        sig, gc = headless.final_state_signature(st), headless.ground_clearance(st)
        out[name] = dict(warmup=warm, frames=frames, digest=state_digest(st),
                         finalPositionChecksum=sig["positionChecksum"], finalMaxSpeed=sig["maxSpeed"],
                         finalDynamicMaxHeight=sig["dynamicAabbMax"][2],
                         finalMaxSphereGroundPenetration=gc["maxPenetration"],
                         finalMaxGroundBodyPenetration=gc["maxGroundBodyPenetration"])
        print(name, round(time.time() - t0, 1), out[name], flush=True)
    with open(os.path.join(HERE, "headless_gates.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
