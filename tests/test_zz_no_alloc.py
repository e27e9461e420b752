"""avbd_snapshot_avbs_static (the allocation-free snapshot: a static device staging array, no
hipMalloc) against avbd_snapshot_avbs on the same handle: byte-equal output, on two calls in a
row on two handles (the second smaller, so the shared staging keeps the first call's larger
contents), plus its capacity and pinned-output errors."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATIC_MAX_BODIES = 65536


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


def _static_snapshot(gpu, w, frame, scene):
    import torch
    lib = gpu.load()
    name = scene.encode()
    size = int(lib.avbd_snapshot_avbs_size(w._h, len(name)))
    assert size > 0
    buf = torch.zeros(size, dtype=torch.uint8, pin_memory=True)
    n = int(lib.avbd_snapshot_avbs_static(w._h, frame, name, C.c_void_p(buf.data_ptr()), size))
    assert n == size, gpu.load().avbd_last_error().decode()
    w.synchronize()
    return bytes(buf.numpy().tobytes())


def test_static_snapshot_equals_allocating_snapshot(gpu):
    from avbd import scenes
    from avbd.state import default_params
    big = scenes.pile(3000, seed=21)
    small = scenes.pile(400, seed=22)
    worlds = []
    for sc, steps in ((big, 25), (small, 12)):
        st = scenes.build_state(sc, gpu.Maker())
        w = gpu.HipWorld.for_state(st, default_params(gravity=sc.gravity))
        w.step(steps)
        worlds.append(w)
    # first the larger world, then the smaller one: the second call overwrites only part of
    # the staging its predecessor filled
    for frame, w in ((7, worlds[0]), (8, worlds[1])):
        print(f"no-alloc snapshot: {w.counts()[0]} bodies, frame {frame}", flush=True)
        want = w.snapshot_avbs(frame, "no-alloc")
        got = _static_snapshot(gpu, w, frame, "no-alloc")
        assert got == want, f"frame {frame}: static-staging snapshot differs"
    # a step between two static calls on one handle: the records follow the state
    worlds[1].step(3)
    assert _static_snapshot(gpu, worlds[1], 11, "") == worlds[1].snapshot_avbs(11, "")
    for w in worlds:
        w.close()


def test_static_snapshot_errors(gpu):
    from avbd import scenes
    from avbd.state import default_params
    lib = gpu.load()
    sc = scenes.pile(200, seed=23)
    w = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), default_params(gravity=sc.gravity))
    size = int(lib.avbd_snapshot_avbs_size(w._h, 0))
    pageable = np.zeros(size, np.uint8)
    r = int(lib.avbd_snapshot_avbs_static(w._h, 0, b"", pageable.ctypes.data_as(C.c_void_p), size))
    assert r == -1 and b"pinned" in lib.avbd_last_error()
    w.close()
    big = scenes.pile(STATIC_MAX_BODIES + 1000, seed=24)
    st = scenes.build_state(big, gpu.Maker())
    assert len(st.bodies) > STATIC_MAX_BODIES
    w = gpu.HipWorld.for_state(st, default_params(gravity=big.gravity))
    import torch
    size = int(lib.avbd_snapshot_avbs_size(w._h, 0))
    buf = torch.zeros(size, dtype=torch.uint8, pin_memory=True)
    r = int(lib.avbd_snapshot_avbs_static(w._h, 0, b"", C.c_void_p(buf.data_ptr()), size))
    assert r == -2 and b"static staging" in lib.avbd_last_error()
    w.close()
