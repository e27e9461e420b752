"""AVBS v1 snapshots packed on the GPU (avbd_snapshot_avbs) against the oracle's encoder of the
downloaded state: byte-for-byte equality (header, ordering by dense id, every field)."""
import numpy as np
import pytest

from oracle.avbs import decode_avbs, encode_avbs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


@pytest.mark.parametrize("kind", ["pile", "mixed_shuffled_ids"])
def test_snapshot_bytes_equal_oracle(gpu, kind):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(1500, seed=6) if kind == "pile" else scenes.mixed_pile(700, seed=4)
    st = scenes.build_state(sc, gpu.Maker())
    if kind != "pile":
        rng = np.random.default_rng(2)
        st.bodies["dense_id"] = rng.choice(5 * len(st.bodies), len(st.bodies), replace=False)
    print(f"avbs {kind}: {len(st.bodies)} bodies", flush=True)
    w = gpu.HipWorld.for_state(st, default_params(gravity=sc.gravity))
    for frame in (0, 1, 40):
        if frame:
            w.step(frame - (0 if frame == 1 else 1))
        got = w.snapshot_avbs(frame, f"{kind} scene")
        want = encode_avbs(w.state().bodies, frame, f"{kind} scene")
        assert got == want, f"{kind} frame {frame}"
    d = decode_avbs(got)
    assert len(d["bodies"]) == len(st.bodies)
    w.close()


def test_stream_and_capacity_error(gpu):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.sphere_pour(300, steps=60)
    w = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), default_params(gravity=sc.gravity))
    frames = []
    for frame, payload in w.stream_avbs(4, every=5, scene="pour"):
        assert payload == encode_avbs(w.state().bodies, frame, "pour")
        frames.append(frame)
    assert frames == [5, 10, 15, 20]
    small = np.zeros(10, np.uint8)
    rc = w._lib.avbd_snapshot_avbs(w._h, 0, b"pour", small.ctypes.data_as(gpu._p), 10)
    assert rc == -2   # AVBD_ERR_CAPACITY
    w.close()
