"""Every scene of the reference (scenes.h, 46 non-empty scenes: friction ramps, pyramid, ropes,
springs, stacks, round-shape stacks/ramps/pours/contact probes, Newton's cradle with capsule
strings and IgnoreCollision pairs, soft bodies, cloth, the mesh-lattice torus, bridge, breakable
joints), built by the reference's own scene functions (oracle/_ref), stepped on the GPU through
the C ABI and by the reference CPU solver in colour order: full state (bodies, manifolds and
contacts, joints) equal at every checkpoint."""
import numpy as np
import pytest

from conftest import assert_states_equal, body_block

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import oracle.bindings as ob
    if not ob.ref_available():
        pytest.skip("oracle/_ref (the compiled reference harness) is not built")
    ob.load_ref()
    from avbd import capi
    capi.load()
    return ob


# scene indices of scenes.h (0 is the empty scene); nothing is loaded at collection time, so
# torch initialises the GPU runtime before any library of this build is opened
@pytest.mark.parametrize("idx", range(1, 47))
def test_reference_scene_on_gpu_equals_reference(ref, idx):
    from avbd import capi
    from avbd.state import default_params
    name = ref.ref_scene_names()[idx]
    prm = default_params()
    cpu = ref.RefWorld(prm, ref_scene=idx)
    start = cpu.state()
    n = len(start.bodies)
    steps = 60 if n < 2000 else 12
    print(f"reference scene {name}: {n} bodies, {len(start.joints)} joints, {len(start.ignores)} ignore pairs, "
          f"{steps} steps", flush=True)
    gpu = capi.HipWorld.for_state(start, prm)
    done = 0
    for s in (1, steps // 4, steps):
        gpu.step(s - done)
        cpu.step(s - done, colored=True)
        done = s
        assert_states_equal(gpu.state(), cpu.state(), f"{name} step {s}")
    gpu.close()
