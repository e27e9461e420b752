"""The drop-in: the reference's own Solver, with its physicsBackend replaced by the
"hip-avbd" PhysicsBackend adapter (avbd-demo3d_amd/host/hip_avbd_backend.cpp),
steps exactly like the reference CPU solver stepping in colour order.

Both solvers here are the compiled reference (oracle/_ref); only the backend
differs. The adapter packs the Solver object graph through the C ABI, runs a
full GPU step and writes bodies, joints and manifolds back into the objects.
"""
import numpy as np
import pytest

from conftest import assert_states_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import oracle.bindings as ob
    if not ob.ref_available():
        pytest.skip("oracle/_ref (reference harness with the adapter) is not built")
    ob.load_ref()
    return ob


@pytest.mark.parametrize("name", ["stack", "pile", "pyramid_chain", "mixed_round", "sphere_pour"])
def test_hip_backend_steps_like_reference(ref, name):
    from avbd import scenes
    from avbd.state import default_params
    sc = {"stack": scenes.stack(), "pile": scenes.pile(600, seed=12),
          "pyramid_chain": scenes.pyramid_chain(base=8, links=120),
          "mixed_round": scenes.mixed_pile(300, seed=11), "sphere_pour": scenes.sphere_pour(400)}[name]
    prm = default_params(gravity=sc.gravity)
    gpu = ref.RefWorld(prm, scene=sc)
    assert gpu.use_hip(0), "harness built without the hip-avbd adapter"
    cpu = ref.RefWorld(prm, scene=sc)
    done = 0
    for s in (1, 2, 10, 40, 90):
        gpu.step(s - done, colored=False)   # Solver::step -> hip-avbd backend
        cpu.step(s - done, colored=True)    # reference CPU solver, colour order
        done = s
        assert_states_equal(gpu.state(), cpu.state(), f"{name} step {s}")
    assert ref.load_ref().ref_hip_fallbacks() == 0


def test_hip_backend_all_pairs_falls_back(ref):
    """A Solver set to a broadphase the GPU step does not reproduce (here all pairs,
    solver.cpp:165-176) steps the CPU reference through the adapter, counted as fallbacks,
    so its trajectory is the CPU backend's own."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(200, seed=13)
    prm = default_params(gravity=sc.gravity)
    gpu = ref.RefWorld(prm, scene=sc)
    cpu = ref.RefWorld(prm, scene=sc)
    for w in (gpu, cpu):
        assert w.set_broadphase(ref.RefWorld.BROADPHASE_ALL_PAIRS) == 0
    assert gpu.use_hip(0)
    before = ref.load_ref().ref_hip_fallbacks()
    gpu.step(30, colored=False)
    cpu.step(30, colored=False)   # Solver::step with the CPU backend, list order
    assert_states_equal(gpu.state(), cpu.state(), "all-pairs solver through hip-avbd")
    assert ref.load_ref().ref_hip_fallbacks() - before == 30
