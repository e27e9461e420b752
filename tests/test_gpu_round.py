"""Spheres, capsules and cylinders on the GPU, through the C ABI.

Bar: exact FP32 equality (value equality, +0 == -0). Against the reference itself: every round
scene of tests/golden/round_scenes.npz (branch cases, stacks, ramps, a sphere pour, a mixed pile)
stepped on the GPU equals the compiled reference's trajectory. Against the oracle (pinned to the
reference by tests/test_oracle_round.py): larger pours and mixed piles, full state (bodies,
manifolds, contacts) at checkpoints; every launch shape; batched worlds.
"""
import numpy as np
import pytest

from conftest import assert_states_equal, body_block, load_golden, scene_from_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


def _round_names():
    return [str(x) for x in load_golden("round_scenes.npz")["names"]]


@pytest.mark.parametrize("name", _round_names())
def test_round_scene_equals_reference_trajectory(gpu, name):
    from avbd.scenes import build_state
    from avbd.state import default_params
    z = load_golden("round_scenes.npz")
    p = name + "."
    sc = scene_from_golden(z, p)
    every, steps = int(z[p + "every"]), int(z[p + "steps"])
    traj = z[p + "traj"]
    print(f"round scene {name}: {sc.num_bodies} bodies, {steps} steps", flush=True)
    w = gpu.HipWorld.for_state(build_state(sc, gpu.Maker()), default_params(gravity=sc.gravity))
    for k in range(1, steps + 1):
        w.step(1)
        if k % every == 0:
            np.testing.assert_array_equal(body_block(w.state()), traj[k // every - 1], err_msg=f"{name} step {k}")
    w.close()


@pytest.mark.parametrize("kind", ["pour", "mixed"])
def test_round_pile_full_state_equals_oracle(gpu, oracle, kind):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.sphere_pour(1500, width=12, depth=10, steps=150) if kind == "pour" else scenes.mixed_pile(1500, seed=9)
    prm = default_params(gravity=sc.gravity)
    print(f"round pile {kind}: {sc.num_bodies} bodies", flush=True)
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    c = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), prm)
    done = 0
    for s in (1, 5, 30, 80, 150):
        g.step(s - done)
        c.step(s - done)
        done = s
        assert_states_equal(g.state(), c.state(), f"{kind} step {s}")
        np.testing.assert_array_equal(g.colors(), c.colors(), err_msg=f"{kind} colours step {s}")
    assert len(c.state().manifolds) > 500
    g.close()


def test_round_every_launch_shape_identical(gpu):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.mixed_pile(800, seed=10)
    prm = default_params(gravity=sc.gravity)
    st = scenes.build_state(sc, gpu.Maker())
    base = None
    for lanes, threads, dual, graph, lazy in [(16, 256, 256, 0, 1), (8, 64, 64, 0, 1), (32, 128, 128, 1, 1),
                                              (16, 64, 256, 1, 1), (16, 256, 256, 0, 0), (8, 64, 64, 0, 0),
                                              (0, 64, 256, 0, 1), (0, 128, 128, 1, 1), (0, 256, 64, 0, 0)]:
        w = gpu.HipWorld.for_state(st, prm)
        w.set_launch_config(lanes, threads, dual, graph, 0, lazy)
        w.step(60)
        out = w.state()
        w.close()
        if base is None:
            base = out
        else:
            assert_states_equal(out, base, f"shape L{lanes} T{threads} D{dual} G{graph} lazy {lazy}")


def test_round_worlds_batched_equal_reference(gpu):
    from avbd import scenes
    from avbd.state import default_params
    z = load_golden("round_scenes.npz")
    names = ["sphere_stack10", "capsule_ramp6", "cylinder_stack10", "cylinders_caps", "sphere_in_moving_box"]
    scs = [scene_from_golden(z, n + ".") for n in names]
    states = [scenes.build_state(s, gpu.Maker()) for s in scs]
    prm = default_params(gravity=scs[0].gravity)
    w = gpu.HipWorld.for_batch(states, prm)
    steps = min(int(z[n + ".steps"]) for n in names)
    from avbd.state import split_state
    offsets = np.cumsum([0] + [len(s.bodies) for s in states])
    for k in range(1, steps + 1):
        w.step(1)
        if k % 15 == 0 or k == steps:
            parts = split_state(w.state(), offsets)
            for n, part in zip(names, parts):
                np.testing.assert_array_equal(body_block(part), z[n + ".traj"][k - 1], err_msg=f"{n} step {k}")
    w.close()
