"""The uniform-grid broadphase (the reference Solver's default, solver.cpp:284-486) and the all-pairs
broadphase (solver.cpp:176-185) on the GPU, against the compiled reference's fixtures and the oracle:
pair emission (pairs in order, check and hit counts), colours and full state, bit for bit."""
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


@pytest.mark.parametrize("key", ["pile", "spread", "mixed", "allpairs"])
def test_grid_and_all_pairs_match_reference_fixtures(gpu, key):
    z = load_golden("grid.npz")
    from avbd import scenes
    from avbd.state import default_params
    sc = scene_from_golden(z, f"{key}.")
    prm = default_params(gravity=sc.gravity)
    w = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    w.set_broadphase(int(z[f"{key}.mode"]))
    assert w.broadphase == int(z[f"{key}.mode"])
    print(f"grid fixture {key}: {sc.num_bodies} bodies, mode {w.broadphase}", flush=True)
    emit_at = set(int(k) for k in z[f"{key}.emit_at"])
    traj = z[f"{key}.traj"]
    for k in range(int(z[f"{key}.steps"])):
        w.step(1)
        if k in emit_at:
            hits = w.hits()
            np.testing.assert_array_equal(hits[:, :2], z[f"{key}.pairs{k}"], err_msg=f"{key} step {k + 1}: pairs")
            st = w.stats()
            assert int(st["pair_checks"]) == int(z[f"{key}.checks{k}"]), f"{key} step {k + 1}: checks"
            assert int(st["sphere_hits"]) == int(z[f"{key}.hits{k}"]), f"{key} step {k + 1}: hits"
        if (k + 1) % 10 == 0:
            np.testing.assert_array_equal(body_block(w.state()), traj[(k + 1) // 10 - 1],
                                          err_msg=f"{key} step {k + 1}: bodies")
    w.close()


@pytest.mark.parametrize("mode", [1, 0])
def test_grid_trajectory_full_state_vs_oracle(gpu, oracle, mode):
    """Bodies, manifolds (features, anchors, penalties, lambdas), colours and the hit sequence
    against the oracle in the same broadphase mode, over a pile that falls and settles."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(2500 if mode == 1 else 600, seed=17)
    prm = default_params(gravity=sc.gravity)
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    g.set_broadphase(mode)
    c = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), prm, broadphase=mode)
    for k in range(1, 181):
        g.step(1)
        c.step(1)
        if k % 30 == 0:
            np.testing.assert_array_equal(g.hits(), c.hits(), err_msg=f"mode {mode} step {k}: hits")
            np.testing.assert_array_equal(g.colors(), c.colors(), err_msg=f"mode {mode} step {k}: colours")
            assert_states_equal(g.state(), c.state(), f"mode {mode} step {k}")
            assert int(g.stats()["pair_checks"]) == c.pair_checks
    g.close()


def test_grid_headline_pile_step_vs_oracle(gpu, oracle):
    """The 50,000-box pile in grid mode (about 2,500 dynamic boxes are larger than half a cell and take
    the AABB path): after 130 GPU steps, one step equals the oracle's grid step from that state."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(50000)
    prm = default_params(gravity=sc.gravity)
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    g.set_broadphase(1)
    g.step(130)
    pre = g.state()
    c = oracle.OracleWorld(pre, prm, broadphase=1)
    g.step(1)
    c.step(1)
    st = g.stats()
    print(f"grid pile50k step 131: {int(st['pair_checks'])} checks, {int(st['sphere_hits'])} hits, "
          f"{int(st['manifold_count'])} manifolds", flush=True)
    assert int(st["pair_checks"]) == c.pair_checks
    np.testing.assert_array_equal(g.hits(), c.hits(), err_msg="grid pair emission")
    np.testing.assert_array_equal(g.colors(), c.colors(), err_msg="colours")
    assert_states_equal(g.state(), c.state(), "grid pile50k step 131")
    g.close()


def test_grid_modes_validation(gpu):
    from avbd import scenes
    from avbd.state import default_params
    sts = [scenes.build_state(scenes.stack(seed=s), gpu.Maker()) for s in (1, 2)]
    prm = default_params()
    w = gpu.HipWorld.for_batch(sts, prm)
    with pytest.raises(gpu.AvbdError) as e:
        w.set_broadphase(1)
    assert e.value.code == -5
    with pytest.raises(gpu.AvbdError) as e:
        w.set_broadphase(7)
    assert e.value.code == -1
    w.upload(sts[0])  # one world again: every mode is allowed
    w.set_broadphase(1)
    from avbd.state import batch_states
    batch, offsets = batch_states(sts)
    with pytest.raises(gpu.AvbdError) as e:
        w.upload_batch(batch, offsets)
    assert e.value.code == -5
    w.set_broadphase(2)
    w.upload_batch(batch, offsets)
    w.step(3)
    w.close()
