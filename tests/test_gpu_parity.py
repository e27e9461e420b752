"""GPU parity: the HIP step through the C ABI against the CPU oracle and the
reference's golden states.

Bar: exact FP32 equality (value equality, +0 == -0) of poses, velocities,
manifold contents, joint state, broadphase pair emission and colours. The FP64
oracle run states the FP32 error of the same step (tolerance 1e-5 relative,
scaled by max(1, |x|)).
"""
import numpy as np
import pytest

from conftest import (assert_states_equal, body_block, load_golden, scene_from_golden, state_from_golden)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


def _worlds(gpu, oracle, sc):
    from avbd import scenes
    from avbd.state import default_params
    prm = default_params(gravity=sc.gravity)
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    c = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), prm)
    return g, c


def test_stack_300_steps_bit_exact_vs_reference(gpu):
    """10-box stack (the CPU-runnable workload): all 300 steps equal the reference trajectory."""
    z = load_golden("stack.npz")
    sc = scene_from_golden(z)
    from avbd import scenes
    from avbd.state import default_params
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), default_params(gravity=sc.gravity))
    for k in range(300):
        g.step(1)
        if k % 10 == 9 or k < 5:
            np.testing.assert_array_equal(body_block(g.state()), z["traj"][k], err_msg=f"step {k + 1}")
            np.testing.assert_array_equal(g.colors(), z["colors"][k])


def test_pyramid_chain_600_steps_vs_reference_snapshots(gpu):
    """210-box pyramid + 1000-link chain, joints and springs: snapshots equal the reference's."""
    z = load_golden("pyramid_chain.npz")
    sc = scene_from_golden(z)
    from avbd import scenes
    from avbd.state import default_params
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), default_params(gravity=sc.gravity))
    done = 0
    for i, s in enumerate(z["snap_steps"]):
        g.step(int(s) - done)
        done = int(s)
        np.testing.assert_array_equal(body_block(g.state()), z["snaps"][i], err_msg=f"step {s}")


def test_pyramid_chain_single_step_from_reference_state(gpu):
    z = load_golden("pyramid_chain.npz")
    from avbd.state import default_params
    pre, post = state_from_golden(z, "pre"), state_from_golden(z, "post")
    g = gpu.HipWorld.for_state(pre, default_params(gravity=float(z["gravity"])))
    g.step(1)
    assert_states_equal(g.state(), post, "pyramid_chain step 101")
    np.testing.assert_array_equal(g.colors(), z["post_colors"])


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_pile_single_step_pairs_colors_vs_reference(gpu, idx):
    """Reference's own pre-state in, one GPU step: pairs, colours and state equal the reference's."""
    z = load_golden("pile.npz")
    from avbd.state import default_params
    pre, post = state_from_golden(z, f"pre{idx}"), state_from_golden(z, f"post{idx}")
    g = gpu.HipWorld.for_state(pre, default_params(gravity=float(z["gravity"])))
    g.step(1)
    hits = g.hits()
    np.testing.assert_array_equal(hits[:, :2], z[f"pairs{idx}"], err_msg="SAP pair emission")
    st = g.stats()
    assert int(st["pair_checks"]) == int(z[f"checks{idx}"])
    assert int(st["sphere_hits"]) == int(z[f"hits{idx}"])
    assert_states_equal(g.state(), post, f"pile step {int(z['pre_steps'][idx]) + 1}")
    np.testing.assert_array_equal(g.colors(), z[f"colors{idx}"])


def test_pile_trajectory_vs_oracle(gpu, oracle):
    """2,000-box pile from rest to contact-rich steps: every checked step equal to the oracle."""
    from avbd import scenes
    g, c = _worlds(gpu, oracle, scenes.pile(2000, seed=21))
    done = 0
    for s in (1, 5, 40, 120, 200):
        g.step(s - done)
        c.step(s - done)
        done = s
        assert_states_equal(g.state(), c.state(), f"step {s}")
        np.testing.assert_array_equal(g.colors(), c.colors())
        np.testing.assert_array_equal(g.hits(), c.hits())


def test_fp32_error_against_fp64(gpu, oracle):
    """The FP64 oracle states the FP32 GPU step's error: within 1e-5 relative (scale max(1,|x|))."""
    z = load_golden("pile.npz")
    from avbd.state import default_params
    pre = state_from_golden(z, "pre2")
    prm = default_params(gravity=float(z["gravity"]))
    g = gpu.HipWorld.for_state(pre, prm)
    g.step(1)
    w64 = oracle.OracleWorld(pre, prm, fp64=True)
    w64.step(1)
    gb = body_block(g.state()).astype(np.float64)
    ob = w64.bodies_f64()
    rel = np.abs(gb - ob) / np.maximum(1.0, np.abs(ob))
    pos_err = rel[:, :7].max()
    print(f"FP32 GPU vs FP64: max rel error pose {pos_err:.3e}, velocity {rel[:, 7:].max():.3e}")
    assert pos_err < 1e-5


def test_phased_api_equals_step(gpu):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(500, seed=8)
    prm = default_params(gravity=sc.gravity)
    st = scenes.build_state(sc, gpu.Maker())
    a = gpu.HipWorld.for_state(st, prm)
    b = gpu.HipWorld.for_state(st, prm)
    for _ in range(80):
        a.step(1)
        b.prepare()
        b.iterate()
        b.finish()
    assert_states_equal(a.state(), b.state(), "phased vs step")


def test_upload_download_roundtrip(gpu):
    z = load_golden("pile.npz")
    from avbd.state import default_params
    pre = state_from_golden(z, "pre1")
    g = gpu.HipWorld.for_state(pre, default_params())
    back = g.state()
    assert_states_equal(back, pre, "round trip")
    np.testing.assert_array_equal(back.manifolds["serial"], pre.manifolds["serial"])


def test_injected_fault_and_capacity_errors(gpu):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(300, seed=3)
    st = scenes.build_state(sc, gpu.Maker())
    g = gpu.HipWorld.for_state(st, default_params())
    g.inject_fault(1)
    with pytest.raises(gpu.AvbdError) as e:
        g.step(1)
    assert e.value.code == -6
    g.inject_fault(0)
    g.step(1)
    g.synchronize()
    # a pair buffer far too small for the pile: the kernels bound their writes and report it
    small = gpu.HipWorld(default_params(), max_bodies=len(st.bodies), max_pairs=8)
    small.upload(st)
    with pytest.raises(gpu.AvbdError) as e:
        small.step(2)
        small.synchronize()
    assert e.value.code == -2
    assert "max_pairs" in str(e.value)


def test_narrowphase_lds_capacity_redo_is_exact(gpu):
    """The box-only narrowphase keeps its clip polygons in LDS up to a vertex capacity; a manifold
    whose polygon outgrows it is listed and redone by the full-capacity narrowphase. With the
    capacity cut to 3 most face manifolds take that path: the state must stay bit-identical to
    the default capacity's."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(2000, seed=5)
    prm = default_params(gravity=sc.gravity)
    st = scenes.build_state(sc, gpu.Maker())
    a = gpu.HipWorld.for_state(st, prm)
    b = gpu.HipWorld.for_state(st, prm)
    b.debug_narrowphase_capacity(3)
    for _ in range(3):
        a.step(40)
        b.step(40)
        sa, sb = a.state(), b.state()
        assert_states_equal(sa, sb, "LDS capacity 3 vs 8")
    assert len(sa.manifolds) > 500
    with pytest.raises(gpu.AvbdError):
        b.debug_narrowphase_capacity(9)


def test_headline_size_single_step_vs_oracle(gpu, oracle):
    """50,000-box pile at its full size: after 120 GPU steps (falling, first contacts with the ground),
    one more step from that state equals the oracle's step from the same state, including the SAP pair
    emission and colours."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(50000)
    prm = default_params(gravity=sc.gravity)
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    g.step(120)
    pre = g.state()
    assert np.isfinite(body_block(pre)).all()
    c = oracle.OracleWorld(pre, prm)
    g.step(1)
    c.step(1)
    np.testing.assert_array_equal(g.hits(), c.hits())
    np.testing.assert_array_equal(g.colors(), c.colors())
    assert_states_equal(g.state(), c.state(), "pile50k step 121")
