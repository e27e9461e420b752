"""Batched worlds and launch shapes on the GPU.

* A batch of independent worlds stepped in one handle (avbd_batch_upload) equals
  every world stepped alone: against the CPU oracle per world (colour order) and
  against one GPU handle per world, exact FP32 (value equality).
* Every launch shape of the iteration kernels (lanes per body, block sizes, graph
  launch) gives the same state, and the autotuner leaves the state unchanged.
"""
import numpy as np
import pytest

from conftest import assert_states_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


def _batch_scenes(num, seed, hi):
    from avbd import scenes
    ws = scenes.batch(num, seed=seed, hi=hi)
    # edge worlds: a single static body, one box on the ground, a two-link chain
    ws.append(scenes.stacks(0, seed=5))
    ws.append(scenes.stacks(1, seed=6))
    ws.append(scenes.pyramid_chain(seed=7, base=0, links=2))
    return ws


def test_batch_equals_each_world_on_oracle(gpu, oracle):
    from avbd import scenes
    from avbd.state import default_params, split_state
    ws = _batch_scenes(20, seed=21, hi=160)
    prm = default_params()
    g = gpu.HipWorld.for_batch([scenes.build_state(w, gpu.Maker()) for w in ws], prm)
    assert g.num_worlds == len(ws)
    cpu = [oracle.OracleWorld(scenes.build_state(w, oracle.OracleMaker()), prm) for w in ws]
    done = 0
    for s in (1, 4, 15):
        g.step(s - done)
        for c in cpu:
            c.step(s - done)
        done = s
        for w, (a, c) in enumerate(zip(split_state(g.state(), g.world_offsets), cpu)):
            assert_states_equal(a, c.state(), f"world {w} ({ws[w].name}) step {s}")


def test_batch_equals_one_handle_per_world(gpu):
    from avbd import scenes
    from avbd.state import default_params, split_state
    ws = _batch_scenes(60, seed=22, hi=512)
    prm = default_params()
    states = [scenes.build_state(w, gpu.Maker()) for w in ws]
    g = gpu.HipWorld.for_batch(states, prm)
    g.set_launch_config(16, 256, 256, use_graph=1)
    g.step(12)
    got = split_state(g.state(), g.world_offsets)
    for w, st in enumerate(states):
        one = gpu.HipWorld.for_state(st, prm)
        one.step(12)
        assert_states_equal(got[w], one.state(), f"world {w} ({ws[w].name})")
        one.close()


def test_every_launch_shape_gives_the_same_state(gpu):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(1500, seed=31)
    prm = default_params(gravity=sc.gravity)
    st = scenes.build_state(sc, gpu.Maker())
    ref = gpu.HipWorld.for_state(st, prm)
    ref.step(40)
    base = ref.state()
    ref.close()
    shapes = [(l, t, d, gr, 0, 1) for l in (0, 8, 16, 32) for t in (64, 128, 256) for d, gr in ((64, 0), (256, 1))]
    shapes += [(l, t, 128, 0, 1, 1) for l, t in ((8, 128), (16, 64), (16, 256), (32, 256))]
    shapes += [(l, t, 256, 0, 0, 0) for l, t in ((8, 64), (16, 128), (32, 256), (0, 64))]
    for l, t, d, gr, df, lz in shapes:
        w = gpu.HipWorld.for_state(st, prm)
        w.set_launch_config(l, t, d, use_graph=gr, dataflow=df, lazy_dual=lz)
        assert w.launch_config() == {"primal_lanes": l, "primal_threads": t, "dual_threads": d, "use_graph": gr,
                                     "dataflow": df, "lazy_dual": lz}
        w.step(40)
        w.synchronize()
        assert_states_equal(w.state(), base, f"lanes {l} threads {t} dual {d} graph {gr} dataflow {df} lazy {lz}")
        w.close()


def test_lazy_dual_equals_eager_with_joints_and_springs(gpu):
    """The lazy-dual form (contact duals applied by their owner body at the next iteration's
    primal, joint duals per iteration, one contact dual at the end) equals the per-iteration dual
    launches bit for bit, on a pyramid with a chain of joints and springs and on the 8x8x8 soft
    body (joints, IgnoreCollision pairs, ground contacts); so do both with adaptive lanes per body
    (primal_lanes 0)."""
    from avbd import named_scenes, scenes
    from avbd.state import default_params
    cases = [(scenes.pyramid_chain(base=12, links=200), 60), (named_scenes.build("Soft Body 8x8x8"), 40)]
    for sc, steps in cases:
        prm = default_params(gravity=sc.gravity)
        st = scenes.build_state(sc, gpu.Maker())
        out = []
        for lanes, lz in ((16, 0), (16, 1), (0, 1), (0, 0)):
            w = gpu.HipWorld.for_state(st, prm)
            w.set_launch_config(lanes, 64, 256, 0, 0, lazy_dual=lz)
            w.step(steps)
            w.synchronize()
            out.append(w.state())
            w.close()
        assert_states_equal(out[1], out[0], f"{sc.name}: lazy vs eager dual")
        assert_states_equal(out[2], out[0], f"{sc.name}: adaptive lanes, lazy dual")
        assert_states_equal(out[3], out[0], f"{sc.name}: adaptive lanes, eager dual")


@pytest.mark.parametrize("lanes,threads", [(16, 64), (8, 128)])
def test_flow_primal_on_settled_headline_pile(gpu, lanes, threads):
    """The flow primal (one launch per step for every iteration and colour, each body waiting on
    the progress counters of its partners) equals the per-colour launches bit for bit on the
    settled 50,000-box pile (about 240k contacts, 7 colours)."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(50000)
    prm = default_params(gravity=sc.gravity)
    st = scenes.build_state(sc, gpu.Maker())
    a = gpu.HipWorld.for_state(st, prm)
    a.set_launch_config(16, 64, 128, use_graph=0, dataflow=0)
    a.step(600)
    mid = a.state()
    a.step(4)
    ref = a.state()
    stats = a.stats()
    a.close()
    assert int(stats["color_count"]) >= 5 and int(stats["contact_count"]) > 150000
    print(f"flow L{lanes} T{threads}: {int(stats['contact_count'])} contacts, {int(stats['color_count'])} colours",
          flush=True)
    b = gpu.HipWorld.for_state(mid, prm)
    b.set_launch_config(lanes, threads, 128, use_graph=0, dataflow=1)
    b.step(4)
    b.synchronize()
    assert_states_equal(b.state(), ref, f"flow L{lanes} T{threads}, 50k pile, steps 601-604")
    b.close()


def test_autotune_keeps_state_and_picks_a_valid_shape(gpu):
    from avbd import scenes
    from avbd.state import default_params, split_state
    ws = _batch_scenes(40, seed=23, hi=256)
    prm = default_params()
    states = [scenes.build_state(w, gpu.Maker()) for w in ws]
    plain = gpu.HipWorld.for_batch(states, prm)
    plain.step(5)
    tuned = gpu.HipWorld.for_batch(states, prm)
    tuned.step(2)
    cfg = tuned.autotune(trial_steps=2)
    assert cfg["primal_lanes"] in (0, 8, 16, 32) and cfg["primal_threads"] in (64, 128, 256)
    assert cfg["dual_threads"] in (64, 128, 256) and cfg["use_graph"] in (0, 1)
    assert tuned.launch_config() == cfg
    tuned.step(3)
    for w, (a, b) in enumerate(zip(split_state(tuned.state(), tuned.world_offsets),
                                   split_state(plain.state(), plain.world_offsets))):
        assert_states_equal(a, b, f"world {w}")


def test_batch_upload_rejects_bad_worlds(gpu):
    from avbd import scenes
    from avbd.state import batch_states, default_params
    prm = default_params()
    states = [scenes.build_state(scenes.pyramid_chain(seed=3, base=0, links=6), gpu.Maker()),
              scenes.build_state(scenes.stacks(4, seed=4), gpu.Maker())]
    batch, off = batch_states(states)
    w = gpu.HipWorld(prm, max_bodies=4000, max_joints=16, max_springs=16)
    bad = batch.copy()
    bad.joints["body_b"][-1] = off[1]          # a joint reaching into the next world
    with pytest.raises(gpu.AvbdError, match="spans two worlds"):
        w.upload_batch(bad, off)
    with pytest.raises(gpu.AvbdError, match="offsets"):
        w.upload_batch(batch, np.array([0, 3, len(batch.bodies) + 1], np.int32))
    with pytest.raises(gpu.AvbdError, match="at least one body"):
        w.upload_batch(batch, np.array([0, 0, len(batch.bodies)], np.int32))
    big = scenes.build_state(scenes.stacks(2100, seed=8), gpu.Maker())
    two, off2 = batch_states([big, states[1]])
    w2 = gpu.HipWorld(prm, max_bodies=len(two.bodies))
    with pytest.raises(gpu.AvbdError, match="AVBD_MAX_WORLD_BODIES"):
        w2.upload_batch(two, off2)
    with pytest.raises(gpu.AvbdError):
        w.set_launch_config(12, 256, 256, 0)
    w.upload_batch(batch, off)
    w.step(2)
    w.synchronize()


def _grid_world(n_side, pitch_x, pitch_y, offset=(0.0, 0.0), name="grid"):
    """Unit boxes on an n_side x n_side grid resting on a ground slab centred under it: with equal
    pitches the centres' x and y variances are equal, so the sweep axis is decided by the float
    rounding of the reference's serial sums."""
    from avbd.scenes import _Builder
    sb = _Builder()
    ox, oy = offset
    span = n_side * max(pitch_x, pitch_y) + 4.0
    sb.body((span, span, 1.0), 0.0, 0.5, (ox, oy, -0.5))
    for i in range(n_side):
        for j in range(n_side):
            sb.body((1.0, 1.0, 1.0), 1.0, 0.5,
                    (ox + (i - (n_side - 1) / 2.0) * pitch_x, oy + (j - (n_side - 1) / 2.0) * pitch_y, 0.5))
    return sb.scene(name, -9.81, 10)


@pytest.mark.parametrize("batched", [False, True], ids=["single", "batched"])
def test_sweep_axis_ties_and_far_worlds_take_the_serial_pass(gpu, oracle, batched):
    """Worlds whose two largest centre variances are equal (a square grid) or within the float
    rounding of the reference's serial mean for a world 10 km from the origin (pitches 0.02 %
    apart): the double-precision estimate cannot tell which axis the reference's float sums pick,
    so the exact serial pass runs (axis_fallbacks > 0) and every step equals the oracle."""
    from avbd import scenes
    from avbd.state import default_params, split_state
    prm = default_params(gravity=-9.81)
    ws = [_grid_world(20, 1.5, 1.5, name="tie"), _grid_world(20, 1.5, 1.5003, (1.0e4, -1.0e4), name="far")]
    cpu = [oracle.OracleWorld(scenes.build_state(w, oracle.OracleMaker()), prm) for w in ws]
    if batched:
        handles = [gpu.HipWorld.for_batch([scenes.build_state(w, gpu.Maker()) for w in ws], prm)]
    else:
        handles = [gpu.HipWorld.for_state(scenes.build_state(w, gpu.Maker()), prm) for w in ws]
    for s in range(4):
        for h in handles:
            h.step(1)
        for c in cpu:
            c.step(1)
        got = split_state(handles[0].state(), handles[0].world_offsets) if batched else [h.state() for h in handles]
        for w, (a, c) in enumerate(zip(got, cpu)):
            assert_states_equal(a, c.state(), f"{ws[w].name} step {s + 1}")
    for w, h in enumerate(handles):
        fb = int(h.stats()["axis_fallbacks"])
        print(f"{'batch' if batched else ws[w].name}: axis fallbacks {fb}, axis {int(h.stats()['sweep_axis'])}",
              flush=True)
        assert fb > 0
        h.close()


def test_batch_with_a_world_of_the_largest_size(gpu, oracle):
    """A world of 2048 bodies, the per-world broadphase's LDS capacity (bitonic sort of 2048 keys),
    batched with small worlds: equal to one handle per world, and the large world to the oracle."""
    from avbd import scenes
    from avbd.state import default_params, split_state
    big = scenes.pile(2043, seed=41)
    assert big.num_bodies == 2048
    ws = [big, scenes.stacks(2, seed=42), scenes.pile(300, seed=43)]
    prm = default_params(gravity=-9.81)
    states = [scenes.build_state(w, gpu.Maker()) for w in ws]
    g = gpu.HipWorld.for_batch(states, prm)
    cpu = oracle.OracleWorld(scenes.build_state(big, oracle.OracleMaker()), prm)
    g.step(150)
    cpu.step(150)
    got = split_state(g.state(), g.world_offsets)
    assert_states_equal(got[0], cpu.state(), "2048-body world, step 150")
    assert len(got[0].manifolds) > 500
    for w, st in enumerate(states):
        one = gpu.HipWorld.for_state(st, prm)
        one.step(150)
        assert_states_equal(got[w], one.state(), f"world {w} ({ws[w].name}) vs its own handle")
        one.close()
    g.close()
