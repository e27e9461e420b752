"""The CPU oracle against golden fixtures written by the compiled reference.

Every comparison is exact FP32 equality: the oracle restates solver.cpp,
manifold.cpp, collide.cpp, joint.cpp, spring.cpp and maths.h operation by
operation, and the fixtures come from those very sources (tests/golden/make_golden.py).
"""
import numpy as np
import pytest

from conftest import (assert_states_equal, body_block, load_golden, scene_from_golden, state_from_golden)


def test_collide_boxes_matches_reference(oracle):
    """collide.cpp:898-955 on 4000 seeded box pairs: counts, feature keys, offsets, basis."""
    from avbd.state import BODY
    z = load_golden("collide_boxes.npz")
    n = len(z["n"])
    mismatches = 0
    for k in range(n):
        a, b = np.zeros((), BODY), np.zeros((), BODY)
        a["size"], a["pos"], a["rot"] = z["a_size"][k], z["a_pos"][k], z["a_rot"][k]
        b["size"], b["pos"], b["rot"] = z["b_size"][k], z["b_pos"][k], z["b_rot"][k]
        cnt, c, basis = oracle.oracle_collide(a, b)
        ok = cnt == z["n"][k]
        if ok and cnt:
            ok = (np.array_equal(c["feature"], z["feature"][k, :cnt]) and np.array_equal(c["r_a"], z["r_a"][k, :cnt])
                  and np.array_equal(c["r_b"], z["r_b"][k, :cnt]) and np.array_equal(basis, z["basis"][k]))
        mismatches += 0 if ok else 1
    assert mismatches == 0, f"{mismatches} of {n} box pairs differ from the reference"
    assert (z["n"] > 0).sum() > n // 2 and z["n"].max() >= 4


def test_stack_trajectory_bit_exact(oracle):
    """10-box stack, 300 steps: every step's bodies and colours equal the reference's."""
    z = load_golden("stack.npz")
    sc = scene_from_golden(z)
    from avbd.scenes import build_state
    from avbd.state import default_params
    w = oracle.OracleWorld(build_state(sc, oracle.OracleMaker()), default_params(gravity=sc.gravity))
    for k in range(300):
        w.step(1)
        st = w.state()
        np.testing.assert_array_equal(body_block(st), z["traj"][k], err_msg=f"step {k + 1}")
        np.testing.assert_array_equal(w.colors(), z["colors"][k], err_msg=f"colours step {k + 1}")
        assert len(st.manifolds) == z["manifold_count"][k]


def test_pyramid_chain_snapshots_bit_exact(oracle):
    """Pyramid + 1000-link chain with joints and springs, 600 steps."""
    z = load_golden("pyramid_chain.npz")
    sc = scene_from_golden(z)
    from avbd.scenes import build_state
    from avbd.state import default_params
    w = oracle.OracleWorld(build_state(sc, oracle.OracleMaker()), default_params(gravity=sc.gravity))
    steps = list(z["snap_steps"])
    done = 0
    for i, s in enumerate(steps):
        w.step(int(s) - done)
        done = int(s)
        np.testing.assert_array_equal(body_block(w.state()), z["snaps"][i], err_msg=f"step {s}")


def test_pyramid_chain_single_step_from_reference_state(oracle):
    """Identical inputs (the reference's state at step 100): one step reproduces its step 101."""
    z = load_golden("pyramid_chain.npz")
    from avbd.state import default_params
    pre, post = state_from_golden(z, "pre"), state_from_golden(z, "post")
    w = oracle.OracleWorld(pre, default_params(gravity=float(z["gravity"])))
    w.step(1)
    assert_states_equal(w.state(), post, "pyramid_chain step 101")
    np.testing.assert_array_equal(w.colors(), z["post_colors"])


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_pile_single_step_pairs_colors(oracle, idx):
    """1000-box pile: SAP pair emission, colours and the stepped state equal the reference's."""
    z = load_golden("pile.npz")
    from avbd.state import default_params
    pre, post = state_from_golden(z, f"pre{idx}"), state_from_golden(z, f"post{idx}")
    w = oracle.OracleWorld(pre, default_params(gravity=float(z["gravity"])))
    w.prepare()
    hits = w.hits()
    np.testing.assert_array_equal(hits[:, :2], z[f"pairs{idx}"], err_msg="SAP pair emission")
    assert w.pair_checks == int(z[f"checks{idx}"])
    assert len(hits) == int(z[f"hits{idx}"])
    w.iterate()
    w.finish()
    assert_states_equal(w.state(), post, f"pile step {int(z['pre_steps'][idx]) + 1}")
    np.testing.assert_array_equal(w.colors(), z[f"colors{idx}"])


def test_oracle_list_order_matches_plain_reference_step(oracle):
    """ORDER_LIST reproduces Solver::step itself (body-list Gauss-Seidel), when the reference is built."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(400, seed=9)
    prm = default_params(gravity=sc.gravity)
    ref = oracle.RefWorld(prm, scene=sc)
    w = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), prm, order=oracle.OracleWorld.ORDER_LIST)
    for k in range(90):
        ref.step(1, colored=False)
        w.step(1)
    np.testing.assert_array_equal(body_block(w.state()), body_block(ref.state()))


def test_record_builders_match_reference_constructors(oracle):
    """Rigid/Joint/Spring constructor-derived fields equal the reference's (rigid.cpp:24-42)."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    from avbd import scenes
    sc = scenes.pyramid_chain(base=6, links=40)
    a = scenes.build_state(sc, oracle.OracleMaker())
    b = scenes.build_state(sc, oracle.RefMaker())
    for f in a.bodies.dtype.names:
        np.testing.assert_array_equal(a.bodies[f], b.bodies[f], err_msg=f)
    np.testing.assert_array_equal(a.joints, b.joints)
    np.testing.assert_array_equal(a.springs, b.springs)


@pytest.mark.parametrize("key", ["pile", "spread", "mixed", "allpairs"])
def test_grid_and_all_pairs_broadphase_match_reference(oracle, key):
    """The uniform-grid broadphase (the reference Solver's default mode, solver.cpp:284-486) and the
    all-pairs broadphase (solver.cpp:176-185): the pair emission (pairs, checks, hits) before the
    recorded steps and the bodies every 10th step equal the reference's. `spread` is sparse enough
    that the grid's cell size is grown below 2^22 cells (solver.cpp:336-344)."""
    z = load_golden("grid.npz")
    from avbd import scenes
    from avbd.state import default_params
    sc = scene_from_golden(z, f"{key}.")
    prm = default_params(gravity=sc.gravity)
    w = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), prm, broadphase=int(z[f"{key}.mode"]))
    emit_at = set(int(k) for k in z[f"{key}.emit_at"])
    traj = z[f"{key}.traj"]
    for k in range(int(z[f"{key}.steps"])):
        w.step(1)
        if k in emit_at:
            hits = w.hits()
            np.testing.assert_array_equal(hits[:, :2], z[f"{key}.pairs{k}"], err_msg=f"{key} step {k + 1}: pairs")
            assert w.pair_checks == int(z[f"{key}.checks{k}"]), f"{key} step {k + 1}: checks"
            assert len(hits) == int(z[f"{key}.hits{k}"]), f"{key} step {k + 1}: hits"
        if (k + 1) % 10 == 0:
            np.testing.assert_array_equal(body_block(w.state()), traj[(k + 1) // 10 - 1],
                                          err_msg=f"{key} step {k + 1}: bodies")


@pytest.mark.parametrize("key", ["boxes", "mixed"])
def test_external_pairs_and_contacts_match_reference(oracle, key):
    """External pairs and an external contact table (Solver::stepCpuReferenceWithExternalBroadphase,
    solver.cpp:939-957; broadphaseExternalPairs, solver.cpp:579-626; the table's use in
    Manifold::initialize, manifold.cpp:110-121): shuffled and reversed pairs, duplicates, invalid and
    separated pairs; table entries with perturbed anchors, zero counts, preset penalties, lambdas and
    stick flags, made-up contacts for separated pairs, a repeated key, entries whose dense ids are
    descending (lookup only). The full state and the step's checks, hits and creations after every
    step equal the reference's, with the pairs and with the table alone."""
    z = load_golden("external.npz")
    from avbd.state import default_params
    st = state_from_golden(z, f"{key}.start")
    prm = default_params(gravity=float(z[f"{key}.gravity"]))
    P, Cx = z[f"{key}.pairs"], z[f"{key}.contacts"]
    w = oracle.OracleWorld(st, prm)
    pair_steps = int(z[f"{key}.pair_steps"])
    counts = z[f"{key}.counts"]
    for k in range(int(z[f"{key}.steps"])):
        w.set_external(P, Cx, pairs_on=k < pair_steps)
        w.step(1)
        hits = w.hits()
        assert (w.pair_checks, len(hits), int(len(hits) - hits[:, 2].sum()), int(hits[:, 2].sum())) == \
            tuple(int(x) for x in counts[k]), f"{key} step {k}: counts"
        assert_states_equal(w.state(), state_from_golden(z, f"{key}.s{k}"), f"{key} step {k}")
