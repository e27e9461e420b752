"""Spheres, capsules and cylinders in the oracle, against the compiled reference's fixtures
(tests/golden/make_golden.py: collide_round.npz, round_scenes.npz).

Exact FP32 equality throughout: Manifold::collide on 6000 seeded pairs over every shape pair
(boxes, thin static slabs, spheres, capsules, cylinders, static bodies, fast relative motion),
and every step of the round scenes (branch cases, stacks, ramps, a sphere pour, a mixed pile)
stepped in colour order. Those trajectories also pin the record builders (mass and moments feed
the dynamics) and the rolling-friction pass. Coverage: the oracle counts the branches of the
round-shape narrowphase and of the rolling pass, and the branch cases reach every one.
"""
import math

import numpy as np
import pytest

from conftest import body_block, load_golden, scene_from_golden


def test_collide_round_pairs_match_reference(oracle):
    z = load_golden("collide_round.npz")
    A, B, dt = z["a"], z["b"], float(z["dt"])
    n = len(z["n"])
    bad = []
    for k in range(n):
        cnt, c, basis = oracle.oracle_collide_pair(A[k], B[k], dt)
        ok = cnt == z["n"][k]
        if ok and cnt:
            ok = (np.array_equal(c["feature"], z["feature"][k, :cnt]) and np.array_equal(c["r_a"], z["r_a"][k, :cnt])
                  and np.array_equal(c["r_b"], z["r_b"][k, :cnt]) and np.array_equal(basis, z["basis"][k]))
        if not ok:
            bad.append(k)
    assert not bad, f"{len(bad)} of {n} pairs differ from the reference (first {bad[:5]})"
    # the pairs cover every shape pair and touch often
    shapes = set(zip(A["shape"].tolist(), B["shape"].tolist()))
    assert len(shapes) == 16
    assert (z["n"] > 0).sum() > n // 2


def _round_names():
    return [str(x) for x in load_golden("round_scenes.npz")["names"]]


@pytest.mark.parametrize("name", _round_names())
def test_round_scene_trajectory_bit_exact(oracle, name):
    from avbd.scenes import build_state
    from avbd.state import default_params
    z = load_golden("round_scenes.npz")
    p = name + "."
    sc = scene_from_golden(z, p)
    every, steps = int(z[p + "every"]), int(z[p + "steps"])
    traj = z[p + "traj"]
    w = oracle.OracleWorld(build_state(sc, oracle.OracleMaker()), default_params(gravity=sc.gravity))
    for k in range(1, steps + 1):
        w.step(1)
        if k % every == 0:
            np.testing.assert_array_equal(body_block(w.state()), traj[k // every - 1], err_msg=f"{name} step {k}")


def test_every_round_branch_is_exercised(oracle):
    from avbd import scenes
    from avbd.state import default_params
    oracle.round_branch_hits(reset=True)
    for sc in scenes.round_cases(90):
        w = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), default_params(gravity=sc.gravity))
        w.step(sc.steps)
    hits = oracle.round_branch_hits()
    missing = [k for k, v in hits.items() if v == 0]
    assert not missing, f"branches never reached: {missing}"


@pytest.mark.parametrize("shape", [1, 2, 3])
def test_round_records_closed_forms(oracle, shape):
    """Mass and moments of the round records against the solids' formulas (float64, 1e-6)."""
    r, hl, rho = 0.37, 0.61, 1.3
    rec = oracle.OracleMaker().make_round_body(shape, r, hl, rho, 0.4, (1, 2, 3), (0, 0, 0, 1), (0, 0, 0),
                                               (0, 0, 0))
    ball = 4.0 / 3.0 * math.pi * r ** 3 * rho
    if shape == 1:
        mass, axial, across, bound = ball, 0.4 * ball * r * r, 0.4 * ball * r * r, r
    else:
        cyl = math.pi * r * r * 2 * hl * rho
        if shape == 2:
            mass = cyl + ball
            axial = 0.5 * cyl * r * r + 0.4 * ball * r * r
            across = (3 * r * r + 4 * hl * hl) / 12 * cyl + 0.4 * ball * r * r + ball * hl * hl
            bound = hl + r
        else:
            mass, axial, across, bound = cyl, 0.5 * cyl * r * r, (3 * r * r + 4 * hl * hl) / 12 * cyl, math.hypot(r, hl)
    assert rec["shape"] == shape and rec["shape_radius"] == np.float32(r)
    assert rec["mass"] == pytest.approx(mass, rel=1e-6)
    assert rec["moment"][2] == pytest.approx(axial, rel=1e-6)
    assert rec["moment"][0] == pytest.approx(across, rel=1e-6) and rec["moment"][1] == rec["moment"][0]
    assert rec["radius"] == pytest.approx(bound, rel=1e-6)


def test_round_manifold_rules(oracle):
    """manifold.cpp:100-174 on round pairs: a sphere makes the pair frictionless; with any round
    shape, matched contacts restart their friction rows (lambda_t = 0, penalty_t = 1, no stick)."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.round_cases(60)
    by = {s.name: s for s in sc}
    for name, frictionless in (("sphere_on_box", True), ("capsule_flat", False), ("cylinder_standing", False)):
        s = by[name]
        w = oracle.OracleWorld(scenes.build_state(s, oracle.OracleMaker()), default_params(gravity=s.gravity))
        w.step(30)
        w.prepare()
        st = w.state()
        shape = st.bodies["shape"]
        m = [rec for rec in st.manifolds if shape[rec["body_a"]] != 0 or shape[rec["body_b"]] != 0]
        assert len(m) > 0
        for rec in m:
            mu = float(rec["friction"])
            assert (mu == 0.0) == frictionless, name
            c = rec["contacts"][: rec["num_contacts"]]
            np.testing.assert_array_equal(c["lambda"][:, 1:], 0.0)
            np.testing.assert_array_equal(c["penalty"][:, 1:], 1.0)
            assert not c["stick"].any()
