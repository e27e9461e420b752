"""The forward-mode reference gradient (oracle World<DualT>) against FP64 central differences.

orc_tangent_run with an FP64 value half follows the FP64 trajectory, the one central finite
differences of orc_f64_run see; where that trajectory is smooth in the perturbation (no
contact, stick, clamp or feature change within +-h) the two must agree to FD accuracy. The
FP32-value variant (the adjoint's reference) runs the same template code, so this pins the
derivative rules (every DualT operation, sqrt, abs, min/max branches) it uses.

Also pinned by closed form: a free box's ballistic step (x1 = x0 + v dt + g dt^2 w, w = 0 from
rest-start acceleration, v1 = (x1 - x0)/dt) has d(z_final)/d(v0_z) = steps * dt.
"""
import numpy as np
import pytest


def _fd(oracle, st, prm, steps, key, base, u, h):
    p = oracle.oracle_f64_run(st, prm, steps, **{key: base + h * u})
    m = oracle.oracle_f64_run(st, prm, steps, **{key: base - h * u})
    return (p - m) / (2 * h)


def test_free_flight_closed_form(oracle):
    from avbd import scenes
    from avbd.state import default_params
    st = scenes.build_state(scenes.stacks(1, seed=2), oracle.OracleMaker())
    st.bodies["pos"][0, 2] += 5.0
    prm = default_params()
    u = np.zeros((2, 3))
    u[0, 2] = 1.0
    for fp64 in (False, True):
        _, tan = oracle.oracle_tangent_run(st, prm, 10, dvel=u, fp64=fp64)
        # warm start x1 = x0 + v dt (+ gravity weighted), finishStep v1 = (x1 - x0)/dt: z moves v dt per step
        assert tan[0, 2] == pytest.approx(10 * float(prm["dt"]), rel=1e-9)
        assert tan[0, 9] == pytest.approx(1.0, rel=1e-9)


def _slider(oracle):
    """One box sliding at 3 m/s on the ground with a spin: a contact that stays sliding."""
    from avbd import scenes
    from avbd.state import default_params
    st = scenes.build_state(scenes.stacks(1, seed=4), oracle.OracleMaker())
    st.bodies["vel"][0] = (3.0, 0.5, 0.0)
    st.bodies["ang_vel"][0] = (0.0, 0.0, 0.8)
    o = oracle.OracleWorld(st, default_params())
    o.step(3)
    return o.state(), default_params()


def _pile(oracle):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(30, seed=5)
    return scenes.build_state(sc, oracle.OracleMaker()), default_params(gravity=sc.gravity)


# (case, steps, kind): each a direction along which the FP64 trajectory is smooth within +-h.
CASES = [("pile30", 40, k) for k in ("dvel", "dangvel", "dmass", "dfriction")] + \
        [("pile30", 120, k) for k in ("dvel", "dmass", "dfriction")] + \
        [("slider", 30, k) for k in ("dvel", "dangvel", "dmass", "dfriction")]


@pytest.mark.parametrize("case,steps,kind", CASES)
def test_fp64_tangent_matches_central_differences(oracle, case, steps, kind):
    st, prm = _pile(oracle) if case == "pile30" else _slider(oracle)
    nb = len(st.bodies)
    dyn = st.bodies["mass"] > 0
    rng = np.random.default_rng(1)
    u = {"dvel": rng.normal(size=(nb, 3)) * dyn[:, None], "dangvel": rng.normal(size=(nb, 3)) * dyn[:, None],
         "dmass": rng.normal(size=nb) * dyn * st.bodies["mass"], "dfriction": rng.normal(size=nb) * 0.1}[kind]
    base = {"dvel": st.bodies["vel"], "dangvel": st.bodies["ang_vel"], "dmass": st.bodies["mass"],
            "dfriction": st.bodies["friction"]}[kind].astype(np.float64)
    _, t64 = oracle.oracle_tangent_run(st, prm, steps, fp64=True, **{kind: u})
    fd = _fd(oracle, st, prm, steps, {"dvel": "vel", "dangvel": "angvel", "dmass": "mass",
                                      "dfriction": "friction"}[kind], base, u, 1e-6)
    a, b = t64[dyn, 0:7], fd[dyn, 0:7]
    rel = np.abs(a - b).max() / max(np.abs(b).max(), 1e-12)
    assert rel < 1e-4, f"{case} {steps} {kind}: FP64 tangent vs central differences {rel:.2e}"


def test_fp32_branch_tangent_follows_the_fp32_trajectory(oracle):
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(30, seed=5)
    st = scenes.build_state(sc, oracle.OracleMaker())
    prm = default_params(gravity=sc.gravity)
    val, _ = oracle.oracle_tangent_run(st, prm, 90, dvel=np.ones((len(st.bodies), 3)))
    w = oracle.OracleWorld(st, prm)
    w.step(90)
    b = w.state().bodies
    np.testing.assert_array_equal(val[:, 0:3].astype(np.float32), b["pos"])
    np.testing.assert_array_equal(val[:, 3:7].astype(np.float32), b["rot"])
    np.testing.assert_array_equal(val[:, 7:10].astype(np.float32), b["vel"])


def test_frozen_branch_run_keeps_the_fp32_trajectory_and_converges(oracle):
    """orc_shadow_run: its FP32 shadows are the FP32 oracle's trajectory bit for bit (every branch
    the frozen-branch FP64 run takes is the FP32 run's), and on a 30-box pile falling and settling
    over 120 steps its central differences converge (h = 1e-4 and 1e-6 agree) to within 1e-3 of
    the FP32-path tangent, the derivative the GPU adjoint computes."""
    from avbd import scenes
    from avbd.state import default_params
    from conftest import body_block
    sc = scenes.pile(30, seed=5)
    st = scenes.build_state(sc, oracle.OracleMaker())
    prm = default_params(gravity=sc.gravity)
    steps = 120
    val, sh = oracle.oracle_shadow_run(st, prm, steps)
    w = oracle.OracleWorld(st, prm)
    w.step(steps)
    np.testing.assert_array_equal(sh.astype(np.float32), body_block(w.state()))
    assert np.abs(val - sh).max() < 1e-2
    nb = len(st.bodies)
    dyn = st.bodies["mass"] > 0
    rng = np.random.default_rng(1)
    for kind, key in (("dvel", "vel"), ("dangvel", "angvel"), ("dmass", "mass"), ("dfriction", "friction")):
        u = {"dvel": rng.normal(size=(nb, 3)) * dyn[:, None], "dangvel": rng.normal(size=(nb, 3)) * dyn[:, None],
             "dmass": rng.normal(size=nb) * dyn * st.bodies["mass"], "dfriction": rng.normal(size=nb) * 0.1}[kind]
        base = {"dvel": st.bodies["vel"], "dangvel": st.bodies["ang_vel"], "dmass": st.bodies["mass"],
                "dfriction": st.bodies["friction"]}[kind].astype(np.float64)
        _, t32 = oracle.oracle_tangent_run(st, prm, steps, **{kind: u})
        fds = []
        for h in (1e-4, 1e-6):
            p, _ = oracle.oracle_shadow_run(st, prm, steps, **{key: base + h * u})
            m, _ = oracle.oracle_shadow_run(st, prm, steps, **{key: base - h * u})
            fds.append((p - m) / (2 * h))
        scale = max(np.abs(fds[0][dyn, 0:7]).max(), 1e-12)
        assert np.abs(fds[0] - fds[1])[dyn, 0:7].max() / scale < 1e-5, f"{kind}: frozen FD not converged"
        rel = np.abs(t32[dyn, 0:7] - fds[0][dyn, 0:7]).max() / scale
        assert rel < 1e-3, f"{kind}: FP32-path tangent vs frozen-branch FD {rel:.2e}"


def _spring_chain(oracle):
    """The chain with its previous velocities at rest: a body built with prevVel = vel starts at the
    kink of the warm start's adaptive gravity weight (clamp(dv_z / (dt |g|), 0, 1) at dv = 0),
    where central differences see half of a one-sided slope; with prevVel = 0 every weight starts
    clamped well inside one side."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.spring_chain(6, seed=12)
    st = scenes.build_state(sc, oracle.OracleMaker())
    st.bodies["prev_vel"] = 0.0
    return st, default_params(gravity=sc.gravity)


@pytest.mark.parametrize("kind", ["dvel", "dangvel", "dmass"])
def test_fp64_tangent_with_springs_matches_central_differences(oracle, kind):
    """Spring stamps (spring.cpp:27-69) in the forward-mode reference: a free-flying chain of six
    boxes tied by springs (no contacts, so the FP64 map is smooth), 60 steps."""
    st, prm = _spring_chain(oracle)
    nb = len(st.bodies)
    rng = np.random.default_rng(3)
    u = {"dvel": rng.normal(size=(nb, 3)), "dangvel": rng.normal(size=(nb, 3)),
         "dmass": rng.normal(size=nb) * st.bodies["mass"]}[kind]
    base = {"dvel": st.bodies["vel"], "dangvel": st.bodies["ang_vel"], "dmass": st.bodies["mass"]}[kind]
    _, t64 = oracle.oracle_tangent_run(st, prm, 60, fp64=True, **{kind: u})
    fd = _fd(oracle, st, prm, 60, {"dvel": "vel", "dangvel": "angvel", "dmass": "mass"}[kind],
             base.astype(np.float64), u, 1e-6)
    rel = np.abs(t64[:, 0:7] - fd[:, 0:7]).max() / max(np.abs(fd[:, 0:7]).max(), 1e-12)
    assert rel < 1e-4, f"spring chain {kind}: FP64 tangent vs central differences {rel:.2e}"
    # the springs act: the chain's relative motion depends on them
    assert np.abs(t64[:, 0:3]).max() > 0.0


@pytest.mark.parametrize("k_ang", [0.0, 1000.0], ids=["ball", "stiff_angular"])
@pytest.mark.parametrize("kind", ["dvel", "dangvel", "dmass"])
def test_fp64_tangent_with_joints_matches_central_differences(oracle, k_ang, kind):
    """Joint stamps, duals and inits (joint.cpp:35-182) in the forward-mode reference: a free-flying
    chain of six boxes joined face to face (infinite linear stiffness; ball joints, or an angular
    stiffness of 1e3 for the angular rows), previous velocities at rest as for the spring chain."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.joint_chain(6, seed=14, k_ang=k_ang)
    st = scenes.build_state(sc, oracle.OracleMaker())
    st.bodies["prev_vel"] = 0.0
    prm = default_params(gravity=sc.gravity)
    nb = len(st.bodies)
    rng = np.random.default_rng(3)
    u = {"dvel": rng.normal(size=(nb, 3)), "dangvel": rng.normal(size=(nb, 3)),
         "dmass": rng.normal(size=nb) * st.bodies["mass"]}[kind]
    base = {"dvel": st.bodies["vel"], "dangvel": st.bodies["ang_vel"], "dmass": st.bodies["mass"]}[kind]
    _, t64 = oracle.oracle_tangent_run(st, prm, 30, fp64=True, **{kind: u})
    fd = _fd(oracle, st, prm, 30, {"dvel": "vel", "dangvel": "angvel", "dmass": "mass"}[kind],
             base.astype(np.float64), u, 1e-5)
    rel = np.abs(t64[:, 0:7] - fd[:, 0:7]).max() / max(np.abs(fd[:, 0:7]).max(), 1e-12)
    assert rel < 1e-6, f"joint chain {kind}: FP64 tangent vs central differences {rel:.2e}"


def _joined_pair(oracle, pair):
    from avbd import named_scenes, scenes
    from avbd.state import default_params
    sb = named_scenes.SceneBuilder()
    a = sb.box((0.8, 0.8, 0.8), 1.0, 0.5, (0.0, 0.0, 5.0))
    if pair:
        b = sb.box((0.8, 0.8, 0.8), 1.0, 0.5, (0.8, 0.0, 5.0))
        sb.joint(a, b, (0.4, 0.0, 0.0), (-0.4, 0.0, 0.0), 5000.0, 1000.0)
    sc = sb.scene("pair" if pair else "lone", -10.0, 20)
    return scenes.build_state(sc, oracle.OracleMaker()), default_params(gravity=sc.gravity)


@pytest.mark.parametrize("fp64", [False, True], ids=["fp32_values", "fp64_values"])
def test_zero_force_joint_takes_the_zero_subgradient(oracle, fp64):
    """Two boxes at rest joined face to face with finite stiffness (the Soft Body scene's joints,
    joint.cpp:35-182): the joint's C is exactly 0, so the ball joint's geometric-stiffness
    columns are zero vectors, whose length (maths.h:476) has a kink there, and the dual grows the
    penalty by beta |C| at |C|'s kink. A uniform velocity seed translates the pair rigidly and
    keeps C at 0, so each body's tangent is a lone box's: steps * dt * u in position (and finite,
    where sqrt' at 0 would give NaN). The other seeds must give finite tangents too."""
    pair, prm = _joined_pair(oracle, True)
    lone, _ = _joined_pair(oracle, False)
    u = np.array([0.3, -0.2, 1.0])
    _, tp = oracle.oracle_tangent_run(pair, prm, 20, dvel=np.tile(u, (2, 1)), fp64=fp64)
    _, t1 = oracle.oracle_tangent_run(lone, prm, 20, dvel=u[None, :], fp64=fp64)
    np.testing.assert_allclose(tp, np.tile(t1, (2, 1)), rtol=0, atol=1e-12)
    np.testing.assert_allclose(tp[:, 0:3], np.tile(u * 20 * float(prm["dt"]), (2, 1)), rtol=1e-6)
    for kind, seed in (("dangvel", np.ones((2, 3))), ("dmass", np.ones(2)), ("dfriction", np.ones(2))):
        _, t = oracle.oracle_tangent_run(pair, prm, 20, fp64=fp64, **{kind: seed})
        assert np.isfinite(t).all(), kind
