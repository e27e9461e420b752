"""Closed-form and invariant checks that pin the oracle's building blocks
independently of the reference fixtures (each would catch a dropped term, a
wrong sign or a transposed operand)."""
import math

import numpy as np
import pytest


def _spd(rng, n=6, cond=1e3):
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.exp(rng.uniform(0, math.log(cond), n))
    return q @ np.diag(ev) @ q.T


def test_ldl_solve_matches_dense_solve(oracle):
    """maths.h:481-547 reads the lower triangle of [[Lin, Cross^T],[Cross, Ang]]."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        A = _spd(rng)
        b = rng.normal(size=6)
        lin, cross, ang = A[:3, :3], A[3:, :3], A[3:, 3:]
        # poison the upper triangles: the solve must not read them
        lin_u, ang_u = lin.copy(), ang.copy()
        lin_u[np.triu_indices(3, 1)] = 1e9
        ang_u[np.triu_indices(3, 1)] = -1e9
        x = oracle.oracle_solve6(lin_u, ang_u, cross, b, fp64=True)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
        x32 = oracle.oracle_solve6(lin_u, ang_u, cross, b, fp64=False)
        np.testing.assert_allclose(x32, np.linalg.solve(A, b), rtol=2e-3, atol=1e-5)


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def test_quaternion_identities(oracle):
    """q - q = 0; q + w is unit and matches the exponential map to first order; rotate is orthogonal."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.normal(size=3)
        w = rng.normal(size=3) * 1e-3
        d, integ, rot = oracle.oracle_quat_ops(q, q, w, v)
        np.testing.assert_allclose(d, 0, atol=2e-6)
        assert abs(np.linalg.norm(integ) - 1.0) < 2e-6
        # exact composition exp(w) * q
        th = np.linalg.norm(w)
        e = np.concatenate([np.sin(th / 2) * w / th, [np.cos(th / 2)]])
        np.testing.assert_allclose(integ, _qmul(e, q), atol=1e-5)
        # difference recovers w: (exp(w) q) - q ~= w
        d2, _, _ = oracle.oracle_quat_ops(_qmul(e, q), q, w, v)
        np.testing.assert_allclose(d2, w, rtol=1e-3, atol=1e-7)
        # rotation: |Rv| = |v| and R from the quaternion's matrix form
        x, y, z, ww = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - ww * z), 2 * (x * z + ww * y)],
                      [2 * (x * y + ww * z), 1 - 2 * (x * x + z * z), 2 * (y * z - ww * x)],
                      [2 * (x * z - ww * y), 2 * (y * z + ww * x), 1 - 2 * (x * x + y * y)]])
        np.testing.assert_allclose(rot, R @ v, atol=1e-5)


def test_box_mass_properties_closed_form(oracle):
    rng = np.random.default_rng(2)
    mk = oracle.OracleMaker()
    for _ in range(50):
        s = rng.uniform(0.1, 3, 3)
        rho = rng.uniform(0.1, 5)
        b = mk.make_box_body(s, rho, 0.5, (0, 0, 0), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0))
        s32 = s.astype(np.float32).astype(np.float64)
        m = s32.prod() * np.float32(rho)
        np.testing.assert_allclose(b["mass"], m, rtol=1e-6)
        np.testing.assert_allclose(b["moment"], m * np.array([s32[1] ** 2 + s32[2] ** 2, s32[0] ** 2 + s32[2] ** 2,
                                                              s32[0] ** 2 + s32[1] ** 2]) / 12, rtol=1e-5)
        np.testing.assert_allclose(b["radius"], np.linalg.norm(s32 / 2), rtol=1e-6)


def _one_body_state(oracle, pos, vel, size=(1, 1, 1)):
    from avbd.state import State
    b = oracle.OracleMaker().make_box_body(size, 1.0, 0.5, pos, (0, 0, 0, 1), vel, (0, 0, 0))
    return State(np.array([b]))


def test_free_body_is_ballistic(oracle):
    """An unconstrained body solves to its inertial prediction: x_k = x0 + k v0 dt + g dt^2 k(k+1)/2."""
    from avbd.state import default_params
    prm = default_params(gravity=-9.81)
    st = _one_body_state(oracle, (1.0, 2.0, 50.0), (0.5, -0.25, 3.0))
    w = oracle.OracleWorld(st, prm, fp64=True)
    dt, g = float(np.float32(1 / 60)), float(np.float32(-9.81))
    for k in range(1, 61):
        w.step(1)
        p = w.bodies_f64()[0]
        exp_z = 50.0 + k * 3.0 * dt + g * dt * dt * k * (k + 1) / 2
        np.testing.assert_allclose(p[2], exp_z, rtol=1e-12)
        np.testing.assert_allclose(p[9], 3.0 + k * g * dt, rtol=1e-10)
        np.testing.assert_allclose(p[0], 1.0 + k * 0.5 * dt, rtol=1e-12)


def test_spring_one_step_minimises_quadratic_energy(oracle):
    """Body hung under a static anchor by a spring (no gravity): one primal solve is the exact minimiser
    of m/(2dt^2)(z - z~)^2 + k/2 (|d| - rest)^2 along the axis (spring.cpp:27-69, maths.h:481)."""
    from avbd.state import SPRING, State, default_params
    mk = oracle.OracleMaker()
    anchor = mk.make_box_body((1, 1, 1), 0.0, 0.5, (0, 0, 10), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0))
    body = mk.make_box_body((1, 1, 1), 2.0, 0.5, (0, 0, 5), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0))
    sp = np.zeros(1, SPRING)
    sp[0]["body_a"], sp[0]["body_b"], sp[0]["serial"] = 0, 1, 0
    sp[0]["rest"], sp[0]["stiffness"] = 4.0, 300.0
    st = State(np.array([anchor, body]), springs=sp, next_serial=1)
    w = oracle.OracleWorld(st, default_params(gravity=0.0), fp64=True)
    w.step(1)
    z = w.bodies_f64()[1][2]
    m, dt = 2.0, float(np.float32(1 / 60))
    a = m / dt ** 2
    expect = (a * 5.0 + 300.0 * (10.0 - 4.0)) / (a + 300.0)
    np.testing.assert_allclose(z, expect, rtol=1e-10)


def test_world_joint_holds_pendulum(oracle):
    """A world-anchored ball joint (joint.cpp:56-123) keeps its anchor error small while the box swings."""
    from avbd.state import JOINT, State, default_params
    mk = oracle.OracleMaker()
    body = mk.make_box_body((1, 0.5, 0.5), 1.0, 0.5, (1.5, 0, 10), (0, 0, 0, 1), (0, 0, 0), (0, 0, 0))
    j = np.zeros(1, JOINT)
    j[0]["body_a"], j[0]["body_b"], j[0]["serial"] = -1, 0, 0
    j[0]["r_a"] = (1.0, 0, 10)
    j[0]["r_b"] = (-0.5, 0, 0)
    j[0]["stiffness_lin"], j[0]["stiffness_ang"], j[0]["fracture"] = np.inf, 0.0, np.inf
    j[0]["torque_arm"] = 1.0 + 0.25 + 0.25
    st = State(np.array([body]), joints=j, next_serial=1)
    w = oracle.OracleWorld(st, default_params(gravity=-9.81), fp64=True)
    maxerr = 0.0
    for _ in range(120):
        w.step(1)
        p = w.bodies_f64()[0]
        x, q = p[:3], p[3:7]
        xx, yy, zz, ww = q
        R = np.array([[1 - 2 * (yy * yy + zz * zz), 2 * (xx * yy - ww * zz), 2 * (xx * zz + ww * yy)],
                      [2 * (xx * yy + ww * zz), 1 - 2 * (xx * xx + zz * zz), 2 * (yy * zz - ww * xx)],
                      [2 * (xx * zz - ww * yy), 2 * (yy * zz + ww * xx), 1 - 2 * (xx * xx + yy * yy)]])
        maxerr = max(maxerr, np.linalg.norm(x + R @ np.array([-0.5, 0, 0]) - np.array([1.0, 0, 10])))
    assert maxerr < 5e-3
    assert w.bodies_f64()[0][2] < 9.9  # it swung down


def test_fp32_and_fp64_oracles_agree_on_one_step(oracle):
    """The FP64 run states the FP32 error of a single step (relative to the pose scale)."""
    from conftest import load_golden, state_from_golden
    from avbd.state import default_params
    z = load_golden("pile.npz")
    pre = state_from_golden(z, "pre2")
    prm = default_params(gravity=float(z["gravity"]))
    w32 = oracle.OracleWorld(pre, prm)
    w64 = oracle.OracleWorld(pre, prm, fp64=True)
    w32.step(1)
    w64.step(1)
    a, b = w32.bodies_f64(), w64.bodies_f64()
    scale = np.maximum(1.0, np.abs(b[:, :3]))
    assert np.max(np.abs(a[:, :3] - b[:, :3]) / scale) < 1e-4
