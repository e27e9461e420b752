"""Shared pytest setup: import paths, markers, and the small loaders tests use.

CPU tests (-m "not gpu") check the oracle against the reference's golden
fixtures, the invariants, and the C-ABI library's host side. GPU tests
(-m gpu) run the HIP path through the C ABI and compare it with the oracle.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "avbd-demo3d_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_oracle()


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def scene_from_golden(z, prefix=""):
    """The Scene stored in a fixture (keys under `prefix`, e.g. "sphere_stack10.")."""
    from avbd.scenes import Scene
    g = lambda k: z[prefix + k]
    has_shape = (prefix + "s_shape") in z
    return Scene(name=prefix.rstrip(".") or "golden", size=g("s_size"), density=g("s_density"),
                 friction=g("s_friction"), pos=g("s_pos"), rot=g("s_rot"), vel=g("s_vel"), angvel=g("s_angvel"),
                 ftype=g("f_type"), fa=g("f_a"), fb=g("f_b"), ra=g("f_ra"), rb=g("f_rb"), p0=g("f_p0"), p1=g("f_p1"),
                 p2=g("f_p2"), gravity=float(g("gravity")),
                 shape=g("s_shape") if has_shape else None,
                 shape_radius=g("s_shape_radius") if has_shape else None,
                 half_length=g("s_half_length") if has_shape else None)


def state_from_golden(z, prefix):
    from avbd.state import State
    return State(z[f"{prefix}_bodies"], z[f"{prefix}_joints"], z[f"{prefix}_springs"], z[f"{prefix}_manifolds"],
                 int(z[f"{prefix}_next_serial"]))


def body_block(state):
    b = state.bodies
    return np.concatenate([b["pos"], b["rot"], b["vel"], b["ang_vel"]], axis=1).astype(np.float32)


def manifold_block(state):
    """(pairs, per-contact float payload, features, counts) in ascending serial order."""
    m = state.manifolds
    pairs = np.stack([m["body_a"], m["body_b"]], axis=1) if len(m) else np.zeros((0, 2), np.int32)
    mask = np.arange(8)[None, :] < m["num_contacts"][:, None]
    c = m["contacts"]
    payload = np.concatenate([c["r_a"], c["r_b"], c["c0"], c["penalty"], c["lambda"]], axis=2)
    payload = np.where(mask[:, :, None], payload, 0).astype(np.float32)
    feats = np.where(mask, c["feature"], 0)
    sticks = np.where(mask, c["stick"], 0)
    return pairs, payload, feats, sticks, m["num_contacts"].copy(), m["basis"].copy(), m["friction"].copy()


def assert_states_equal(a, b, what=""):
    """Exact FP32 equality of bodies and manifolds (value equality: +0 == -0)."""
    assert len(a.bodies) == len(b.bodies), what
    np.testing.assert_array_equal(body_block(a), body_block(b), err_msg=f"{what}: bodies")
    pa, xa, fa, sa, na, ba, ra = manifold_block(a)
    pb, xb, fb, sb, nb, bb, rb = manifold_block(b)
    np.testing.assert_array_equal(pa, pb, err_msg=f"{what}: manifold pairs/order")
    np.testing.assert_array_equal(na, nb, err_msg=f"{what}: contact counts")
    np.testing.assert_array_equal(fa, fb, err_msg=f"{what}: feature keys")
    np.testing.assert_array_equal(sa, sb, err_msg=f"{what}: stick flags")
    np.testing.assert_array_equal(xa, xb, err_msg=f"{what}: contact payload")
    np.testing.assert_array_equal(ba, bb, err_msg=f"{what}: basis")
    np.testing.assert_array_equal(ra, rb, err_msg=f"{what}: friction")
    ja, jb = a.joints, b.joints
    assert len(ja) == len(jb), what
    for f in ("penalty_lin", "penalty_ang", "lambda_lin", "lambda_ang", "c0_lin", "c0_ang", "broken"):
        np.testing.assert_array_equal(ja[f], jb[f], err_msg=f"{what}: joint {f}")


@pytest.fixture(scope="session")
def oracle():
    import oracle.bindings as ob
    ob.load_oracle()
    return ob
