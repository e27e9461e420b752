"""Invariants of the oracle's broadphase, colouring and dual update that do not
depend on the reference fixtures."""
import numpy as np
import pytest

from conftest import load_golden, state_from_golden


def _random_bodies(oracle, rng, n, spread):
    from avbd.state import State
    mk = oracle.OracleMaker()
    bodies = []
    for _ in range(n):
        s = rng.uniform(0.2, 2.0, 3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        bodies.append(mk.make_box_body(s, 1.0, 0.5, rng.uniform(-spread, spread, 3), q, (0, 0, 0), (0, 0, 0)))
    return State(np.array(bodies))


@pytest.mark.parametrize("seed,n,spread", [(0, 2, 1.0), (1, 40, 3.0), (2, 300, 8.0), (3, 600, 4.0)])
def test_sap_hits_equal_brute_force(oracle, seed, n, spread):
    """The sweep finds exactly the bounding-sphere-overlapping pairs (the set is axis independent)."""
    from avbd.state import default_params
    rng = np.random.default_rng(seed)
    st = _random_bodies(oracle, rng, n, spread)
    w = oracle.OracleWorld(st, default_params())
    w.prepare()
    hits = w.hits()
    got = {(min(a, b), max(a, b)) for a, b, _ in hits}
    assert len(got) == len(hits), "a pair was emitted twice"
    pos = st.bodies["pos"]
    rad = st.bodies["radius"]
    exp = set()
    for i in range(n):
        for j in range(i + 1, n):
            dp = (pos[i] - pos[j]).astype(np.float32)
            d2 = np.float32(np.float32(dp[0] * dp[0]) + np.float32(dp[1] * dp[1])) + np.float32(dp[2] * dp[2])
            r = np.float32(rad[i] + rad[j])
            if not d2 > np.float32(r * r):
                exp.add((i, j))
    assert got == exp


def _adjacency(state):
    n = len(state.bodies)
    adj = [set() for _ in range(n)]
    for tab in (state.manifolds, state.joints, state.springs):
        for r in tab:
            a, b = int(r["body_a"]), int(r["body_b"])
            if a >= 0 and b >= 0:
                adj[a].add(b)
                adj[b].add(a)
    return adj


@pytest.mark.parametrize("fixture,prefix", [("pile.npz", "pre1"), ("pile.npz", "pre2"), ("pyramid_chain.npz", "pre")])
def test_colouring_is_valid_and_greedy(oracle, fixture, prefix):
    """No two adjacent solved bodies share a colour, and each colour is the smallest not used by a
    lower-index neighbour (sequential greedy, avbd_gpu_solver.cpp:1812-1865)."""
    from avbd.state import default_params
    z = load_golden(fixture)
    w = oracle.OracleWorld(state_from_golden(z, prefix), default_params(gravity=float(z["gravity"])))
    w.prepare()
    st = w.state()
    col = w.colors()
    adj = _adjacency(st)
    mass = st.bodies["mass"]
    for i in range(len(col)):
        if col[i] < 0:
            continue
        used = {int(col[j]) for j in adj[i] if j < i and col[j] >= 0}
        mex = 0
        while mex in used:
            mex += 1
        assert col[i] == mex
        for j in adj[i]:
            if col[j] >= 0:
                assert col[j] != col[i]
    solved = (mass > 0) & np.array([len(a) > 0 or False for a in adj])
    # bodies touching only static bodies are solved too; unsolved ones have colour -1
    assert np.all(col[mass <= 0] == -1)
    assert np.all(col[solved] >= 0)


def test_dual_invariants_after_steps(oracle):
    """lambda_n <= 0, |lambda_t| <= mu |lambda_n|, 1 <= penalty <= 1e10 (manifold.cpp:231-280)."""
    from avbd.state import default_params
    z = load_golden("pile.npz")
    w = oracle.OracleWorld(state_from_golden(z, "pre1"), default_params(gravity=float(z["gravity"])))
    for _ in range(3):
        w.step(1)
        st = w.state()
        m = st.manifolds
        mask = np.arange(8)[None, :] < m["num_contacts"][:, None]
        lam = m["contacts"]["lambda"][mask]
        pen = m["contacts"]["penalty"][mask]
        mu = np.repeat(m["friction"], m["num_contacts"])
        assert np.all(lam[:, 0] <= 0)
        assert np.all(np.hypot(lam[:, 1], lam[:, 2]) <= mu * np.abs(lam[:, 0]) * (1 + 1e-5) + 1e-30)
        assert np.all(pen >= 1.0) and np.all(pen <= 1e10)


def test_warm_start_decay_of_persisting_contacts(oracle):
    """A contact kept across a step starts from lambda * alpha * gamma and penalty * gamma (Eq. 19)."""
    from avbd.state import default_params
    z = load_golden("pile.npz")
    pre = state_from_golden(z, "pre2")
    prm = default_params(gravity=float(z["gravity"]))
    w = oracle.OracleWorld(pre, prm)
    w.prepare()
    st = w.state()
    old = {(int(m["body_a"]), int(m["body_b"])): m for m in pre.manifolds}
    checked = 0
    a, g = np.float32(prm["alpha"]), np.float32(prm["gamma"])
    for m in st.manifolds:
        o = old.get((int(m["body_a"]), int(m["body_b"])))
        if o is None:
            continue
        for i in range(m["num_contacts"]):
            for j in range(o["num_contacts"]):
                if m["contacts"][i]["feature"] == o["contacts"][j]["feature"]:
                    exp = (o["contacts"][j]["lambda"] * a * g).astype(np.float32)
                    np.testing.assert_array_equal(m["contacts"][i]["lambda"], exp)
                    pen = np.clip((o["contacts"][j]["penalty"] * g).astype(np.float32), 1.0, 1e10)
                    np.testing.assert_array_equal(m["contacts"][i]["penalty"], pen)
                    checked += 1
                    break
    assert checked > 100
