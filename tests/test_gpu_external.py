"""External pair and contact sources on the GPU (avbd_set_external_pairs / _contacts: the reference's
Solver::stepCpuReferenceWithExternalBroadphase and external contact table, solver.cpp:579-626,
939-987, manifold.cpp:110-121) against the compiled reference's fixtures and the oracle: full state,
the hit sequence and the step's checks, hits and creations, bit for bit."""
import numpy as np
import pytest

from conftest import assert_states_equal, load_golden, state_from_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


def _counts(w):
    st = w.stats()
    return (int(st["pair_checks"]), int(st["sphere_hits"]), int(st["constrained_hits"]), int(st["manifolds_created"]))


@pytest.mark.parametrize("key", ["boxes", "mixed"])
def test_external_sources_match_reference_fixture(gpu, key):
    z = load_golden("external.npz")
    from avbd.state import default_params
    st = state_from_golden(z, f"{key}.start")
    prm = default_params(gravity=float(z[f"{key}.gravity"]))
    w = gpu.HipWorld.for_state(st, prm)
    w.set_external_pairs(z[f"{key}.pairs"])
    w.set_external_contacts(z[f"{key}.contacts"])
    pair_steps = int(z[f"{key}.pair_steps"])
    print(f"external fixture {key}: {len(st.bodies)} bodies, {len(z[f'{key}.pairs'])} pairs, "
          f"{len(z[f'{key}.contacts'])} table entries", flush=True)
    for k in range(int(z[f"{key}.steps"])):
        if k == pair_steps:
            w.clear_external_pairs()  # the table stays; the handle's own broadphase (SAP) runs
        w.step(1)
        assert _counts(w) == tuple(int(x) for x in z[f"{key}.counts"][k]), f"{key} step {k}: counts"
        assert_states_equal(w.state(), state_from_golden(z, f"{key}.s{k}"), f"{key} step {k}")
    w.close()


def _user_broadphase(state, rng, slack=0.5):
    """Every body pair whose bounding spheres come within `slack`, shuffled and randomly oriented,
    with some repeated: a caller's own broadphase."""
    pos = state.bodies["pos"].astype(np.float64)
    rad = state.bodies["radius"].astype(np.float64)
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2)
    a, b = np.nonzero(np.triu(d <= rad[:, None] + rad[None, :] + slack, 1))
    P = np.stack([a, b], 1)
    P = P[rng.permutation(len(P))]
    flip = rng.random(len(P)) < 0.5
    P[flip] = P[flip][:, ::-1]
    return np.concatenate([P, P[rng.choice(len(P), len(P) // 20)]]).astype(np.int32)


def test_external_pairs_pile_vs_oracle(gpu, oracle):
    """A 1500-box pile stepped on the GPU into contact, then 6 steps with a caller's pair list (all
    pairs within 0.5 m of touching, recomputed each step, shuffled, reversed, 5 % repeated) and a
    contact table taken from the oracle's own manifolds of the previous step (anchors moved by
    1 mm): hits, checks, creations and full state equal the oracle's every step."""
    from avbd import scenes
    from avbd.state import EXTERNAL_CONTACT, default_params
    sc = scenes.pile(1500, seed=21)
    prm = default_params(gravity=sc.gravity)
    g = gpu.HipWorld.for_state(scenes.build_state(sc, gpu.Maker()), prm)
    g.step(110)
    pre = g.state()
    c = oracle.OracleWorld(pre, prm)
    rng = np.random.default_rng(5)
    for k in range(6):
        cur = c.state()
        P = _user_broadphase(cur, rng)
        m = cur.manifolds[rng.permutation(len(cur.manifolds))[: len(cur.manifolds) // 4]]
        T = np.zeros(len(m), EXTERNAL_CONTACT)
        T["body_a"], T["body_b"], T["num_contacts"] = m["body_a"], m["body_b"], m["num_contacts"]
        T["basis"], T["contacts"] = m["basis"], m["contacts"]
        T["contacts"]["r_a"] += np.float32(1e-3)
        g.set_external_pairs(P)
        g.set_external_contacts(T)
        c.set_external(P, T)
        g.step(1)
        c.step(1)
        print(f"external pile step {k}: {len(P)} pairs, {len(T)} table entries, counts {_counts(g)}", flush=True)
        np.testing.assert_array_equal(g.hits(), c.hits(), err_msg=f"step {k}: hits")
        assert int(g.stats()["pair_checks"]) == c.pair_checks
        assert_states_equal(g.state(), c.state(), f"external pile step {k}")
    g.clear_external_pairs()
    g.set_external_contacts(None)
    c.set_external(None, None, pairs_on=False)
    g.step(1)
    c.step(1)
    assert_states_equal(g.state(), c.state(), "after clearing the external sources")
    g.close()


def test_external_sources_validation(gpu):
    from avbd import scenes
    from avbd.state import EXTERNAL_CONTACT, batch_states, default_params
    sts = [scenes.build_state(scenes.stack(seed=s), gpu.Maker()) for s in (1, 2)]
    prm = default_params()
    w = gpu.HipWorld.for_batch(sts, prm)
    with pytest.raises(gpu.AvbdError) as e:
        w.set_external_pairs([[0, 1]])
    assert e.value.code == -5
    with pytest.raises(gpu.AvbdError) as e:
        w.set_external_contacts(np.zeros(1, EXTERNAL_CONTACT))
    assert e.value.code == -5
    w.upload(sts[0])
    w.set_external_pairs(np.zeros((0, 2), np.int32))  # no pairs: no manifolds from the broadphase
    w.step(1)
    assert int(w.stats()["pair_checks"]) == 0 and int(w.stats()["manifold_count"]) == 0
    batch, offsets = batch_states(sts)
    with pytest.raises(gpu.AvbdError) as e:
        w.upload_batch(batch, offsets)
    assert e.value.code == -5
    with pytest.raises(gpu.AvbdError) as e:
        w.adjoint_forward(2, 1)
    assert e.value.code == -5
    w.clear_external_pairs()
    w.upload_batch(batch, offsets)
    w.step(2)
    w.close()
