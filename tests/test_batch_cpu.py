"""Host side of the batched step: the world generator and the batch/split of
state tables (avbd_batch_upload's layout), without a GPU."""
import math

import numpy as np

from conftest import assert_states_equal


def test_batch_generator_sizes_and_kinds():
    from avbd import scenes
    ws = scenes.batch(512, seed=4)
    sizes = np.array([w.num_bodies for w in ws])
    assert sizes.min() >= 8 and sizes.max() <= 512 + 1
    # log-uniform in [8, 512]: the median sits near sqrt(8 * 512) = 64
    assert 40 < np.median(sizes) < 100
    kinds = {w.name.rstrip("0123456789").replace("pyramid0_chain", "chain") for w in ws}
    assert kinds == {"stacks", "pile", "chain"}
    again = scenes.batch(512, seed=4)
    assert all(np.array_equal(a.pos, b.pos) and a.name == b.name for a, b in zip(ws, again))


def test_default_batch_is_4096_worlds():
    from avbd import scenes
    ws = scenes.batch(4096, seed=4, kinds=("stacks",))
    assert len(ws) == 4096
    total = sum(w.num_bodies for w in ws)
    # E[n] of log-uniform [8, 512] is 504 / ln 64 = 121.2 bodies per world
    assert abs(total / 4096 - 504 / math.log(64)) < 6


def test_batch_split_round_trip(oracle):
    from avbd import scenes
    from avbd.state import batch_states, split_state, default_params
    ws = scenes.batch(12, seed=11, hi=64)
    sts = [scenes.build_state(w, oracle.OracleMaker()) for w in ws]
    # give the worlds manifolds: step each a few times on the oracle
    prm = default_params()
    stepped = []
    for st in sts:
        o = oracle.OracleWorld(st, prm)
        o.step(3)
        stepped.append(o.state())
    batch, off = batch_states(stepped)
    assert off[-1] == len(batch.bodies) and len(off) == 13
    assert batch.next_serial == sum(s.next_serial for s in stepped)
    assert len(np.unique(np.concatenate([batch.joints["serial"], batch.springs["serial"],
                                         batch.manifolds["serial"]]))) == \
        len(batch.joints) + len(batch.springs) + len(batch.manifolds)
    for w, (a, b) in enumerate(zip(stepped, split_state(batch, off))):
        assert_states_equal(a, b, f"world {w}")
        for tbl in ("joints", "springs", "manifolds"):
            x, y = getattr(a, tbl), getattr(b, tbl)
            assert np.array_equal(np.argsort(x["serial"], kind="stable"), np.argsort(y["serial"], kind="stable"))
