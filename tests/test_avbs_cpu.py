"""AVBS v1 snapshot encoding (viewer_bridge.cpp:796-828) in the oracle: the encoder against the
viewer's decoder restated (viewer/src/main.js:986-1050), the header against the format's fixed
bytes, and the ordering / material rules on states with every shape and shuffled dense ids."""
import numpy as np

from oracle.avbs import MAGIC, decode_avbs, encode_avbs


def _mixed_state(oracle, seed=3, n=60):
    from avbd import scenes
    st = scenes.build_state(scenes.mixed_pile(n, seed=seed), oracle.OracleMaker())
    rng = np.random.default_rng(seed)
    ids = rng.choice(10 * len(st.bodies), len(st.bodies), replace=False)   # holes and any order
    st.bodies["dense_id"] = ids
    return st


def test_header_bytes():
    payload = encode_avbs(np.zeros(0, __import__("avbd.state", fromlist=["BODY"]).BODY), 7, "Pour")
    assert payload[:4] == b"AVBS"            # 0x53425641 little-endian
    assert int.from_bytes(payload[:4], "little") == MAGIC
    assert payload[4:8] == (1).to_bytes(4, "little")
    assert payload[8:16] == (7).to_bytes(8, "little")
    assert payload[16:20] == (4).to_bytes(4, "little") and payload[20:24] == (0).to_bytes(4, "little")
    assert payload[24:] == b"Pour"


def test_encode_decode_round_trip(oracle):
    st = _mixed_state(oracle)
    payload = encode_avbs(st.bodies, 123456789012, "Mixed pile")
    d = decode_avbs(payload)
    assert d["frame"] == 123456789012 and d["scene"] == "Mixed pile"
    assert len(payload) == 24 + len("Mixed pile") + 64 * len(st.bodies)
    b = st.bodies
    order = np.argsort(b["dense_id"], kind="stable")
    assert [x["id"] for x in d["bodies"]] == sorted(b["dense_id"].tolist())
    for rec, i in zip(d["bodies"], order):
        assert rec["shape"] == b["shape"][i] and rec["dynamic"] == (b["mass"][i] > 0)
        np.testing.assert_array_equal(np.float32(rec["position"]), b["pos"][i])
        np.testing.assert_array_equal(np.float32(rec["rotation"]), b["rot"][i])
        np.testing.assert_array_equal(np.float32(rec["size"]), b["size"][i])
        assert np.float32(rec["radius"]) == b["shape_radius"][i]
        assert np.float32(rec["half_length"]) == b["half_length"][i]
        expect = 0 if b["mass"][i] <= 0 else {0: 1, 1: 3, 2: 5, 3: 4}[int(b["shape"][i])]
        assert rec["material"] == expect
    # every shape and the static material occur
    assert {x["shape"] for x in d["bodies"]} == {0, 1, 2, 3}
    assert 0 in {x["material"] for x in d["bodies"]}
