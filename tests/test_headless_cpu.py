"""Headless benchmark front end (avbd/headless.py) on the host: the JSON lines' keys and order
against the reference's (tests/golden/headless_keys.json, extracted from headless_server.cpp),
the number format, the command line's error paths, and the final-state measurements against
closed forms (resting and sunk bodies of every shape, tilted and undersized grounds, dense-id
order of the checksums, q / -q in the comparison)."""
import json
import math
import os

import numpy as np
import pytest

from avbd import headless as hl
from avbd.state import BODY, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_CYLINDER, SHAPE_SPHERE, STATS, State, default_params

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "headless_keys.json")
F = np.float32


def _body(shape, pos, size=(1, 1, 1), mass=1.0, rot=(0, 0, 0, 1), radius=0.5, srad=0.0, hlen=0.0, vel=(0, 0, 0),
          dense_id=0):
    b = np.zeros((), BODY)
    b["shape"], b["pos"], b["size"], b["mass"], b["rot"] = shape, pos, size, mass, rot
    b["radius"], b["shape_radius"], b["half_length"], b["vel"], b["dense_id"] = radius, srad, hlen, vel, dense_id
    return b


def _state(bodies):
    arr = np.array(bodies, BODY)
    return State(arr[::-1].copy())   # list order is newest first


def _ground(dense_id=0, size=(100, 100, 1), rot=(0, 0, 0, 1)):
    return _body(SHAPE_BOX, (0, 0, 0), size=size, mass=0.0, rot=rot, radius=70.0, dense_id=dense_id)


def _axis_angle(axis, ang):
    s = math.sin(ang / 2)
    return tuple(float(F(a * s)) for a in axis) + (float(F(math.cos(ang / 2))),)


def test_json_keys_follow_reference_order():
    keys = json.load(open(GOLDEN))
    st = _state([_ground(0), _body(SHAPE_SPHERE, (0, 0, 1), radius=0.5, srad=0.5, dense_id=1)])
    stat = {k: hl._Stat() for k in ("frame", "physics", "serialize", "iterate", "broadphase", "narrowphase",
                                    "coloring", "finish")}
    fields = hl.benchmark_fields("Sphere Stack", 10, 0, False, True, default_params(), st, np.zeros((), STATS), 3,
                                 stat, 1.0, 0, 0)
    assert [k for k, _ in fields] == keys["headlessBenchmark"]
    cfields = hl.comparison_fields("Rope", 120, 10, "CPU Reference", 1.0, 0.5, 0, st, st)
    assert [k for k, _ in cfields] == keys["backendComparison"]
    line = hl._json(fields)
    parsed = json.loads(line)
    assert parsed["type"] == "headlessBenchmark" and parsed["gravity"] == -10.0
    # std::fixed with 6 decimals for floats, plain integers and booleans
    assert '"gravity":-10.000000' in line and '"frames":10,' in line and '"collisionOnly":false' in line
    assert '"finalAabbMin":[' in line


def test_signature_uses_dense_order_and_shape_weights():
    # two bodies created in order 0, 1; the list holds them newest first
    a = _body(SHAPE_BOX, (1, 2, 3), vel=(1, 0, 0), dense_id=0, radius=0.25)
    b = _body(SHAPE_SPHERE, (-1, 0, 2), vel=(0, 0, -2), dense_id=1, radius=0.5, srad=0.5)
    sig = hl.final_state_signature(_state([a, b]))
    w_box, w_sph = 1.0, 1.0 + 1 * 0.125
    pos = 1 * w_box * (1 * 0.73 + 2 * 1.31 + 3 * 1.91) + 2 * w_sph * (-1 * 0.73 + 0 + 2 * 1.91)
    assert sig["positionChecksum"] == pytest.approx(pos, rel=1e-12)
    assert sig["positionAbsChecksum"] == pytest.approx(1 * (1 + 2 + 3) + 2 * w_sph * (1 + 0 + 2), rel=1e-12)
    assert sig["rotationChecksum"] == pytest.approx(1 * 1.43 + 2 * 1.43, rel=1e-12)
    assert sig["linearVelocityChecksum"] == pytest.approx(1 * 0.67 + 2 * (-2 * 1.61), rel=1e-12)
    assert sig["maxSpeed"] == 2.0 and sig["spheres"] == 1 and sig["boxes"] == 1
    assert sig["aabbMin"] == [-1.5, -0.5, 1.5] and sig["aabbMax"] == [1.25, 2.25, 3.25]
    assert sig["dynamicAabbMax"][2] == 3.25
    # non-finite fields count as zero
    c = _body(SHAPE_BOX, (np.nan, 0, 0), vel=(np.inf, 0, 0), dense_id=2)
    sig = hl.final_state_signature(_state([a, b, c]))
    assert math.isfinite(sig["positionChecksum"]) and sig["maxSpeed"] == 2.0


@pytest.mark.parametrize("shape,kw,extent", [
    (SHAPE_SPHERE, dict(srad=0.3, radius=0.3), 0.3),
    (SHAPE_BOX, dict(size=(2.0, 1.0, 0.5), radius=1.2), 0.25),
    (SHAPE_CAPSULE, dict(srad=0.2, hlen=0.6, radius=0.8), 0.8),       # upright: half length + radius
    (SHAPE_CYLINDER, dict(srad=0.4, hlen=0.3, radius=0.5), 0.3),      # upright: half length
])
def test_resting_and_sunk_bodies(shape, kw, extent):
    top = F(0.5)
    for sink in (0.0, 0.03):
        z = float(top + F(extent) - F(sink))
        gc = hl.ground_clearance(_state([_ground(0), _body(shape, (0, 0, z), dense_id=1, **kw)]))
        assert gc["hasGround"] and gc["groundTop"] == 0.5
        assert gc["maxGroundBodyPenetration"] == pytest.approx(sink, abs=2e-6)
        assert gc["belowGroundBodies"] == (1 if sink else 0)
        if shape == SHAPE_SPHERE:
            assert gc["maxPenetration"] == pytest.approx(sink, abs=2e-6)
            assert gc["belowGroundSpheres"] == (1 if sink else 0)
        else:
            assert gc["dynamicSpheres"] == 0 and gc["maxPenetration"] == 0.0


def test_rotated_extents_closed_forms():
    th = 0.6
    q = _axis_angle((1, 0, 0), th)
    b = np.array([_body(SHAPE_BOX, (0, 0, 0), size=(2, 1, 0.5), rot=q),
                  _body(SHAPE_CAPSULE, (0, 0, 0), rot=q, srad=0.2, hlen=0.6),
                  _body(SHAPE_CYLINDER, (0, 0, 0), rot=q, srad=0.4, hlen=0.3),
                  _body(SHAPE_CAPSULE, (0, 0, 0), rot=_axis_angle((0, 1, 0), math.pi / 2), srad=0.2, hlen=0.6)], BODY)
    e = hl.ground_contact_extent_z(b)
    c, s = math.cos(th), math.sin(th)
    np.testing.assert_allclose(e, [s * 0.5 + c * 0.25, c * 0.6 + 0.2, c * 0.3 + s * 0.4, 0.2], rtol=0, atol=2e-6)


def test_ground_receivers():
    sph = _body(SHAPE_SPHERE, (0, 0, 0.2), srad=0.3, radius=0.3, dense_id=1)
    # tilted slab (normal z below 0.98), too-small slab and too-thick slab are not grounds
    for g in (_ground(rot=_axis_angle((1, 0, 0), 0.3)), _ground(size=(7, 100, 1)), _ground(size=(100, 100, 3))):
        gc = hl.ground_clearance(_state([g, sph]))
        assert not gc["hasGround"] and gc["dynamicGroundBodies"] == 1 and gc["maxPenetration"] == 0.0
    # the highest receiver wins
    g2 = _body(SHAPE_BOX, (0, 0, 1.0), size=(10, 10, 1), mass=0.0, dense_id=2)
    gc = hl.ground_clearance(_state([_ground(0), sph, g2]))
    assert gc["groundTop"] == 1.5 and gc["maxPenetration"] == pytest.approx(1.5 - 0.2 + 0.3, abs=1e-6)


def test_compare_worlds():
    a = _body(SHAPE_BOX, (0, 0, 1), rot=_axis_angle((0, 0, 1), 0.4), dense_id=0)
    b = a.copy()
    b["rot"] = -b["rot"]                    # same rotation
    b["pos"] = (0.3, 0.4, 1)
    c = hl.compare_worlds(_state([a]), _state([b]))
    assert c["maxRotationError"] == 0.0 and c["maxPositionError"] == pytest.approx(0.5)
    assert c["rmsPositionError"] == pytest.approx(0.5) and c["bodyCountMismatch"] == 0
    c = hl.compare_worlds(_state([a]), _state([a, _body(SHAPE_BOX, (0, 0, 0), dense_id=1)]))
    assert c["bodyCountMismatch"] == 1 and c["comparedBodies"] == 1


def test_cli_errors():
    assert hl.main(["--benchmark-scene", "no such scene", "--no-stream"]) == 2
    assert hl.main(["--benchmark-scene", "Pyramid", "--physics-backend", "webgpu-avbd"]) == 2
    assert hl.main(["--benchmark-scene", "Pyramid", "--collision-only"]) == 2
    assert hl.main(["--compare-scene", "Rope"]) == 2
    assert hl.main([]) == 2


def test_broadphase_names_follow_the_reference():
    """applyBroadphaseMode (headless_server.cpp:579-591); no flag keeps the Solver's default grid."""
    assert hl.broadphase_mode(None) == 1
    assert [hl.broadphase_mode(n) for n in ("allpairs", "grid", "uniformgrid", "spatialhash", "sap",
                                             "sweepandprune")] == [0, 1, 1, 1, 2, 2]
    assert hl.broadphase_mode("octree") == 1
