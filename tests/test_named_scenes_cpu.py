"""Named scene registry (avbd/named_scenes.py) and the mesh-lattice importer (avbd/mesh_lattice.py).

Each of the reference's 46 non-empty scenes, built by this build's registry and record builders,
must equal the state the reference's own scene function creates (oracle/_ref): every body field the
constructors set, every joint and spring, the IgnoreCollision pairs and the serial counter. The
lattice importer is also checked against closed forms: a box mesh fills a full grid with face
joints only, a thin sheet is a surface-only lattice, and stiffness follows the omega*dt rule.
"""
import math

import numpy as np
import pytest

from avbd import mesh_lattice as ml
from avbd import named_scenes as ns
from avbd import scenes
from avbd.state import default_params

# fields the reference sets only inside a step (start-of-step copies, joint C0): uninitialised in a
# freshly built scene
_STEP_SCRATCH = ("initial_", "inertial_", "c0_")


def _fields(dtype):
    return [f for f in dtype.names if not f.startswith(_STEP_SCRATCH)]


@pytest.fixture(scope="module")
def ref():
    import oracle.bindings as ob
    if not ob.ref_available():
        pytest.skip("oracle/_ref (the compiled reference harness) is not built")
    ob.load_ref()
    return ob


def test_names_and_lookup():
    assert len(ns.NAMES) == 47 and ns.NAMES[0] == "Empty"
    assert ns.index_of("Soft Body 8x8x8") == ns.index_of("softbody8x8x8") == 42
    assert ns.index_of("NEWTON'S CRADLE") == 40
    assert ns.normalize_name("Sphere Pour 5000 on Cylinders") == "spherepour5000oncylinders"
    with pytest.raises(KeyError):
        ns.index_of("no such scene")


def test_reference_scene_names(ref):
    assert ref.ref_scene_names() == ns.NAMES


@pytest.mark.parametrize("idx", range(1, 47))
def test_registry_scene_equals_reference(ref, idx):
    from avbd import capi
    mine = scenes.build_state(ns.build(idx), capi.Maker())
    theirs = ref.RefWorld(default_params(), ref_scene=idx).state()
    for part in ("bodies", "joints", "springs"):
        a, b = getattr(mine, part), getattr(theirs, part)
        assert len(a) == len(b), part
        for f in _fields(b.dtype):
            np.testing.assert_array_equal(a[f], b[f], err_msg=f"{ns.NAMES[idx]}: {part}.{f}")
    np.testing.assert_array_equal(mine.ignores, theirs.ignores)
    assert mine.next_serial == theirs.next_serial


def _box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    tris = []
    for a, b, c, d in quads:
        tris += [a, b, c, a, c, d]
    return ml.TriMesh(np.asarray(v, np.float32), np.asarray(tris, np.uint32))


@pytest.mark.parametrize("n", [1, 2, 4, 5])
def test_box_mesh_fills_full_grid_with_face_joints(n):
    """An axis box of side n*s - e: an n^3 grid, 3 n^2 (n-1) face joints, no diagonals (every
    diagonal has an occupied face path)."""
    s = 0.25
    L = n * s - 0.01
    sb = ns.SceneBuilder()
    build = ml.build_lattice(sb, _box_mesh((0, 0, 0), (L, L, L)), (0, 0, 0), 1.0, ml.LatticeParams(spacing=s))
    sc = sb.scene("box", -10.0, 1)
    assert len(build.particles) == sc.num_bodies == n ** 3
    assert len(sc.ftype) == 3 * n * n * (n - 1)
    assert (sc.shape == scenes.SHAPE_SPHERE).all()
    np.testing.assert_array_equal(sc.shape_radius, np.float32(s) * np.float32(0.5))
    # particle centres sit at cell centres; creation order z, y, x
    c = build.centers
    k = np.arange(n ** 3)
    expect = np.stack([k % n, (k // n) % n, k // (n * n)], 1).astype(np.float32)
    np.testing.assert_allclose(c, (expect + 0.5) * s, rtol=0, atol=1e-6)
    # anchors meet at the shared face: |ra| = s/2 along one axis, rb = -ra
    np.testing.assert_array_equal(sc.rb, -sc.ra)
    np.testing.assert_allclose(np.abs(sc.ra).sum(1), s / 2, rtol=1e-6)
    # automatic stiffness: 7000 * particle mass, angular 0.2 of it
    r = s / 2
    mass = 4.0 / 3.0 * math.pi * r ** 3 * 12.0
    np.testing.assert_allclose(sc.p0, 7000.0 * mass, rtol=1e-5)
    np.testing.assert_allclose(sc.p1, 0.2 * 7000.0 * mass, rtol=1e-5)


def test_rigid_lattice_uses_infinite_stiffness():
    sb = ns.SceneBuilder()
    p = ml.LatticeParams(spacing=0.25, stiffness_lin=math.inf, stiffness_ang=math.inf)
    ml.build_lattice(sb, _box_mesh((0, 0, 0), (0.49, 0.49, 0.49)), (0, 0, 0), 1.0, p)
    sc = sb.scene("rigid", -10.0, 1)
    assert len(sc.ftype) == 12 and np.isinf(sc.p0).all() and np.isinf(sc.p1).all()


def test_diagonal_joint_only_without_face_path():
    """Two cells touching along an edge only: the (1,1,0) diagonal is the one joint."""
    sb = ns.SceneBuilder()
    # two thin boxes whose cells are (0,0,0) and (1,1,0) of a 2x2x1 grid: a single-triangle
    # sliver per box so the parity fill finds nothing and the surface marks one cell each
    s = 0.25
    v = np.asarray([(0.02, 0.02, 0.02), (0.2, 0.02, 0.02), (0.02, 0.2, 0.02),
                    (0.3, 0.3, 0.02), (0.48, 0.3, 0.02), (0.3, 0.48, 0.02)], np.float32)
    mesh = ml.TriMesh(v, np.asarray([0, 1, 2, 3, 4, 5], np.uint32))
    build = ml.build_lattice(sb, mesh, (0, 0, 0), 1.0, ml.LatticeParams(spacing=s))
    sc = sb.scene("diag", -10.0, 1)
    assert len(build.particles) == 2 and len(sc.ftype) == 1
    np.testing.assert_array_equal(sc.ra[0], np.float32([s / 2, s / 2, 0]))


def test_empty_and_degenerate_meshes():
    sb = ns.SceneBuilder()
    assert ml.build_lattice(sb, ml.TriMesh(np.zeros((0, 3), np.float32), np.zeros(0, np.uint32)), (0, 0, 0)) is None
    # spacing far below the mesh size: more than 2^22 cells is refused
    assert ml.build_lattice(sb, _box_mesh((0, 0, 0), (10, 10, 10)), (0, 0, 0), 1.0, ml.LatticeParams(spacing=0.05)) is None


def test_torus_mesh_and_obj_round_trip():
    t = ml.torus_mesh(1.0, 0.45, 28, 14)
    assert t.verts.shape == (28 * 14, 3) and len(t.tris) == 28 * 14 * 6
    rho = np.hypot(t.verts[:, 0], t.verts[:, 1])
    np.testing.assert_allclose(np.hypot(rho - 1.0, t.verts[:, 2]), 0.45, rtol=1e-5)
    text = "# torus\n" + "".join(f"v {x!r} {y!r} {z!r}\n" for x, y, z in t.verts.tolist())
    text += "".join(f"f {a + 1}/1/1 {b + 1} {c + 1}\n" for a, b, c in t.tris.reshape(-1, 3).tolist())
    back = ml.read_obj(text)
    np.testing.assert_array_equal(back.verts, t.verts)
    np.testing.assert_array_equal(back.tris, t.tris)
    quad = ml.read_obj("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf -4 -3 -2 -1\n")
    np.testing.assert_array_equal(quad.tris, [0, 1, 2, 0, 2, 3])
    with pytest.raises(ValueError):
        ml.read_obj("v 0 0 0\nf 1 2 3\n")


def test_skin_bindings_weights():
    sb = ns.SceneBuilder()
    mesh = _box_mesh((0, 0, 0), (0.74, 0.74, 0.74))
    build = ml.build_lattice(sb, mesh, (0, 0, 0), 1.0, ml.LatticeParams(spacing=0.25))
    binds = ml.skin_bindings(mesh, (0, 0, 0), 1.0, build, 0.25)
    assert len(binds) == 8
    for b in binds:
        assert 1 <= len(b) <= 4
        assert abs(sum(w for _, w, *_ in b) - 1.0) < 1e-5
        # nearest first
        d = [math.dist((0, 0, 0), o) for _, _, *o in b]
        assert d == sorted(d)
