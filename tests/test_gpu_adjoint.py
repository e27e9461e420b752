"""The hand-written adjoint of the step (avbd_adjoint_forward / _backward) on the GPU.

Reference: the oracle's forward-mode run (orc_tangent_run). Its value half performs the FP32
operations of the step, so it follows the GPU's FP32 trajectory bit for bit, branches
included, and its FP64 tangent is the directional derivative of that trajectory. For a loss
L = sum c . pos_final, the adjoint's gradient g must satisfy g . u = dL[u] for directions u
over each parameter kind (initial velocities, angular velocities, masses with moments scaling,
friction coefficients). Bar: 1e-4 relative. Both are FP64 derivatives of the same FP32 path;
they differ only in the point each linearises at (the forward-mode run uses the FP32-rounded
intermediates, the adjoint recomputes them in FP64 from the FP32 state), which the stiff
penalty rows amplify to ~1e-5 over tens of steps.
The tangent itself is pinned against FP64 central differences in test_oracle_adjoint.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ("vel", "ang_vel", "mass", "friction")


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from avbd import capi
    capi.load()
    return capi


def _directions(state, seed):
    rng = np.random.default_rng(seed)
    nb = len(state.bodies)
    dyn = (state.bodies["mass"] > 0).astype(np.float64)
    return {"vel": rng.normal(size=(nb, 3)) * dyn[:, None], "ang_vel": rng.normal(size=(nb, 3)) * dyn[:, None],
            "mass": rng.normal(size=nb) * dyn * state.bodies["mass"], "friction": rng.normal(size=nb) * 0.1}


def _check(gpu, oracle, state, params, steps, every, seed=0, tol=1e-4, label="", terms=False, floor=1e-12):
    """terms: measure the error against the sum of the absolute per-body terms the projection adds
    up (sum |c_i . dpos_i|) instead of against the projection, for losses whose terms cancel.
    floor: the smallest scale the error is measured against (derivatives that vanish, e.g. a
    sphere's spin on a static sphere, are compared in absolute terms)."""
    nb = len(state.bodies)
    rng = np.random.default_rng(seed + 100)
    c = rng.normal(size=(nb, 3)) * (state.bodies["mass"] > 0)[:, None]
    w = gpu.HipWorld.for_state(state, params)
    w.adjoint_forward(steps, every)
    g = w.adjoint_backward(d_pos=c)
    final = w.state()
    w.close()
    u = _directions(state, seed)
    seeds = {"vel": "dvel", "ang_vel": "dangvel", "mass": "dmass", "friction": "dfriction"}
    errs = {}
    for k in KINDS:
        val, tan = oracle.oracle_tangent_run(state, params, steps, **{seeds[k]: u[k]})
        # the forward-mode run and the GPU forward are the same FP32 trajectory
        np.testing.assert_array_equal(val[:, 0:3].astype(np.float32), final.bodies["pos"], err_msg=f"{label} {k}")
        ref = float((c * tan[:, 0:3]).sum())
        got = float((g[k] * u[k]).sum())
        scale = float(np.abs(c * tan[:, 0:3]).sum()) if terms else abs(ref)
        errs[k] = abs(got - ref) / max(scale, floor)
        print(f"{label} {k:8s} adjoint {got: .10e}  forward-mode {ref: .10e}  rel {abs(got - ref) / max(abs(ref), 1e-12):.2e}"
              f"  of sum|terms| {abs(got - ref) / max(float(np.abs(c * tan[:, 0:3]).sum()), 1e-12):.2e}", flush=True)
    for k in KINDS:
        assert errs[k] < tol, f"{label} {k}: rel error {errs[k]:.2e}"


def test_adjoint_free_flight(gpu, oracle):
    from avbd import capi, scenes
    from avbd.state import default_params
    sc = scenes.stacks(1, seed=2)
    st = scenes.build_state(sc, oracle.OracleMaker())
    st.bodies["pos"][0, 2] += 3.0          # the box starts in the air
    st.bodies["vel"][0] = (0.3, -0.2, 1.0)
    st.bodies["ang_vel"][0] = (0.5, 0.1, -0.7)
    _check(gpu, oracle, st, default_params(), 12, 4, label="free flight")


def _settled(oracle, sc, steps):
    from avbd import scenes
    from avbd.state import default_params
    prm = default_params(gravity=sc.gravity)
    o = oracle.OracleWorld(scenes.build_state(sc, oracle.OracleMaker()), prm)
    o.step(steps)
    return o.state(), prm


def test_adjoint_stack(gpu, oracle):
    from avbd import scenes
    st, prm = _settled(oracle, scenes.stack(count=3), 5)
    _check(gpu, oracle, st, prm, 6, 3, label="stack3")


def test_adjoint_pile_with_checkpoints(gpu, oracle):
    from avbd import scenes
    st, prm = _settled(oracle, scenes.pile(24, seed=7), 70)
    assert len(st.manifolds) > 0
    _check(gpu, oracle, st, prm, 20, 5, label="pile24")


def test_adjoint_checkpoint_interval_does_not_change_the_gradient(gpu, oracle):
    from avbd import capi, scenes
    st, prm = _settled(oracle, scenes.pile(24, seed=8), 70)
    nb = len(st.bodies)
    c = np.random.default_rng(3).normal(size=(nb, 3))
    out = []
    # 0: the forward tapes every step and the backward only sweeps
    for every in (1, 4, 16, 0):
        w = gpu.HipWorld.for_state(st, prm)
        w.adjoint_forward(16, every)
        out.append(w.adjoint_backward(d_pos=c))
        w.close()
    for k in KINDS:
        for o in out[1:]:
            np.testing.assert_allclose(o[k], out[0][k], rtol=1e-9, atol=1e-12)


def test_adjoint_springs_free_chain(gpu, oracle):
    """Spring stamps (spring.cpp:27-69) reversed: a free-flying chain of six boxes tied by springs."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.spring_chain(6, seed=12)
    st = scenes.build_state(sc, oracle.OracleMaker())
    st.bodies["prev_vel"] = 0.0
    # the projection cancels (about 0.04 out of terms summing to 4): the bar is on the terms
    _check(gpu, oracle, st, default_params(gravity=sc.gravity), 30, 10, tol=1e-5, label="spring chain", terms=True)


def test_adjoint_springs_and_contacts(gpu, oracle):
    """A chain of eight boxes tied by springs dropped onto a 60-box pile: spring and contact
    stamps in the same bodies' systems, with checkpoints."""
    from avbd import scenes
    sc = scenes.spring_drop(60, 8, seed=13)
    st, prm = _settled(oracle, sc, 40)
    assert len(st.springs) == 7 and len(st.manifolds) > 0
    _check(gpu, oracle, st, prm, 30, 10, tol=1e-5, label="spring drop", terms=True)


@pytest.mark.parametrize("k_ang", [0.0, 1000.0], ids=["ball", "stiff_angular"])
def test_adjoint_joints_free_chain(gpu, oracle, k_ang):
    """Joint stamps, duals and inits (joint.cpp:35-182) reversed: a free-flying chain of six boxes
    joined face to face (infinite linear stiffness; ball joints or angular stiffness 1e3)."""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.joint_chain(6, seed=14, k_ang=k_ang)
    st = scenes.build_state(sc, oracle.OracleMaker())
    st.bodies["prev_vel"] = 0.0
    _check(gpu, oracle, st, default_params(gravity=sc.gravity), 30, 10, tol=1e-5, label=f"joint chain {k_ang:g}",
           terms=True)


def test_adjoint_joints_and_contacts(gpu, oracle):
    """A chain of eight ball-jointed boxes dropped onto a 60-box pile: joint and contact stamps
    in the same bodies' systems, with checkpoints."""
    from avbd import scenes
    sc = scenes.joint_chain(8, seed=15, pile_count=60)
    st, prm = _settled(oracle, sc, 40)
    assert len(st.joints) == 7 and len(st.manifolds) > 0
    _check(gpu, oracle, st, prm, 30, 10, tol=1e-5, label="joint drop", terms=True)


def test_adjoint_soft_body_with_ignore_pairs(gpu, oracle):
    """The reference's Soft Body scene (scenes.h): 192 boxes joined by 432 joints, 648
    IgnoreCollision pairs, resting on the ground: bodies the pairs make constrained are solved
    (solver.cpp:1161-1166, 1190-1191), and the reversed warm start follows that decision."""
    from avbd import named_scenes, scenes
    sc = named_scenes.build("Soft Body")
    st, prm = _settled(oracle, sc, 20)
    # the pairs are list indices; stepping keeps the body list, and the oracle's state omits them
    st.ignores = scenes.build_state(sc, oracle.OracleMaker()).ignores
    assert len(st.ignores) > 0 and len(st.joints) > 0 and len(st.manifolds) > 0
    _check(gpu, oracle, st, prm, 20, 5, tol=1e-5, label="soft body", terms=True)


@pytest.mark.parametrize("kind", ["sphere", "capsule", "cylinder"])
def test_adjoint_round_stacks(gpu, oracle, kind):
    """Columns of three spheres, crosswise capsules or standing cylinders on a static slab, settled
    30 steps: round contact offsets (manifold.cpp:29-90) in the reversed stamps, duals and force
    init, the round narrowphase's pose Jacobian (collide.cpp:435-1181), the spheres' rolling
    friction against the slab (solver.cpp:628-660) and their frictionless manifolds."""
    from avbd import scenes
    st, prm = _settled(oracle, scenes.round_stack(kind, count=3), 30)
    assert len(st.manifolds) == 3
    _check(gpu, oracle, st, prm, 20, 5, tol=1e-5, label=f"{kind} stack", terms=True)


def _round_case_names():
    from avbd import scenes
    return [sc.name for sc in scenes.round_cases()]


@pytest.mark.parametrize("name", _round_case_names())
def test_adjoint_round_branch_cases(gpu, oracle, name):
    """Every branch case of the round narrowphase and contact offsets (thin static tops, swept
    spheres whose Jacobian includes the velocities, push-out faces, capsule cores, cylinder caps,
    rims, sides and crossings, round pairs), 30 steps from rest with checkpoints every 10; error
    against sum |c_i . dpos_i| with a 1e-6 floor (measured up to 2e-5)."""
    from avbd import scenes
    from avbd.state import default_params
    sc = next(c for c in scenes.round_cases() if c.name == name)
    st = scenes.build_state(sc, oracle.OracleMaker())
    _check(gpu, oracle, st, default_params(gravity=sc.gravity), 30, 10, tol=1e-4, label=name, terms=True,
           floor=1e-6)


def test_adjoint_mixed_pile(gpu, oracle):
    """60 boxes, spheres, capsules and cylinders in four walls, settled 60 steps (every pair kind)."""
    from avbd import scenes
    st, prm = _settled(oracle, scenes.mixed_pile(60, seed=3), 60)
    assert len(st.manifolds) > 60
    _check(gpu, oracle, st, prm, 30, 10, tol=1e-5, label="mixed pile", terms=True)


def test_adjoint_rejects_external_sources(gpu, oracle):
    from avbd import scenes
    from avbd.state import default_params
    st = scenes.build_state(scenes.round_stack("sphere", count=3), gpu.Maker())
    w = gpu.HipWorld.for_state(st, default_params())
    w.set_external_pairs(np.array([[1, 2]], np.int32))
    with pytest.raises(gpu.AvbdError, match="external"):
        w.adjoint_forward(2, 1)
    with pytest.raises(gpu.AvbdError, match="no recorded forward"):
        w.adjoint_backward()
    w.close()


def _com_loss_parts(pos, mass, pile, target):
    w = mass[pile]
    com = (pos[pile] * w[:, None]).sum(0) / w.sum()
    return com, ((com - target) ** 2).sum()


@pytest.mark.parametrize("count,projectiles,seed,steps,tol", [(64, 4, 9, 90, 1e-4), (1024, 64, 5, 120, 1e-3)],
                         ids=["small", "full_C5"])
def test_rollout_autograd_function_matches_forward_mode(gpu, oracle, count, projectiles, seed, steps, tol):
    """torch.autograd.Function over the C ABI: the projectile workload, loss = |COM(block) - target|^2
    with the COM weighted by the (differentiable) masses; at small size and at the differentiable
    benchmark configuration itself (1024-box block + 64 projectiles, 120 steps, checkpoints every 10)."""
    import torch
    from avbd import adjoint, scenes
    from avbd.state import default_params
    sc = scenes.projectile_pile(count=count, projectiles=projectiles, seed=seed)
    st, pile, proj, target = scenes.projectile_state(sc, gpu.Maker(), seed=seed)
    prm = default_params(gravity=sc.gravity)
    print(f"rollout {count}+{projectiles} boxes, {steps} steps", flush=True)
    sim = adjoint.Rollout(st, prm, steps, checkpoint_every=10)
    b = st.bodies
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)
    vel, ang, mass, fric = t(b["vel"]), t(b["ang_vel"]), t(b["mass"]), t(b["friction"])
    pos = sim(vel, ang, mass, fric)
    com, loss = _com_loss_parts(pos, mass, torch.as_tensor(pile), torch.as_tensor(target))
    loss.backward()
    np.testing.assert_array_equal(pos.detach().numpy().astype(np.float32), sim.final.bodies["pos"])
    # forward-mode reference of the same FP32 trajectory, plus the loss's explicit mass term
    rng = np.random.default_rng(4)
    nb = len(b)
    com_np = com.detach().numpy()
    dcom = 2.0 * (com_np - target)
    wsum = b["mass"][pile].astype(np.float64).sum()
    dirs = {"dvel": np.zeros((nb, 3)), "dangvel": np.zeros((nb, 3))}
    dirs["dvel"][proj] = rng.normal(size=(len(proj), 3))
    dirs["dangvel"][proj] = rng.normal(size=(len(proj), 3))
    dirs["dmass"] = rng.normal(size=nb) * (b["mass"] > 0)
    dirs["dfriction"] = rng.normal(size=nb) * 0.05
    grads = {"dvel": vel.grad.numpy(), "dangvel": ang.grad.numpy(), "dmass": mass.grad.numpy(),
             "dfriction": fric.grad.numpy()}
    pos_np = pos.detach().numpy()
    for k, u in dirs.items():
        _, tan = oracle.oracle_tangent_run(sim.initial_state(b["vel"], b["ang_vel"], b["mass"], b["friction"]), prm,
                                           steps, **{k: u})
        ref = float(dcom @ ((tan[pile, 0:3] * b["mass"][pile, None]).sum(0) / wsum))
        if k == "dmass":
            ref += float(dcom @ (((pos_np[pile] - com_np) * u[pile, None]).sum(0) / wsum))
        got = float((grads[k] * u).sum())
        rel = abs(got - ref) / max(abs(ref), 1e-12)
        print(f"rollout {k:9s} autograd {got: .10e}  forward-mode {ref: .10e}  rel {rel:.2e}", flush=True)
        assert rel < tol, f"{k}: {rel:.2e}"
    sim.close()


def test_rollout_device_tensors_equal_the_host_path(gpu):
    """The Rollout with CUDA tensors (avbd_set_body_inputs / avbd_get_body_state_f64 /
    avbd_adjoint_rewind, pointers into device memory, no host copies of the tensors) gives the
    host path's results: the same position bits, and the same gradients up to the order in which
    FP64 atomics accumulate them (1e-11 of the largest), on a first call, on a second call with
    other inputs (the rewind to the recorded start), and after a host-path call on the same
    Rollout (which uploads, so the device path uploads the original state again)."""
    import torch
    from avbd import adjoint, scenes
    from avbd.state import default_params
    sc = scenes.projectile_pile(count=64, projectiles=4, seed=9)
    st, pile, proj, target = scenes.projectile_state(sc, gpu.Maker(), seed=9)
    prm = default_params(gravity=sc.gravity)
    b = st.bodies
    rng = np.random.default_rng(2)
    inputs = [(b["vel"], b["ang_vel"], b["mass"], b["friction"]),
              (b["vel"] + rng.normal(size=b["vel"].shape) * 0.1, b["ang_vel"] * 0.5,
               b["mass"] * (1.0 + 0.1 * rng.random(len(b))), b["friction"] * 0.9)]

    # a loss linear in the positions, so that dL/dpos is the same bits on both devices (torch's
    # own reductions round differently on the CPU and the GPU)
    c = rng.normal(size=(len(b), 3))

    def run(sim, xs, dev):
        ts = [torch.tensor(np.asarray(x, np.float64), device=dev, requires_grad=True) for x in xs]
        pos = sim(*ts)
        pos.backward(torch.as_tensor(c, device=dev))
        return [pos.detach().cpu().numpy()] + [t.grad.cpu().numpy() for t in ts]

    host = adjoint.Rollout(st, prm, 60, checkpoint_every=10)
    dev = adjoint.Rollout(st, prm, 60, checkpoint_every=10)
    for k, xs in enumerate(inputs + inputs[:1]):
        if k == 2:
            run(dev, inputs[1], "cpu")   # a host-path call in between
        want, got = run(host, xs, "cpu"), run(dev, xs, "cuda")
        np.testing.assert_array_equal(got[0], want[0], err_msg=f"call {k}: pos")
        for name, a, g in zip(("vel", "ang_vel", "mass", "friction"), want[1:], got[1:]):
            err = np.abs(g - a).max() / max(np.abs(a).max(), 1e-300)
            assert err < 1e-11, f"call {k}: {name} differs by {err:.1e} of the largest"
    np.testing.assert_array_equal(dev.final.bodies["pos"], host.final.bodies["pos"])
    host.close()
    dev.close()


def test_body_state_io_through_host_pointers(gpu):
    """avbd_get_body_state_f64 into host memory equals the downloaded state; avbd_set_body_inputs
    from host memory sets what a download then shows (masses with their moments scaled, a static
    body left static); avbd_adjoint_rewind without a recorded forward is refused."""
    from avbd import scenes
    from avbd.state import default_params
    st = scenes.build_state(scenes.pile(20, seed=3), gpu.Maker())
    w = gpu.HipWorld.for_state(st, default_params())
    w.step(5)
    ref = w.state().bodies
    nb = len(ref)
    out = {k: np.zeros((nb, n)) for k, n in (("pos", 3), ("rot", 4), ("vel", 3), ("ang_vel", 3))}
    w.get_body_state_f64_ptr(*[out[k].ctypes.data for k in ("pos", "rot", "vel", "ang_vel")])
    for k in out:
        np.testing.assert_array_equal(out[k], ref[k].astype(np.float64), err_msg=k)
    rng = np.random.default_rng(1)
    v, a = rng.normal(size=(nb, 3)), rng.normal(size=(nb, 3))
    m, f = 1.0 + rng.random(nb), rng.random(nb)
    w.set_body_inputs_ptr(v.ctypes.data, a.ctypes.data, m.ctypes.data, f.ctypes.data)
    got = w.state().bodies
    dyn = st.bodies["mass"] > 0
    assert (~dyn).any() and dyn.any()
    np.testing.assert_array_equal(got["vel"], v.astype(np.float32))
    np.testing.assert_array_equal(got["ang_vel"], a.astype(np.float32))
    np.testing.assert_array_equal(got["friction"], f.astype(np.float32))
    np.testing.assert_array_equal(got["mass"], np.where(dyn, m, 0.0).astype(np.float32))
    ratio = st.bodies["moment"][dyn].astype(np.float64) / st.bodies["mass"][dyn, None].astype(np.float64)
    np.testing.assert_array_equal(got["moment"][dyn], (ratio * m[dyn, None]).astype(np.float32))
    np.testing.assert_array_equal(got["moment"][~dyn], 0.0)
    with pytest.raises(gpu.AvbdError, match="no recorded forward"):
        w.adjoint_rewind()
    w.close()


def test_adjoint_against_frozen_branch_central_differences(gpu, oracle):
    """The GPU adjoint against FP64 central differences (h = 1e-4) of the frozen-branch run
    (orc_shadow_run: the FP32 trajectory's contact sets, clamps and stick decisions, FP64
    arithmetic), on a 30-box pile falling and settling over 120 steps with checkpoints every 10:
    the directional derivatives of L = sum c . pos_final agree for initial velocities, angular
    velocities, masses and friction within 1e-3 of sum |c_i . dpos_i| (the size of the terms the
    projection sums; measured 3e-5 to 3e-4) and within 3e-3 of the projection itself (measured
    4e-4 to 1.5e-3). The adjoint is the derivative of the FP32 trajectory and the differences are
    the derivative of the FP64 map along the same branches: they differ by the FP32 rounding of
    the linearisation point, about 1e-3 after 120 steps, while the adjoint agrees with the FP32
    forward-mode tangent to about 1e-6. (On the 1024+64 projectile workload no finite
    difference converges: the FP32 and FP64 trajectories part by metres and the frozen-branch
    FP64 run leaves the FP32 path at the impacts; that workload is checked against the forward-mode
    reference above.)"""
    from avbd import scenes
    from avbd.state import default_params
    sc = scenes.pile(30, seed=5)
    st = scenes.build_state(sc, gpu.Maker())
    prm = default_params(gravity=sc.gravity)
    steps = 120
    nb = len(st.bodies)
    dyn = st.bodies["mass"] > 0
    rng = np.random.default_rng(7)
    c = rng.normal(size=(nb, 3)) * dyn[:, None]
    w = gpu.HipWorld.for_state(st, prm)
    w.adjoint_forward(steps, 10)
    g = w.adjoint_backward(d_pos=c)
    w.close()
    u = _directions(st, 3)
    keys = {"vel": "vel", "ang_vel": "angvel", "mass": "mass", "friction": "friction"}
    base = {"vel": st.bodies["vel"], "ang_vel": st.bodies["ang_vel"], "mass": st.bodies["mass"],
            "friction": st.bodies["friction"]}
    h = 1e-5
    for k in KINDS:
        x0 = base[k].astype(np.float64)
        p, _ = oracle.oracle_shadow_run(st, prm, steps, **{keys[k]: x0 + h * u[k]})
        m, _ = oracle.oracle_shadow_run(st, prm, steps, **{keys[k]: x0 - h * u[k]})
        terms = c * (p[:, 0:3] - m[:, 0:3]) / (2 * h)
        fd = float(terms.sum())
        got = float((g[k] * u[k]).sum())
        rel = abs(got - fd) / max(abs(fd), 1e-12)
        rel_terms = abs(got - fd) / max(float(np.abs(terms).sum()), 1e-12)
        print(f"pile30 frozen-branch FD {k:8s} adjoint {got: .10e}  FD64 {fd: .10e}  rel {rel:.2e}  "
              f"rel to sum|terms| {rel_terms:.2e}", flush=True)
        assert rel_terms < 1e-3, f"{k}: adjoint vs frozen-branch FD {rel_terms:.2e} of sum|terms|"
        assert rel < 3e-3, f"{k}: adjoint vs frozen-branch FD {rel:.2e}"
