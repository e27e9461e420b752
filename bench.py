#!/usr/bin/env python3
"""AVBD step throughput on one MI355X: the 50,000-box pile at 10 iterations.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full resident AVBD step (sweep-and-prune broadphase, box-box
narrowphase and warm start, colouring, 10 primal/dual iterations, velocity
update) over the whole pile, with the state resident in HBM. The timed region
is K steps after W warm-up steps, bracketed by device synchronisation.

Printed JSON line:
  value        dynamic bodies x K / wall seconds (body-steps/s)
  roofline     the dominant kernel, the primal (k_primal_adaptive or k_primal_color:
               one launch per colour per iteration; k_primal_flow: one per
               step): algorithmic bytes per launch (from the step's own
               counts: solved bodies, contact/joint/spring endpoints) over its
               average HIP-event duration, against 8 TB/s HBM; the whole
               iteration loop and k_dual are reported beside it
  cpu_baseline the reference CPU solver (oracle/_ref, its own WorkerPool) on a
               bounded sample of the same workload: the settled state the timed
               window starts from, stepped for ~25 s of CPU time

The pile is first stepped --settle untimed steps (default 900) so that the
timed window is the settled pile (about 240k contacts, 7 colours) and not the
free fall of the first steps; the run fails if the timed window has no colours.
The default window (500 steps, 923-1422) stays inside that regime: a box that
fell through the ground is far enough below it by step 1909 to turn the sweep
axis to z, along which the flat pile's intervals overlap, and from then on each
step's sweep visits ~7.5e8 pairs instead of 31 M (as the reference's does).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "avbd-demo3d_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 PMC summary, if present."""
    try:
        with open(os.path.join(PROFILES, f"{kernel}_pile50k_traffic.txt")) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(state, params, target_s, first_step):
    import oracle.bindings as ob
    n_dyn = state.num_dynamic
    if ob.ref_available():
        world = ob.RefWorld(params, state=state)
        kind = "reference"
        workers = min(7, max((os.cpu_count() or 1) - 1, 0))
        cores = workers + 1
        run = lambda n: world.time_steps(n, colored=False)
        what = "reference Solver::step (oracle/_ref, SAP broadphase)"
    else:
        world = ob.OracleWorld(state, params, order=ob.OracleWorld.ORDER_LIST)
        kind = "port"
        cores = 1
        run = world.time_steps
        what = "CPU oracle (single thread)"
    t1 = run(1)
    n = max(1, min(200, int(target_s / max(t1, 1e-6))))
    t = run(n)
    return {"value": n_dyn * n / t, "unit": "body-steps/s", "cores": cores, "kind": kind,
            "sample": f"{what}: {n} steps of the {len(state.bodies)}-body pile from the settled state the GPU's "
                      f"timed window starts from (step {first_step}; {t:.1f} s; broadphase/force init on {cores} "
                      f"threads, primal/dual loops single-threaded as in the reference)"}


def calibrate_copy(mib):
    """A plain device-to-device copy of known size, run before the bench in the same process so
    that one rocprofv3 PMC pass reads FETCH_SIZE/WRITE_SIZE for both the copy and the step kernels."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copied {mib} MiB x5")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bodies", type=int, default=50000)
    ap.add_argument("--settle", type=int, default=900,
                    help="untimed steps before the warm-up, so that the timed window is the settled pile")
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--launch", default="", metavar="KEY=V,...",
                    help="set launch-shape fields (primal_lanes, primal_threads, dual_threads, use_graph, "
                         "dataflow, lazy_dual) instead of autotuning, e.g. primal_lanes=16,primal_threads=64")
    ap.add_argument("--batch-worlds", type=int, default=4096, help="0 skips the batched measurement")
    ap.add_argument("--batch-launch", default="", metavar="KEY=V,...",
                    help="as --launch, for the batched leg (instead of autotuning it)")
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--batch-warmup", type=int, default=20)
    ap.add_argument("--batch-sample", type=int, default=64,
                    help="worlds stepped one handle at a time for the batched-vs-separate comparison")
    ap.add_argument("--pour-spheres", type=int, default=20000, choices=[0, 5000, 20000],
                    help="the reference's benchmark sphere pour on cylinders; 0 skips it")
    ap.add_argument("--pour-steps", type=int, default=600)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout measurement")
    ap.add_argument("--adjoint-steps", type=int, default=120)
    ap.add_argument("--adjoint-checkpoint", type=int, default=10)
    ap.add_argument("--no-adjoint-check", action="store_true",
                    help="skip the CPU forward-mode and FP64 finite-difference gradient checks")
    args = ap.parse_args()

    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py measures one GPU: the AVBD step here runs on a single MI355X "
              "(independent worlds shard across GPUs without communication). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)

    from avbd import capi, scenes
    from avbd.state import default_params

    sc = scenes.pile(args.bodies)
    params = default_params(gravity=sc.gravity)
    log(f"building {sc.name}: {sc.num_bodies} bodies")
    state = scenes.build_state(sc, capi.Maker())
    world = capi.HipWorld.for_state(state, params, timing=True)
    n_dyn = state.num_dynamic

    t_settle = time.perf_counter()
    done = 0
    while done < args.settle:
        k = min(100, args.settle - done)
        world.step(k)
        done += k
    world.synchronize()
    st = world.stats()
    log(f"settled {args.settle} untimed steps in {time.perf_counter() - t_settle:.2f} s: "
        f"{int(st['manifold_count'])} manifolds, {int(st['contact_count'])} contacts, "
        f"{int(st['color_count'])} colours")
    log(f"warmup {args.warmup} steps")
    world.step(args.warmup)
    world.synchronize()
    launch = world.launch_config()
    if args.launch:
        for kv in args.launch.split(","):
            key, val = kv.split("=")
            launch[key.strip()] = int(val)
        world.set_launch_config(**launch)
        launch = world.launch_config()
        log(f"launch shape set to {launch}")
    elif not args.no_tune:
        # direct launches only: the per-kernel timing marks need them; 20 steps per trial so
        # that the steps right after each trial's state restore do not decide the shape
        launch = world.autotune(trial_steps=20, allow_graph=False)
        log(f"autotuned launch shape {launch}")
        world.step(2)
        world.synchronize()
    torch.cuda.synchronize()
    # the CPU baseline steps this same state (downloaded outside the timed region)
    start_state = None if args.no_cpu_baseline else world.state()
    world.reset_stats()
    # the autotuner restores the state it started from; two steps follow it
    first = args.settle + args.warmup + (0 if args.no_tune or args.launch else 2) + 1
    log(f"timing {args.steps} steps ({first}-{first + args.steps - 1})")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    chunk = 100
    while done < args.steps:
        k = min(chunk, args.steps - done)
        world.step(k)
        done += k
    world.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = world.stats()
    log(f"timed region {wall:.3f} s; last step: {int(st['manifold_count'])} manifolds, "
        f"{int(st['contact_count'])} contacts, {int(st['color_count'])} colours "
        f"({int(st['coloring_rounds'])} rounds), {int(st['sphere_hits'])} hits, {int(st['pair_checks'])} checks; "
        f"phase ms: broadphase {st['broadphase_ms']:.3f} init {st['narrowphase_ms']:.3f} "
        f"colour {st['coloring_ms']:.3f} iterate {st['iterate_ms']:.3f} finish {st['finish_ms']:.3f}")

    if int(st["color_count"]) == 0 or int(st["contact_count"]) == 0:
        raise SystemExit(f"bench: the timed window ({first}-{first + args.steps - 1}) has no coloured bodies "
                         f"({int(st['contact_count'])} contacts): the pile is not in contact; raise --settle")

    steps_counted = max(int(st["counted_steps"]), 1)
    iter_ms = float(st["iterate_ms_total"]) / max(int(st["timed_steps"]), 1)
    iter_bytes = float(st["iterate_bytes_total"]) / steps_counted
    step_bytes = float(st["step_bytes_total"]) / steps_counted
    # dominant kernel: k_primal_color (one launch per colour per iteration)
    iters = int(params["iterations"])
    timed = max(int(st["timed_steps"]), 1)
    pr_launches = max(int(st["primal_launches"]), 1)
    pr_ms = float(st["primal_ms_total"]) / pr_launches
    # the spans cover the last iteration's colour launches, or (flow form) the one launch per step
    # that runs every iteration
    pr_iters = max(int(st["primal_iterations"]), 1)
    flow = pr_iters > timed
    # per-colour launches: k_primal_adaptive (lanes per body by stamp count) or k_primal_color
    pr_kernel = "k_primal_flow" if flow else ("k_primal_adaptive" if launch.get("primal_lanes") == 0
                                              else "k_primal_color")
    pr_bytes = float(st["primal_bytes_total"]) / steps_counted / iters * pr_iters / pr_launches
    achieved = pr_bytes / (pr_ms * 1e-3) / 1e9
    du_launches = max(int(st["dual_launches"]), 1)
    du_ms = float(st["dual_ms_total"]) / du_launches
    # dual launches per step: one per iteration, or (lazy dual, contact-only worlds) one per step
    du_per_step = 1 if launch.get("lazy_dual") else iters
    du_bytes = float(st["dual_bytes_total"]) / steps_counted / du_per_step / (du_launches / timed)
    loop_gbs = iter_bytes / (iter_ms * 1e-3) / 1e9
    log(f"{pr_kernel} {pr_ms * 1e3:.2f} us/launch {pr_bytes / 1e6:.2f} MB/launch -> {achieved:.0f} GB/s; "
        f"k_dual {du_ms * 1e3:.2f} us/launch {du_bytes / 1e6:.2f} MB/launch -> "
        f"{du_bytes / (du_ms * 1e-3) / 1e9:.0f} GB/s; iteration loop {iter_ms:.3f} ms -> {loop_gbs:.0f} GB/s")

    result = {
        "metric": "body-steps/s @50k boxes,10 iters; %HBM peak; batched & adjoint vs plain",
        "value": n_dyn * args.steps / wall,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pile{args.bodies}: {args.bodies} boxes (half-extents U[0.25,1] m, random orientations, "
                        f"v~N(0,0.5) m/s) on a jittered 50x40x25 grid dropping onto ground inside 4 walls; "
                        f"dt=1/60, 10 iterations, friction 0.5, gravity -9.81; {args.settle} untimed settling steps and "
                        f"{args.warmup} warm-up steps, then steps {first}-{first + args.steps - 1} timed "
                        f"(last timed step: {int(st['manifold_count'])} manifolds, {int(st['contact_count'])} "
                        f"contacts, {int(st['color_count'])} colours)",
            "timed_steps": [first, first + args.steps - 1],
            "manifolds": int(st["manifold_count"]),
            "contacts": int(st["contact_count"]),
            "colours": int(st["color_count"]),
            "bodies": int(n_dyn),
            "iterations": int(params["iterations"]),
            "parallelism": "single GPU",
            "step_bytes_per_step": step_bytes,
            "iterate_ms_per_step": iter_ms,
            "step_gbs_algorithmic": step_bytes / (wall / args.steps) / 1e9,
            "launch": launch,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": pr_kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(pr_kernel),
            "bytes_per_launch": pr_bytes,
            "launch_us": pr_ms * 1e3,
            "iteration_loop": {"kernels": ("k_primal_flow (all iterations) + k_dual" if flow else
                                           f"{pr_kernel} x colours + k_dual, x10"), "ms_per_step": iter_ms,
                               "bytes_per_step": iter_bytes, "achieved": loop_gbs,
                               "frac": loop_gbs / HBM_PEAK_GBS},
            "k_dual": {"launch_us": du_ms * 1e3, "bytes_per_launch": du_bytes,
                       "achieved": du_bytes / (du_ms * 1e-3) / 1e9},
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(start_state, params, args.cpu_seconds, first)
    world.close()
    torch.cuda.synchronize()
    if args.batch_worlds > 0:
        result["batched"] = batched(args, params)
    if args.pour_spheres > 0:
        result["pour"] = pour_leg(args)
    if not args.no_adjoint:
        result["adjoint"] = adjoint_leg(args)
    print(json.dumps(result), flush=True)


def adjoint_leg(args):
    """The differentiable rollout: a 1024-box block plus 64 projectiles over 120 steps, loss =
    |COM(block) - target|^2, gradients w.r.t. projectile velocities and angular velocities, all
    masses and friction coefficients; cost against the plain forward, and the gradient checked
    against the CPU forward-mode reference (same FP32 trajectory) and against central differences
    of the frozen-branch FP64 run where those exist (adjoint_checks)."""
    import numpy as np
    import torch
    from avbd import adjoint, capi, scenes
    from avbd.state import default_params
    sc = scenes.projectile_pile(1024, 64, seed=5)
    st, pile, proj, target = scenes.projectile_state(sc, capi.Maker(), seed=5)
    prm = default_params(gravity=sc.gravity)
    steps, every = args.adjoint_steps, args.adjoint_checkpoint
    n_dyn = int((st.bodies["mass"] > 0).sum())
    # plain forward (same state, same steps)
    plain = capi.HipWorld.for_state(st, prm)
    plain.step(steps)
    plain.synchronize()
    times = []
    for _ in range(3):
        plain.upload(st)
        plain.synchronize()
        t0 = time.perf_counter()
        plain.step(steps)
        plain.synchronize()
        times.append(time.perf_counter() - t0)
    plain.close()
    t_plain = min(times)
    sim = adjoint.Rollout(st, prm, steps, checkpoint_every=every)
    b = st.bodies
    # inputs, positions and gradients stay in device tensors (the rollout's device-pointer path)
    dev = torch.device("cuda", 0)
    t = lambda a: torch.tensor(np.asarray(a, np.float64), device=dev, requires_grad=True)
    pile_t, target_t = torch.as_tensor(pile, device=dev), torch.as_tensor(target, device=dev)
    best = None
    for _ in range(3):
        vel, ang, mass, fric = t(b["vel"]), t(b["ang_vel"]), t(b["mass"]), t(b["friction"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pos = sim(vel, ang, mass, fric)
        t1 = time.perf_counter()
        w = mass[pile_t]
        com = (pos[pile_t] * w[:, None]).sum(0) / w.sum()
        loss = ((com - target_t) ** 2).sum()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if best is None or (t2 - t0) < best[0] + best[1]:
            best = (t1 - t0, t2 - t1)
    t_fwd, t_bwd = best
    log(f"adjoint: plain {t_plain * 1e3:.1f} ms, forward with checkpoints {t_fwd * 1e3:.1f} ms, "
        f"backward {t_bwd * 1e3:.1f} ms -> {(t_fwd + t_bwd) / t_plain:.2f}x the plain forward")
    grads = {"vel": vel.grad.cpu().numpy(), "ang_vel": ang.grad.cpu().numpy(), "mass": mass.grad.cpu().numpy(),
             "friction": fric.grad.cpu().numpy()}
    # the same rollout with every step taped in the forward (checkpoint_every = 0: no recompute)
    sim_all = adjoint.Rollout(st, prm, steps, checkpoint_every=0)
    best_all = None
    for _ in range(3):
        va, aa, ma, fa = t(b["vel"]), t(b["ang_vel"]), t(b["mass"]), t(b["friction"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pa = sim_all(va, aa, ma, fa)
        t1 = time.perf_counter()
        wa = ma[pile_t]
        ca = (pa[pile_t] * wa[:, None]).sum(0) / wa.sum()
        ((ca - target_t) ** 2).sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if best_all is None or (t2 - t0) < best_all[0] + best_all[1]:
            best_all = (t1 - t0, t2 - t1)
    ga = {"vel": va.grad.cpu().numpy(), "ang_vel": aa.grad.cpu().numpy(), "mass": ma.grad.cpu().numpy(),
          "friction": fa.grad.cpu().numpy()}
    sim_all.close()
    same = max(float(np.abs(ga[k] - grads[k]).max() / max(np.abs(grads[k]).max(), 1e-300)) for k in grads)
    log(f"adjoint, every step taped in the forward: forward {best_all[0] * 1e3:.1f} ms, backward "
        f"{best_all[1] * 1e3:.1f} ms -> {sum(best_all) / t_plain:.2f}x the plain forward; gradients vs the "
        f"checkpointed run: max rel diff {same:.1e}")
    out = {
        "workload": f"1024-box block + 64 projectiles (masses U[0.5,2], moments scaled, friction U[0.3,0.8]), "
                    f"{steps} steps, dt=1/60, 10 iterations, checkpoints every {every} steps; loss "
                    f"|COM(block) - target|^2",
        "bodies": int(len(b)), "dynamic_bodies": n_dyn, "steps": steps, "checkpoint_every": every,
        "plain_forward_ms": t_plain * 1e3, "forward_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3,
        "cost_vs_plain_forward": (t_fwd + t_bwd) / t_plain, "loss": float(loss.detach()),
        "dtype": "f32 forward, f64 adjoint",
        "tape_every_step": {"forward_ms": best_all[0] * 1e3, "backward_ms": best_all[1] * 1e3,
                            "cost_vs_plain_forward": sum(best_all) / t_plain,
                            "max_rel_diff_vs_checkpointed": same},
    }
    if not args.no_adjoint_check:
        out["checks"] = adjoint_checks(sim, st, prm, steps, pile, proj, target, pos.detach().cpu().numpy(), grads)
    sim.close()
    return out


def adjoint_checks(sim, st, prm, steps, pile, proj, target, pos, grads):
    """Directional derivatives of the loss against the CPU oracle.

    - forward_mode: the oracle's forward-mode run on the same FP32 trajectory (the derivative the
      adjoint computes); the error is given relative to the derivative and relative to the sum of
      the absolute per-body terms it adds up (the projection can cancel: mass, 0.07 out of terms
      near 1).
    - fd_fp64_free: FP64 central differences (h = 1e-4) of the FP64 trajectory, with how far the
      FP32 and FP64 trajectories end apart: through the impacts they part by metres, so these
      differences are of another trajectory and are printed, not compared.
    - frozen-branch run (the FP32 trajectory's contact sets, clamps and stick decisions in FP64
      arithmetic): how far its unperturbed run ends from the FP32 trajectory (at the impacts the
      frozen map leaves it, so no finite difference of it exists here), and on a 30-box pile
      falling and settling for 120 steps, where it does, the GPU adjoint against its central
      differences (h = 1e-5)."""
    import numpy as np
    import oracle.bindings as ob
    b = st.bodies
    nb = len(b)
    rng = np.random.default_rng(11)
    w = b["mass"][pile].astype(np.float64)
    com = (pos[pile] * w[:, None]).sum(0) / w.sum()
    dcom = 2.0 * (com - target)
    dirs = {"vel": np.zeros((nb, 3)), "ang_vel": np.zeros((nb, 3))}
    dirs["vel"][proj] = rng.normal(size=(len(proj), 3))
    dirs["ang_vel"][proj] = rng.normal(size=(len(proj), 3))
    dirs["mass"] = rng.normal(size=nb) * (b["mass"] > 0) * 0.1
    dirs["friction"] = rng.normal(size=nb) * 0.05
    seed = {"vel": "dvel", "ang_vel": "dangvel", "mass": "dmass", "friction": "dfriction"}
    fdkey = {"vel": "vel", "ang_vel": "angvel", "mass": "mass", "friction": "friction"}
    base = {"vel": b["vel"], "ang_vel": b["ang_vel"], "mass": b["mass"], "friction": b["friction"]}
    init = sim.initial_state(b["vel"], b["ang_vel"], b["mass"], b["friction"])

    def loss_of(p, m):
        ww = m[pile]
        c = (p[pile] * ww[:, None]).sum(0) / ww.sum()
        return float(((c - target) ** 2).sum())

    out = {}
    t0 = time.perf_counter()
    p64 = ob.oracle_f64_run(init, prm, steps)
    fr, sh = ob.oracle_shadow_run(init, prm, steps)
    gap = np.abs(fr[:, 0:3] - sh[:, 0:3])
    out["trajectories"] = {
        "fp32_vs_fp64_max_pos_m": float(np.abs(p64[:, 0:3] - pos).max()),
        "frozen_branch_vs_fp32_max_pos_m": float(np.nanmax(gap)) if np.isfinite(gap).any() else None,
        "frozen_branch_finite": bool(np.isfinite(fr).all()),
    }
    log(f"adjoint check trajectories: {out['trajectories']}")
    for k, u in dirs.items():
        got = float((grads[k] * u).sum())
        _, tan = ob.oracle_tangent_run(init, prm, steps, **{seed[k]: u})
        terms = (tan[pile, 0:3] * w[:, None]) @ dcom / w.sum()
        if k == "mass":
            terms = np.concatenate([terms, ((pos[pile] - com) * u[pile, None]) @ dcom / w.sum()])
        ref = float(terms.sum())
        h = 1e-4
        x0 = base[k].astype(np.float64)
        lp, lm = [], []
        for sgn, acc in ((1.0, lp), (-1.0, lm)):
            xs = x0 + sgn * h * u
            r = ob.oracle_f64_run(init, prm, steps, **{fdkey[k]: xs})
            acc.append(loss_of(r[:, 0:3], xs if k == "mass" else b["mass"].astype(np.float64)))
        fd = (lp[0] - lm[0]) / (2 * h)
        out[k] = {"adjoint": got, "forward_mode_fp32_path": ref,
                  "rel_vs_forward_mode": abs(got - ref) / max(abs(ref), 1e-30),
                  "rel_vs_forward_mode_terms": abs(got - ref) / max(float(np.abs(terms).sum()), 1e-30),
                  "fd_fp64_free_h1e-4": fd}
        log(f"adjoint check {k:8s}: adjoint {got: .6e} forward-mode {ref: .6e} "
            f"(rel {out[k]['rel_vs_forward_mode']:.1e}, of sum|terms| {out[k]['rel_vs_forward_mode_terms']:.1e}) "
            f"free FD64 {fd: .6e}")
    out["pile30_frozen_branch_fd"] = pile30_frozen_check()
    out["seconds"] = time.perf_counter() - t0
    return out


def pile30_frozen_check(steps=120, h=1e-5):
    """GPU adjoint of L = sum c . pos_final on a 30-box pile (falls and settles over 120 steps,
    checkpoints every 10) against central differences of the oracle's frozen-branch FP64 run."""
    import numpy as np
    import oracle.bindings as ob
    from avbd import capi, scenes
    from avbd.state import default_params
    sc = scenes.pile(30, seed=5)
    st = scenes.build_state(sc, capi.Maker())
    prm = default_params(gravity=sc.gravity)
    nb = len(st.bodies)
    dyn = (st.bodies["mass"] > 0).astype(np.float64)
    rng = np.random.default_rng(7)
    c = rng.normal(size=(nb, 3)) * dyn[:, None]
    w = capi.HipWorld.for_state(st, prm)
    w.adjoint_forward(steps, 10)
    g = w.adjoint_backward(d_pos=c)
    w.close()
    rng = np.random.default_rng(3)
    u = {"vel": rng.normal(size=(nb, 3)) * dyn[:, None], "ang_vel": rng.normal(size=(nb, 3)) * dyn[:, None],
         "mass": rng.normal(size=nb) * dyn * st.bodies["mass"], "friction": rng.normal(size=nb) * 0.1}
    keys = {"vel": "vel", "ang_vel": "angvel", "mass": "mass", "friction": "friction"}
    base = {"vel": st.bodies["vel"], "ang_vel": st.bodies["ang_vel"], "mass": st.bodies["mass"],
            "friction": st.bodies["friction"]}
    out = {"workload": f"30-box pile, {steps} steps, checkpoints every 10, L = sum c . pos_final, h = {h:g}"}
    for k in ("vel", "ang_vel", "mass", "friction"):
        x0 = base[k].astype(np.float64)
        p, _ = ob.oracle_shadow_run(st, prm, steps, **{keys[k]: x0 + h * u[k]})
        m, _ = ob.oracle_shadow_run(st, prm, steps, **{keys[k]: x0 - h * u[k]})
        terms = c * (p[:, 0:3] - m[:, 0:3]) / (2 * h)
        fd = float(terms.sum())
        got = float((g[k] * u[k]).sum())
        out[k] = {"adjoint": got, "fd": fd, "rel": abs(got - fd) / max(abs(fd), 1e-30),
                  "rel_terms": abs(got - fd) / max(float(np.abs(terms).sum()), 1e-30)}
        log(f"adjoint check pile30 frozen-branch FD {k:8s}: adjoint {got: .6e} FD {fd: .6e} "
            f"(rel {out[k]['rel']:.1e}, of sum|terms| {out[k]['rel_terms']:.1e})")
    return out


def pour_leg(args):
    """The reference's round-shape benchmark: spheres poured on seven cylinders
    (sceneSpherePour{5000,20000}OnCylinders, scenes.h:332-342), 5 warm-up steps then the timed
    steps (viewer/package.json:21-22: --frames 600 --warmup-frames 5), and its gates on the final
    state. The launch shape is tuned on the pour's state at step 300, then the run restarts."""
    import torch
    from avbd import capi, gates, named_scenes, scenes
    from avbd.state import default_params
    sc = named_scenes.build(f"Sphere Pour {args.pour_spheres} on Cylinders")
    gate = scenes.POUR_GATES[args.pour_spheres]
    prm = default_params(gravity=sc.gravity)
    start = scenes.build_state(sc, capi.Maker())
    world = capi.HipWorld.for_state(start, prm)
    world.step(300)
    world.synchronize()
    launch = world.autotune(trial_steps=6, allow_graph=False)
    world.upload(start)
    world.step(5)
    world.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.step(args.pour_steps)
    world.synchronize()
    wall = time.perf_counter() - t0
    st = world.stats()
    final = world.state()
    world.close()
    m = gates.pour_metrics(final)
    ms = wall * 1e3 / args.pour_steps
    n_dyn = m["dynamic_bodies"]
    checks = {"avg_step_ms": ms < gate["avg_ms"],
              "max_sphere_ground_penetration": m.get("max_sphere_ground_penetration", 0.0) <= gate["penetration"],
              "final_max_speed": m["final_max_speed"] <= gate["max_speed"],
              "final_dynamic_max_height": m.get("final_dynamic_max_height", 0.0) <= gate["max_height"]}
    log(f"pour: {args.pour_spheres} spheres, {ms:.3f} ms/step, {int(st['contact_count'])} contacts at the end, "
        f"gates {checks}")
    return {
        "workload": f"{sc.name}: {args.pour_spheres} spheres (r 0.12) + 7 cylinders + ground slab, the "
                    f"reference's scene; dt=1/60, 10 iterations, gravity -10; steps 6-{5 + args.pour_steps} timed",
        "bodies": int(len(final.bodies)),
        "dynamic_bodies": n_dyn,
        "steps": args.pour_steps,
        "ms_per_step": ms,
        "value": n_dyn * args.pour_steps / wall,
        "unit": "body-steps/s",
        "launch": launch,
        "final_contacts": int(st["contact_count"]),
        "metrics": m,
        "reference_gate": gate,
        "gate_passed": checks,
        "gate_note": ("the reference solver's own final state after these steps misses the penetration gate "
                      "(0.171 > 0.08) and this run's final state equals it bit for bit; the gate was set for the "
                      "reference's WebGPU lane") if args.pour_spheres == 20000 else "",
    }


def batched(args, params):
    """4096 independent worlds (body counts log-uniform in [8, 512]) in one handle, against
    the same worlds stepped one handle (one world) at a time."""
    import numpy as np
    import torch
    from avbd import capi, scenes
    log(f"batched: building {args.batch_worlds} worlds")
    worlds = scenes.batch(args.batch_worlds, seed=4)
    maker = capi.Maker()
    states = [scenes.build_state(w, maker) for w in worlds]
    n_bodies = sum(len(s.bodies) for s in states)
    n_dyn = sum(s.num_dynamic for s in states)
    world = capi.HipWorld.for_batch(states, params)
    world.step(2)
    world.synchronize()
    if args.batch_launch:
        launch = world.launch_config()
        for kv in args.batch_launch.split(","):
            key, val = kv.split("=")
            launch[key.strip()] = int(val)
        world.set_launch_config(**launch)
        launch = world.launch_config()
        log(f"batched: {n_bodies} bodies ({n_dyn} dynamic); launch shape set to {launch}")
    else:
        launch = world.autotune(trial_steps=6, allow_graph=True)
        log(f"batched: {n_bodies} bodies ({n_dyn} dynamic); autotuned launch shape {launch}")
    world.step(args.batch_warmup)
    world.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    world.step(args.batch_steps)
    world.synchronize()
    wall = time.perf_counter() - t0
    st = world.stats()
    world.close()
    ms = wall * 1e3 / args.batch_steps
    log(f"batched: {ms:.3f} ms/step, {int(st['color_count'])} colours, {int(st['contact_count'])} contacts")
    # one world at a time: every k-th world on its own handle (same warm-up), stepped in turn
    k = max(1, args.batch_worlds // max(args.batch_sample, 1))
    sample = list(range(0, args.batch_worlds, k))[:args.batch_sample]
    steps_each = 10
    per_world = []
    for w in sample:
        one = capi.HipWorld.for_state(states[w], params)
        one.step(2 + args.batch_warmup)
        one.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps_each):
            one.step(1)
            one.synchronize()
        per_world.append((time.perf_counter() - t1) * 1e3 / steps_each)
        one.close()
    sep_ms = float(np.mean(per_world)) * args.batch_worlds
    log(f"one world at a time: {np.mean(per_world):.3f} ms per world-step over {len(sample)} sampled worlds "
        f"-> {sep_ms:.1f} ms per step of all {args.batch_worlds} worlds")
    # the same window on a fresh handle with the other launch form (iteration graph <-> direct
    # launches), every other launch field as chosen above, so the line reports both
    ms_form = {int(launch["use_graph"]): ms}
    alt = dict(launch, use_graph=1 - int(launch["use_graph"]))
    other = capi.HipWorld.for_batch(states, params)
    other.step(2)
    other.set_launch_config(**alt)
    other.step(args.batch_warmup)
    other.synchronize()
    t2 = time.perf_counter()
    other.step(args.batch_steps)
    other.synchronize()
    ms_form[alt["use_graph"]] = (time.perf_counter() - t2) * 1e3 / args.batch_steps
    other.close()
    forms = {"direct_ms_per_step": ms_form[0], "iteration_graph_ms_per_step": ms_form[1],
             "picked": "iteration_graph" if launch["use_graph"] else "direct"}
    log(f"batched: direct launches {ms_form[0]:.3f} ms/step, iteration graph {ms_form[1]:.3f} ms/step")
    return {
        "workload": f"{args.batch_worlds} worlds, body counts log-uniform in [8, 512] (stack fields, piles, "
                    f"chains); dt=1/60, 10 iterations; steps {args.batch_warmup + 3}-"
                    f"{args.batch_warmup + 2 + args.batch_steps} timed",
        "worlds": args.batch_worlds,
        "bodies": int(n_bodies),
        "dynamic_bodies": int(n_dyn),
        "steps": args.batch_steps,
        "ms_per_step": ms,
        "value": n_dyn * args.batch_steps / wall,
        "unit": "body-steps/s",
        "launch": launch,
        "one_at_a_time": {"sample_worlds": len(sample), "steps_each": steps_each,
                          "ms_per_world_step": float(np.mean(per_world)),
                          "ms_per_step_all_worlds": sep_ms},
        "speedup_vs_one_at_a_time": sep_ms / ms,
        "launch_forms": forms,
    }


if __name__ == "__main__":
    main()
