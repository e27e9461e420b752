/*
 * avbd_hip.h -- C ABI of the MI355X (gfx950) AVBD rigid-body step.
 *
 * This is the drop-in boundary. It replaces, for the AVBD step, the reference's
 * PhysicsBackend implementations:
 *   - struct PhysicsBackend { name(); step(Solver&); }            solver.h:410-415
 *   - makeCpuReferencePhysicsBackend()                             solver.h:417, solver.cpp:671-674
 *   - makeWebGpuAvbdPhysicsBackend(WebGpuDevice*)                  avbd_gpu_solver.h:15
 *   - AvbdGpuSolverStats / avbdGpuSolverStats()                    avbd_gpu_solver.h:19-41
 * A C++ PhysicsBackend adapter over these entry points lives in
 * avbd-demo3d_amd/host/hip_avbd_backend.{h,cpp} (registered as "hip-avbd").
 *
 * Conventions
 *   - Every entry point returns an int status: AVBD_OK (0) or a negative error
 *     code. Nothing throws across the ABI. avbd_last_error() returns a
 *     thread-local message for the most recent failure on the calling thread.
 *   - The caller owns every host array passed in; the library owns all device
 *     buffers. State stays resident in HBM between calls.
 *   - Body arrays are in the solver's body-list order (solver.h:357: the list
 *     head, i.e. the most recently created body, is index 0). Joint/spring/
 *     manifold body references are body-list indices; -1 is the world anchor
 *     (a joint whose bodyA is null, joint.cpp:28).
 *   - Forces carry a creation serial. A body's constraint order (the order in
 *     which the reference accumulates its Hessian, force.cpp:25-36 prepends) is
 *     descending serial.
 *   - One handle drives one HIP stream. Calls on one handle must be serialised
 *     by the caller; distinct handles may run concurrently.
 */
#ifndef AVBD_HIP_H
#define AVBD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AVBD_MAX_CONTACTS 8

enum avbd_status {
    AVBD_OK = 0,
    AVBD_ERR_INVALID_ARGUMENT = -1,
    AVBD_ERR_CAPACITY = -2,      /* a buffer sized at avbd_create would overflow */
    AVBD_ERR_HIP = -3,           /* a HIP runtime call failed */
    AVBD_ERR_NO_DEVICE = -4,
    AVBD_ERR_UNSUPPORTED = -5,   /* feature outside the step (shape code; adjoint with external sources) */
    AVBD_ERR_INJECTED = -6       /* avbd_debug_inject_fault armed this failure */
};

/* RigidShapeType (solver.h:47-53). Boxes are Rigid::Rigid (rigid.cpp:24-42); the round
 * shapes are Rigid::makeSphere / makeCapsule / makeCylinder (rigid.cpp:44-86), whose
 * core segment runs along the body's local z axis. */
enum avbd_shape { AVBD_SHAPE_BOX = 0, AVBD_SHAPE_SPHERE = 1, AVBD_SHAPE_CAPSULE = 2, AVBD_SHAPE_CYLINDER = 3 };

/* Solver parameters: Solver::dt/gravity/iterations/alpha/betaLin/betaAng/gamma
 * (solver.h:348-355), defaults in Solver::defaultParams (solver.cpp:909-932).
 * Gravity acts along -z (the reference is Z-up, solver.cpp:1141). */
typedef struct avbd_params {
    float dt;
    float gravity;
    int32_t iterations;
    float alpha;
    float beta_lin;
    float beta_ang;
    float gamma;
} avbd_params;

/* One rigid body: the Rigid fields the step reads or writes (solver.h:197-228).
 * Quaternions are (x, y, z, w). */
typedef struct avbd_body {
    float size[3];          /* full box widths */
    float mass;             /* <= 0: static */
    float moment[3];        /* principal moments of inertia */
    float friction;
    float radius;           /* bounding-sphere radius */
    float pos[3];           /* positionLin */
    float rot[4];           /* positionAng */
    float initial_lin[3];   /* x- (solver.cpp:1151-1152) */
    float initial_ang[4];
    float inertial_lin[3];  /* x~ (solver.cpp:1138-1141) */
    float inertial_ang[4];
    float vel[3];           /* velocityLin */
    float ang_vel[3];       /* velocityAng */
    float prev_vel[3];      /* prevVelocityLin */
    float shape_radius;     /* RigidShape::radius: sphere/capsule/cylinder radius (box: bounding radius) */
    float half_length;      /* RigidShape::halfLength: capsule/cylinder core half length (0 otherwise) */
    int32_t shape;          /* enum avbd_shape */
    int32_t dense_id;       /* caller's id, carried through unchanged */
} avbd_body;

/* One contact point of a manifold: Manifold::Contact (solver.h:301-310). */
typedef struct avbd_contact {
    int32_t feature;        /* FeaturePair::key */
    float r_a[3];           /* contact offset in A's local frame */
    float r_b[3];           /* contact offset in B's local frame */
    float c0[3];            /* constraint value at x- plus margin */
    float penalty[3];
    float lambda[3];
    int32_t stick;
} avbd_contact;

/* A box-box contact manifold: Manifold (solver.h:284-318). */
typedef struct avbd_manifold {
    int32_t body_a;
    int32_t body_b;
    int64_t serial;
    int32_t num_contacts;
    float friction;
    float basis[9];         /* row 0 normal (B to A), rows 1-2 tangents */
    avbd_contact contacts[AVBD_MAX_CONTACTS];
} avbd_manifold;

/* Ball-socket joint with optional angular row and fracture: Joint (solver.h:251-264). */
/* One entry of the external contact table (ExternalManifoldContact, solver.h:336-343): a body pair
 * by handle body index (-1: none), its basis (row 0 the normal, B to A) and its contacts. */
typedef struct avbd_external_contact {
    int32_t body_a;
    int32_t body_b;
    int32_t num_contacts;   /* the first min(num_contacts, 8) contacts are used */
    int32_t pad_;
    float basis[9];
    avbd_contact contacts[AVBD_MAX_CONTACTS];
} avbd_external_contact;

typedef struct avbd_joint {
    int32_t body_a;         /* -1: anchored to the world point r_a */
    int32_t body_b;
    int64_t serial;
    float r_a[3];
    float r_b[3];
    float c0_lin[3];
    float c0_ang[3];
    float penalty_lin[3];
    float penalty_ang[3];
    float lambda_lin[3];
    float lambda_ang[3];
    float stiffness_lin;    /* INFINITY: hard constraint */
    float stiffness_ang;
    float fracture;
    float torque_arm;
    int32_t broken;
    int32_t pad_;
} avbd_joint;

/* Penalty spring: Spring (solver.h:267-278). */
typedef struct avbd_spring {
    int32_t body_a;
    int32_t body_b;
    int64_t serial;
    float r_a[3];
    float r_b[3];
    float rest;
    float stiffness;
} avbd_spring;

/* Capacities of the device buffers; fixed for the handle's lifetime. */
typedef struct avbd_config {
    avbd_params params;
    int32_t max_bodies;
    int32_t max_joints;
    int32_t max_springs;
    int32_t max_manifolds;   /* persistent manifolds + new candidates in one step */
    int32_t max_pairs;       /* bounding-sphere hits per broadphase */
    int32_t device;          /* HIP device ordinal */
    int32_t enable_timing;   /* record per-phase hipEvent timings in avbd_stats */
} avbd_config;

/* Per-step counters and timings of the last avbd_step call's final step:
 * the subset of SolverStats (solver.h:121-188) and AvbdGpuSolverStats
 * (avbd_gpu_solver.h:19-39) that the resident GPU step produces. */
typedef struct avbd_stats {
    int32_t body_count;
    int32_t active_body_count;
    int32_t solved_body_count;
    int32_t pair_checks;
    int32_t sphere_hits;
    int32_t constrained_hits;
    int32_t manifolds_created;
    int32_t manifold_count;
    int32_t contact_count;
    int32_t joint_count;
    int32_t spring_count;
    int32_t color_count;
    int32_t coloring_rounds;
    int32_t sweep_axis;
    int32_t axis_fallbacks;      /* steps whose sweep axis needed the exact serial pass */
    int32_t cpu_fallbacks;       /* set by host adapters that fell back to the CPU */
    int64_t steps_taken;
    float broadphase_ms;
    float narrowphase_ms;
    float coloring_ms;
    float iterate_ms;
    float finish_ms;
    float step_ms;
    /* Algorithmic HBM bytes of the last step: the iteration loop alone and the
     * whole step, from the step's own counts (solved bodies, contacts and their
     * dynamic endpoints, joints, springs, candidate pairs, manifolds). */
    double iterate_bytes;
    double step_bytes;
    /* Sums since avbd_create / avbd_reset_stats. The *_ms_total sums need
     * enable_timing and cover the last min(timed_steps, 4096) steps. */
    double iterate_bytes_total;
    double step_bytes_total;
    double iterate_ms_total;
    double step_ms_total;
    int64_t counted_steps;
    int64_t timed_steps;
    /* The iteration loop split into its two kernels: per iteration, one primal
     * launch per colour (k_primal_color) then one dual launch (k_dual).
     * *_bytes_total: algorithmic bytes of all iterations of the counted steps
     * (primal: solved bodies and the stamps of their forces; dual: contacts
     * and joints). *_ms_total and *_launches: HIP-event spans around the first
     * iteration's primal launches and its dual launch of each timed step, and
     * the number of launches those spans cover. */
    double primal_bytes_total;
    double dual_bytes_total;
    double primal_ms_total;
    double dual_ms_total;
    int64_t primal_launches;
    int64_t dual_launches;
    /* Solver iterations the primal spans cover: one per timed step with per-colour launches (the
     * last iteration), all of them with the flow form (avbd_launch_config.dataflow), whose one
     * primal launch per step runs every iteration. */
    int64_t primal_iterations;
} avbd_stats;

typedef struct avbd_handle avbd_handle;

/* Thread-local description of the last failure on this thread ("" if none). */
const char *avbd_last_error(void);

/* Solver::defaultParams (solver.cpp:909-932), with the given gravity. */
int avbd_default_params(avbd_params *out);

/* Host-side record builders: the mass properties and warm-start defaults the
 * reference constructors compute.
 *   avbd_make_box_body  <- Rigid::Rigid   (rigid.cpp:24-42)
 *   avbd_make_round_body <- Rigid::makeSphere / makeCapsule / makeCylinder (rigid.cpp:44-86)
 *   avbd_make_joint     <- Joint::Joint   (joint.cpp:25-33)
 *   avbd_make_spring    <- Spring::Spring (spring.cpp:14-25)
 * Bodies referenced by a joint/spring are passed by pointer (NULL = world). */
int avbd_make_box_body(const float size[3], float density, float friction,
                       const float pos[3], const float rot[4], const float vel[3],
                       const float ang_vel[3], avbd_body *out);
/* shape: AVBD_SHAPE_SPHERE (half_length ignored), _CAPSULE or _CYLINDER. */
int avbd_make_round_body(int32_t shape, float radius, float half_length, float density, float friction,
                         const float pos[3], const float rot[4], const float vel[3],
                         const float ang_vel[3], avbd_body *out);
int avbd_make_joint(const avbd_body *body_a, const avbd_body *body_b,
                    int32_t index_a, int32_t index_b, int64_t serial,
                    const float r_a[3], const float r_b[3], float stiffness_lin,
                    float stiffness_ang, float fracture, avbd_joint *out);
int avbd_make_spring(const avbd_body *body_a, const avbd_body *body_b,
                     int32_t index_a, int32_t index_b, int64_t serial,
                     const float r_a[3], const float r_b[3], float stiffness,
                     float rest, avbd_spring *out);

/* Handle lifetime. The replacement for constructing a PhysicsBackend. */
int avbd_create(const avbd_config *config, avbd_handle **out);
int avbd_destroy(avbd_handle *h);
int avbd_set_params(avbd_handle *h, const avbd_params *params);

/* Replace the whole resident state (bodies in list order). */
int avbd_upload_state(avbd_handle *h,
                      const avbd_body *bodies, int32_t num_bodies,
                      const avbd_joint *joints, int32_t num_joints,
                      const avbd_spring *springs, int32_t num_springs,
                      const avbd_manifold *manifolds, int32_t num_manifolds,
                      int64_t next_serial);
int avbd_state_counts(avbd_handle *h, int32_t *num_bodies, int32_t *num_joints,
                      int32_t *num_springs, int32_t *num_manifolds, int64_t *next_serial);
/* Any output pointer may be NULL to skip that table. Manifolds are returned in
 * ascending serial order (the reverse of the solver's force-list order). */
int avbd_download_state(avbd_handle *h, avbd_body *bodies, avbd_joint *joints,
                        avbd_spring *springs, avbd_manifold *manifolds);

/* Full resident step(s): Solver::stepCpuReference (solver.cpp:1036-1041) with
 * sweep-and-prune broadphase (solver.cpp:488-577), box-box manifolds
 * (manifold.cpp:100-174, collide.cpp:898-955), greedy body coloring
 * (avbd_gpu_solver.cpp:1812-1865), colored primal/dual iterations
 * (solver.cpp:1172-1273) and BDF1 velocities (solver.cpp:1275-1298).
 * Asynchronous on the handle's stream unless avbd_synchronize is called. */
int avbd_step(avbd_handle *h, int32_t num_steps);
int avbd_synchronize(avbd_handle *h);

/* Phased form, as Solver::prepareStep / iteratePrimalDualCpu / finishStep
 * (solver.h:394-401). avbd_step == prepare + iterate + finish. */
int avbd_prepare(avbd_handle *h);
int avbd_iterate(avbd_handle *h);
int avbd_finish(avbd_handle *h);

/* Introspection of the last step (for parity tests and the host adapter). */
int avbd_get_stats(avbd_handle *h, avbd_stats *out);
/* Zero the *_total sums of avbd_stats (synchronises the handle's stream). */
int avbd_reset_stats(avbd_handle *h);
/* Per body (list order): color of the last step, -1 for unsolved bodies. */
int avbd_get_colors(avbd_handle *h, int32_t *colors, int32_t capacity);
/* Bounding-sphere hits of the last broadphase, in sweep emission order, as
 * (index_a, index_b, created) triples; created = 1 when a new manifold was made. */
int avbd_get_hits(avbd_handle *h, int32_t *triples, int32_t capacity, int32_t *count);

/* Broadphase mode: Solver::broadphaseMode (solver.h:39-45, 366). AVBD_BROADPHASE_SAP is sweep
 * and prune (solver.cpp:488-577), the default of a new handle; AVBD_BROADPHASE_GRID is the uniform
 * grid (solver.cpp:284-486), the default of a default-constructed reference Solver
 * (solver.cpp:676-680); AVBD_BROADPHASE_ALL_PAIRS tests every pair in list order
 * (solver.cpp:176-185). The modes emit different pair sequences, hence different manifold creation
 * orders and trajectories; each reproduces its reference mode bit for bit. The grid's cell size is
 * computed on the host from two values read back per step (the reference's own float arithmetic,
 * cbrtf included). Batched worlds (more than one) use sweep and prune only. The first switch to a
 * mode other than sweep and prune allocates the grid's buffers (a 16 MiB cell table). */
enum avbd_broadphase { AVBD_BROADPHASE_ALL_PAIRS = 0, AVBD_BROADPHASE_GRID = 1, AVBD_BROADPHASE_SAP = 2 };
int avbd_set_broadphase(avbd_handle *h, int32_t mode);
int avbd_get_broadphase(avbd_handle *h, int32_t *mode);

/* External pair and contact sources (replace Solver::stepCpuReferenceWithExternalBroadphase and
 * Solver::setExternalManifoldContacts / clearExternalManifoldContacts, solver.h:403-406,
 * solver.cpp:939-974; the state they set is solver.h:361-365).
 *
 * avbd_set_external_pairs: every following step takes its candidate pairs from `pairs` (count
 * (index_a, index_b) pairs, handle body indices) instead of the broadphase mode, as
 * broadphaseExternalPairs does (solver.cpp:579-626): in order, each valid pair (both indices in
 * range and different) is one check and gets the bounding-sphere test; then every contact-table
 * entry whose bodies are valid and in ascending dense-id order (avbd_body.dense_id) is a check and
 * a hit with no test. Each hit is classified like a broadphase hit (a pair already joined by a
 * force, or by a manifold an earlier entry of the same step created, is a constrained hit), so new
 * manifolds are created in entry order. count may be 0 (only contact entries then create
 * manifolds). The list is copied; it stays in force until avbd_clear_external_pairs.
 *
 * avbd_set_external_contacts: the contact table Manifold::initialize consults (manifold.cpp:110-121):
 * a manifold (a, b) whose pair has an entry -- the last entry for (a, b), else the last for (b, a)
 * -- takes the entry's basis and contacts (penalty, lambda and stick included) instead of the
 * narrowphase, then merges with its old contacts by feature key as usual. count 0 clears the table.
 * The table is copied and used by every following step.
 *
 * A single world only (AVBD_ERR_UNSUPPORTED on batched handles); the adjoint rejects handles with
 * an external source. */
int avbd_set_external_pairs(avbd_handle *h, const int32_t *pairs, int32_t count);
int avbd_clear_external_pairs(avbd_handle *h);
int avbd_set_external_contacts(avbd_handle *h, const avbd_external_contact *contacts, int32_t count);

/* IgnoreCollision forces (solver.h:283-290, force.cpp:75-79) of the uploaded
 * state: body pairs (list indices, 2 * count values) the broadphase treats as
 * already joined (Rigid::constrainedTo, rigid.cpp:105-116), so no manifold is
 * made between them; they add nothing to the primal or dual, but a dynamic
 * body holding one counts as constrained (solved every iteration rather than
 * moved to its inertial target, solver.cpp:1161-1166, 1190-1191). Replaces the
 * handle's previous set; avbd_upload_state clears it. */
int avbd_set_ignore_pairs(avbd_handle *h, const int32_t *pairs, int32_t count);

/* ---- AVBS binary snapshots of the resident state ----
 * The viewer stream format of makeSnapshotBinary (viewer_bridge.cpp:796-828),
 * decoded by the viewer's decodeBinarySnapshot (viewer/src/main.js:986-1050):
 * little-endian header { uint32 magic 0x53425641 ("AVBS"), uint32 version 1,
 * uint64 frame, uint32 scene_bytes, uint32 body_count }, the scene name bytes,
 * then 64 bytes per body in ascending dense_id: { uint32 id (dense_id), shape,
 * material (materialIdForBody, viewer_bridge.cpp:475-493: 0 static, 1 box,
 * 3 sphere, 4 cylinder, 5 capsule), dynamic; float pos[3], rot[4], size[3],
 * shape_radius, half_length }. The body records are packed on the GPU from the
 * resident state (no state download). */
#define AVBD_AVBS_MAGIC 0x53425641u
#define AVBD_AVBS_HEADER_BYTES 24
/* Bytes of the snapshot of the current state with a scene name of scene_bytes. */
int64_t avbd_snapshot_avbs_size(avbd_handle *h, int32_t scene_bytes);
/* Enqueues the packing kernel and the copy into out (capacity bytes, at least
 * avbd_snapshot_avbs_size) on the handle's stream and returns the byte count;
 * the body records are in out once the stream reaches this point
 * (avbd_synchronize). Pinned out (hipHostMalloc) makes the copy asynchronous,
 * so a viewer can stream frames while the next steps run. scene may be NULL. */
int64_t avbd_snapshot_avbs(avbd_handle *h, uint64_t frame, const char *scene, void *out, int64_t capacity);
/* Allocation-free form of avbd_snapshot_avbs. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving or streaming
 * loops, harnesses that forbid hipMalloc). avbd_snapshot_avbs allocates its device staging on its
 * first call; this form stages the packed records in a fixed 4 MiB device array of the library
 * instead (one per GPU, held until the process ends) and calls no allocating routine. Limits: the
 * handle holds at most 65536 bodies (more: AVBD_ERR_CAPACITY), and out must be pinned host memory
 * (hipHostMalloc; anything else: AVBD_ERR_INVALID_ARGUMENT). Output and return value are those of
 * avbd_snapshot_avbs. All handles on one GPU share the staging array: calls that overlap on one GPU
 * (different handles or threads, before the earlier call's stream has passed its copy) race. */
int64_t avbd_snapshot_avbs_static(avbd_handle *h, uint64_t frame, const char *scene, void *out, int64_t capacity);

/* ---- Batched worlds ----
 * Many independent solvers in one handle. The reference runs one Solver per
 * world (class Solver, solver.h:345-420, stepped by Solver::step,
 * solver.cpp:934-937); here the worlds' bodies are stored one world after
 * another and every launch of a step covers all of them. World w owns bodies
 * [world_body_offsets[w], world_body_offsets[w + 1]), each world in its own
 * list order. Joint, spring and manifold body indices are batch-wide and must
 * stay inside one world; serials must be unique over the batch (any renumbering
 * that keeps each world's order works). With more than one world, a world holds
 * at most AVBD_MAX_WORLD_BODIES bodies. Stepping a batch gives every world
 * exactly the result of stepping it alone. avbd_upload_state resets the handle
 * to one world. */
#define AVBD_MAX_WORLD_BODIES 2048
int avbd_batch_upload(avbd_handle *h, const int32_t *world_body_offsets, int32_t num_worlds,
                      const avbd_body *bodies, int32_t num_bodies,
                      const avbd_joint *joints, int32_t num_joints,
                      const avbd_spring *springs, int32_t num_springs,
                      const avbd_manifold *manifolds, int32_t num_manifolds,
                      int64_t next_serial);
int avbd_batch_num_worlds(avbd_handle *h, int32_t *num_worlds);

/* ---- Launch shapes and the autotuner ----
 * No counterpart in the reference (its WebGPU lane uses fixed 64-thread
 * workgroups, avbd_gpu_solver.cpp:502-714). Every shape gives bit-identical
 * results; shapes differ only in speed. Defaults: adaptive lanes (0), 64
 * threads, dual 256 threads, no graph, no dataflow, lazy dual. */
typedef struct avbd_launch_config {
    int32_t primal_lanes;    /* lanes per body in the primal kernel: 8, 16 or 32; 0: per body the
                                fewest of 8, 16, 32, 64 that hold its stamps (the flow form then
                                uses 16) */
    int32_t primal_threads;  /* threads per primal block: 64, 128 or 256 */
    int32_t dual_threads;    /* threads per dual block: 64, 128 or 256 */
    int32_t use_graph;       /* 1: each solver iteration is one hipGraph launch */
    int32_t dataflow;        /* 1: the flow form: one primal launch per step runs every iteration
                                and colour, each body starting as soon as the partners it reads
                                have finished the iterations it needs (contact-only worlds with
                                lazy_dual and direct launches; otherwise ignored) */
    int32_t lazy_dual;       /* 1: each contact point's dual update is applied at the start of
                                the next iteration's primal by the first of its bodies to be
                                solved, and one dual launch after the last iteration covers
                                them all (direct launches only; results are identical) */
} avbd_launch_config;
int avbd_set_launch_config(avbd_handle *h, const avbd_launch_config *cfg);
int avbd_get_launch_config(avbd_handle *h, avbd_launch_config *out);
#define AVBD_TUNE_NO_GRAPH 1   /* keep direct launches (per-kernel timing marks stay exact) */
/* Times every launch shape on the resident state for trial_steps steps each and
 * keeps the fastest (primal shapes are timed twice). The trials step on from
 * one another; the state is restored at the end, so tuning changes speed only;
 * stats sums are reset. out may be NULL. */
int avbd_autotune(avbd_handle *h, int32_t trial_steps, int32_t flags, avbd_launch_config *out);

/* ---- Adjoint of the step (reverse-mode gradients) ----
 * No counterpart in the reference (its solver is not differentiable). For
 * single worlds of boxes, spheres, capsules and cylinders with contacts, joints,
 * springs and IgnoreCollision pairs (external pair or contact sources:
 * AVBD_ERR_UNSUPPORTED): avbd_adjoint_forward
 * steps num_steps times keeping
 * the state at every checkpoint_every-th step start in HBM; avbd_adjoint_backward
 * then returns, for a loss L of the final state, dL/d(initial pos, rot, vel,
 * ang_vel) and dL/d(mass) (each body's moments scale with its mass) and
 * dL/d(friction) per body. It replays each checkpoint segment, recomputes every
 * step with its iterations taped, and sweeps the step in reverse. With
 * checkpoint_every = 0 the forward itself tapes every step (tapes and step
 * copies of all num_steps steps stay in HBM, sized by the handle's capacities:
 * about (iterations + 1) * (32 B per body + 256 B per contact slot) per step)
 * and the backward only sweeps, with the same result. The sweep: finishStep,
 * then per iteration the dual update and the primal colours in reverse order
 * (one transposed 6x6 LDL^T solve per body), the warm start and the force init
 * (collide's pose Jacobian included). The gradient is that of the FP32
 * trajectory with its branches (contact set, clamps, stick flags) held fixed,
 * evaluated in FP64. d_* inputs (nb x 3, nb x 4 for rot) may be NULL (zero);
 * g_* outputs (same shapes; nb for mass and friction) may be NULL. The handle
 * is left at the forward's final state. */
int avbd_adjoint_forward(avbd_handle *h, int32_t num_steps, int32_t checkpoint_every);
int avbd_adjoint_backward(avbd_handle *h, const double *d_pos, const double *d_rot,
                          const double *d_vel, const double *d_ang_vel,
                          double *g_pos, double *g_rot, double *g_vel, double *g_ang_vel,
                          double *g_mass, double *g_friction);

/* FP64 body inputs and outputs for differentiable rollouts driven from device
 * tensors (no counterpart in the reference). Pointers may be device memory or
 * host memory (copies use hipMemcpyDefault); each call is done with them when
 * it returns; d_* and g_* of avbd_adjoint_backward may be either kind as well.
 * avbd_set_body_inputs overwrites the resident bodies' linear and angular
 * velocities (nb x 3), masses (nb: a dynamic body's moments scale with its
 * mass by the moment/mass ratios of the state as uploaded; a body uploaded
 * static keeps mass 0) and friction coefficients (nb), each rounded to FP32; a
 * NULL pointer keeps that field. avbd_get_body_state_f64 reads pos (nb x 3),
 * rot (nb x 4, x y z w), vel and ang_vel (nb x 3) as FP64; NULL outputs are
 * skipped. avbd_adjoint_rewind returns the handle to the start of the recorded
 * avbd_adjoint_forward (its first checkpoint: poses, velocities, contacts,
 * joints; masses and friction are left as they are), so that new inputs can
 * be set and the forward run again without an upload; AVBD_ERR_INVALID_ARGUMENT
 * when no forward of the uploaded state is recorded (an upload discards it). */
int avbd_set_body_inputs(avbd_handle *h, const double *vel, const double *ang_vel, const double *mass,
                         const double *friction);
int avbd_get_body_state_f64(avbd_handle *h, double *pos, double *rot, double *vel, double *ang_vel);
int avbd_adjoint_rewind(avbd_handle *h);

/* Test hook: make the next avbd_step on this handle fail with AVBD_ERR_INJECTED
 * before any work is enqueued (code 1), or clear it (code 0). */
int avbd_debug_inject_fault(avbd_handle *h, int32_t code);
/* Test hook (no counterpart in the reference): the number of clip-polygon vertices
 * (1..8, default 8) the box-only narrowphase keeps in LDS. Manifolds whose polygon
 * outgrows it are redone by the full-capacity narrowphase, so results do not change;
 * a small value sends most manifolds down that path. */
int avbd_debug_narrowphase_capacity(avbd_handle *h, int32_t vertices);

/* sizeof of the ABI records (0 body, 1 contact, 2 manifold, 3 joint, 4 spring,
 * 5 params, 6 launch config, 7 stats, 8 config, 9 external contact), so bindings can verify their
 * layouts; -1 for other values. */
int avbd_sizeof(int which);

#ifdef __cplusplus
}
#endif

#endif /* AVBD_HIP_H */
