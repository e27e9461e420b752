// hip_avbd_backend.cpp -- the "hip-avbd" PhysicsBackend (see hip_avbd_backend.h).
//
// One Solver::step = pack the object graph -> avbd_upload_state -> avbd_step(1)
// -> avbd_download_state -> apply. The tables use the ABI's conventions:
// bodies in Solver::bodies order, forces carry their creation rank as serial
// (the solver's force list is newest first, force.cpp:17-22), so the per-body
// force order the GPU reproduces is the one the reference accumulates in.
// New manifolds come back with serials past the uploaded ones and are created
// here in ascending serial, which prepends them exactly as the reference's
// broadphase would (solver.cpp:243-247); manifolds and joints missing from the
// download were deleted by their initialisation (manifold.cpp:173,
// joint.cpp:53) and are deleted here.
#include "hip_avbd_backend.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace
{

HipAvbdSolverStats g_stats{};

using Clock = std::chrono::steady_clock;

float ms_since(Clock::time_point t0) { return std::chrono::duration<float, std::milli>(Clock::now() - t0).count(); }

void put3(float *o, float3 v)
{
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}
void put4(float *o, quat q)
{
    o[0] = q.x;
    o[1] = q.y;
    o[2] = q.z;
    o[3] = q.w;
}
float3 get3(const float *p) { return float3{p[0], p[1], p[2]}; }
quat get4(const float *p) { return quat{p[0], p[1], p[2], p[3]}; }

struct HipAvbdBackend : PhysicsBackend
{
    int device;
    avbd_handle *h = nullptr;
    int capBodies = 0, capJoints = 0, capSprings = 0, capManifolds = 0;
    bool broken = false;
    bool warnedMode = false;

    explicit HipAvbdBackend(int dev) : device(dev) {}
    ~HipAvbdBackend() override
    {
        if (h)
            avbd_destroy(h);
    }

    const char *name() const override { return "HIP AVBD"; }

    bool ensure(const avbd_params &p, int nb, int nj, int ns, int nm)
    {
        if (h && nb <= capBodies && nj <= capJoints && ns <= capSprings && 2 * nm + 12 * nb <= capManifolds)
            return avbd_set_params(h, &p) == AVBD_OK;
        if (h)
        {
            avbd_destroy(h);
            h = nullptr;
        }
        avbd_config cfg{};
        cfg.params = p;
        cfg.max_bodies = capBodies = nb + nb / 2 + 16;
        cfg.max_joints = capJoints = nj + 16;
        cfg.max_springs = capSprings = ns + 16;
        cfg.max_manifolds = capManifolds = 12 * capBodies + 4 * nm + 1024;
        cfg.device = device;
        return avbd_create(&cfg, &h) == AVBD_OK;
    }

    void fallback(Solver &solver, const char *why)
    {
        std::fprintf(stderr, "hip-avbd: %s (%s); stepping the CPU reference\n", why, avbd_last_error());
        g_stats.cpuIterateFallbacks++;
        solver.stepCpuReference();
    }

    // The GPU step runs the Solver's own broadphase mode (sweep and prune, the uniform grid or all
    // pairs; solver.cpp:1077-1082), each with the reference's pair order, or its external pairs and
    // external contact table (solver.h:336-343, 361-365; sent to the handle every step). The
    // GPU-narrowphase pair sink (solver.h:379-383) gives a different contact set, and an external
    // contact list the Solver's lookup does not see (filled without setExternalManifoldContacts, so
    // broadphaseExternalPairs creates manifolds from it that Manifold::initialize then collides) has
    // no counterpart on the handle; those steps run on the CPU reference, counted as fallbacks, as the
    // reference's own GPU lane does for steps it cannot take (avbd_gpu_solver.cpp:2152-2170).
    const char *unsupported(const Solver &solver) const
    {
        if (solver.useExternalBroadphasePairs && !solver.useExternalManifoldContacts &&
            !solver.externalManifoldContacts.empty())
            return "an external contact list not installed by setExternalManifoldContacts";
        if (solver.spherePairSink)
            return "a GPU-narrowphase pair sink";
        if (solver.broadphaseMode != BROADPHASE_SWEEP_AND_PRUNE && solver.broadphaseMode != BROADPHASE_UNIFORM_GRID &&
            solver.broadphaseMode != BROADPHASE_ALL_PAIRS)
            return "an unknown broadphase mode";
        return nullptr;
    }

    // The Solver's external pairs and contact table, by dense id, as handle body indices (an id with
    // no body in the Solver's list becomes -1, which the handle skips as the reference skips an id
    // without a source body, solver.cpp:583-596).
    bool sendExternal(const Solver &solver, const std::vector<Rigid *> &bodies)
    {
        std::unordered_map<BodyId, int> byId;
        if (solver.useExternalBroadphasePairs || solver.useExternalManifoldContacts)
        {
            byId.reserve(bodies.size() * 2);
            for (int i = 0; i < (int)bodies.size(); ++i)
                byId[bodies[i]->denseId] = i;
        }
        auto idx = [&](BodyId id) -> int32_t {
            auto it = byId.find(id);
            return it == byId.end() ? -1 : it->second;
        };
        if (solver.useExternalBroadphasePairs)
        {
            std::vector<int32_t> pairs;
            pairs.reserve(2 * solver.externalBroadphasePairs.size());
            for (const BroadphasePair &q : solver.externalBroadphasePairs)
            {
                pairs.push_back(idx(q.bodyA));
                pairs.push_back(idx(q.bodyB));
            }
            if (avbd_set_external_pairs(h, pairs.data(), (int32_t)(pairs.size() / 2)) != AVBD_OK)
                return false;
        }
        else if (avbd_clear_external_pairs(h) != AVBD_OK)
            return false;
        std::vector<avbd_external_contact> table;
        if (solver.useExternalManifoldContacts)
        {
            table.resize(solver.externalManifoldContacts.size());
            for (size_t k = 0; k < table.size(); ++k)
            {
                const ExternalManifoldContact &e = solver.externalManifoldContacts[k];
                avbd_external_contact &o = table[k];
                std::memset(&o, 0, sizeof(o));
                o.body_a = idx(e.bodyA);
                o.body_b = idx(e.bodyB);
                o.num_contacts = e.numContacts;
                for (int r = 0; r < 3; ++r)
                    put3(o.basis + 3 * r, e.basis[r]);
                for (int c = 0; c < AVBD_MAX_CONTACTS; ++c)
                {
                    const Manifold::Contact &ci = e.contacts[c];
                    avbd_contact &co = o.contacts[c];
                    co.feature = ci.feature.key;
                    put3(co.r_a, ci.rA);
                    put3(co.r_b, ci.rB);
                    put3(co.c0, ci.C0);
                    put3(co.penalty, ci.penalty);
                    put3(co.lambda, ci.lambda);
                    co.stick = ci.stick ? 1 : 0;
                }
            }
        }
        return avbd_set_external_contacts(h, table.data(), (int32_t)table.size()) == AVBD_OK;
    }

    void step(Solver &solver) override
    {
        if (broken)
        {
            g_stats.cpuIterateFallbacks++;
            solver.stepCpuReference();
            return;
        }
        if (const char *why = unsupported(solver))
        {
            if (!warnedMode)
                std::fprintf(stderr, "hip-avbd: the Solver is set to %s; stepping the CPU reference\n", why);
            warnedMode = true;
            g_stats.cpuIterateFallbacks++;
            solver.stepCpuReference();
            return;
        }
        Clock::time_point t0 = Clock::now();
        std::vector<Rigid *> bodies;
        for (Rigid *b = solver.bodies; b; b = b->next)
            bodies.push_back(b);
        std::unordered_map<const Rigid *, int> index;
        index.reserve(bodies.size() * 2);
        for (int i = 0; i < (int)bodies.size(); ++i)
            index[bodies[i]] = i;
        std::vector<Force *> forces;
        for (Force *f = solver.forces; f; f = f->next)
            forces.push_back(f);
        std::vector<Force *> created(forces.rbegin(), forces.rend());

        std::vector<avbd_body> tb(bodies.size());
        for (size_t i = 0; i < bodies.size(); ++i)
        {
            const Rigid *r = bodies[i];
            avbd_body &b = tb[i];
            std::memset(&b, 0, sizeof(b));
            put3(b.size, r->size);
            b.mass = r->mass;
            put3(b.moment, r->moment);
            b.friction = r->friction;
            b.radius = r->radius;
            put3(b.pos, r->positionLin);
            put4(b.rot, r->positionAng);
            put3(b.initial_lin, r->initialLin);
            put4(b.initial_ang, r->initialAng);
            put3(b.inertial_lin, r->inertialLin);
            put4(b.inertial_ang, r->inertialAng);
            put3(b.vel, r->velocityLin);
            put3(b.ang_vel, r->velocityAng);
            put3(b.prev_vel, r->prevVelocityLin);
            // RigidShapeType and avbd_shape share their codes (box, sphere, capsule, cylinder)
            b.shape = (int32_t)r->shape.type;
            b.shape_radius = r->shape.radius;
            b.half_length = r->shape.halfLength;
            b.dense_id = (int32_t)r->denseId;
        }
        std::vector<avbd_joint> tj;
        std::vector<avbd_spring> ts;
        std::vector<avbd_manifold> tm;
        std::vector<int32_t> ti; // IgnoreCollision pairs
        std::unordered_map<int64_t, Force *> bySerial;
        for (size_t k = 0; k < created.size(); ++k)
        {
            Force *f = created[k];
            int64_t serial = (int64_t)k;
            bySerial[serial] = f;
            int ia = f->bodyA ? index[f->bodyA] : -1;
            int ib = f->bodyB ? index[f->bodyB] : -1;
            if (f->type == SIM_CONSTRAINT_JOINT)
            {
                const Joint *j = (const Joint *)f;
                avbd_joint o;
                std::memset(&o, 0, sizeof(o));
                o.body_a = ia;
                o.body_b = ib;
                o.serial = serial;
                put3(o.r_a, j->rA);
                put3(o.r_b, j->rB);
                put3(o.c0_lin, j->C0Lin);
                put3(o.c0_ang, j->C0Ang);
                put3(o.penalty_lin, j->penaltyLin);
                put3(o.penalty_ang, j->penaltyAng);
                put3(o.lambda_lin, j->lambdaLin);
                put3(o.lambda_ang, j->lambdaAng);
                o.stiffness_lin = j->stiffnessLin;
                o.stiffness_ang = j->stiffnessAng;
                o.fracture = j->fracture;
                o.torque_arm = j->torqueArm;
                o.broken = j->broken ? 1 : 0;
                tj.push_back(o);
            }
            else if (f->type == SIM_CONSTRAINT_SPRING)
            {
                const Spring *sp = (const Spring *)f;
                avbd_spring o;
                std::memset(&o, 0, sizeof(o));
                o.body_a = ia;
                o.body_b = ib;
                o.serial = serial;
                put3(o.r_a, sp->rA);
                put3(o.r_b, sp->rB);
                o.rest = sp->rest;
                o.stiffness = sp->stiffness;
                ts.push_back(o);
            }
            else if (f->type == SIM_CONSTRAINT_MANIFOLD)
            {
                const Manifold *m = (const Manifold *)f;
                avbd_manifold o;
                std::memset(&o, 0, sizeof(o));
                o.body_a = ia;
                o.body_b = ib;
                o.serial = serial;
                o.num_contacts = m->numContacts;
                o.friction = m->friction;
                for (int r = 0; r < 3; ++r)
                    put3(o.basis + 3 * r, m->basis[r]);
                for (int c = 0; c < m->numContacts; ++c)
                {
                    const Manifold::Contact &ci = m->contacts[c];
                    avbd_contact &co = o.contacts[c];
                    co.feature = ci.feature.key;
                    put3(co.r_a, ci.rA);
                    put3(co.r_b, ci.rB);
                    put3(co.c0, ci.C0);
                    put3(co.penalty, ci.penalty);
                    put3(co.lambda, ci.lambda);
                    co.stick = ci.stick ? 1 : 0;
                }
                tm.push_back(o);
            }
            else if (f->type == SIM_CONSTRAINT_IGNORE_COLLISION)
            {
                ti.push_back(ia);
                ti.push_back(ib);
            }
            else
                return fallback(solver, "force type without a GPU path");
        }
        avbd_params p{solver.dt, solver.gravity, solver.iterations, solver.alpha,
                      solver.betaLin, solver.betaAng, solver.gamma};
        if (!ensure(p, (int)tb.size(), (int)tj.size(), (int)ts.size(), (int)tm.size()))
        {
            broken = true;
            return fallback(solver, "device unavailable");
        }
        g_stats.packMs = ms_since(t0);
        Clock::time_point t1 = Clock::now();
        int32_t nb = 0, nj = 0, ns = 0, nm = 0;
        int64_t nextSerial = 0;
        if (avbd_upload_state(h, tb.data(), (int)tb.size(), tj.data(), (int)tj.size(), ts.data(), (int)ts.size(),
                              tm.data(), (int)tm.size(), (int64_t)created.size()) != AVBD_OK ||
            avbd_set_ignore_pairs(h, ti.data(), (int32_t)(ti.size() / 2)) != AVBD_OK ||
            avbd_set_broadphase(h, (int32_t)solver.broadphaseMode) != AVBD_OK ||  // enum values shared
            !sendExternal(solver, bodies) || avbd_step(h, 1) != AVBD_OK || avbd_synchronize(h) != AVBD_OK ||
            avbd_state_counts(h, &nb, &nj, &ns, &nm, &nextSerial) != AVBD_OK)
        {
            broken = true;
            return fallback(solver, "GPU step failed");
        }
        std::vector<avbd_body> ob(nb);
        std::vector<avbd_joint> oj(nj);
        std::vector<avbd_manifold> om(nm);
        if (avbd_download_state(h, ob.data(), oj.data(), nullptr, om.data()) != AVBD_OK ||
            avbd_get_stats(h, &g_stats.gpu) != AVBD_OK)
        {
            broken = true;
            return fallback(solver, "download failed");
        }
        g_stats.uploads++;
        g_stats.gpuMs = ms_since(t1);
        Clock::time_point t2 = Clock::now();

        // bodies (same list order)
        for (size_t i = 0; i < bodies.size(); ++i)
        {
            Rigid *r = bodies[i];
            const avbd_body &b = ob[i];
            r->positionLin = get3(b.pos);
            r->positionAng = get4(b.rot);
            r->initialLin = get3(b.initial_lin);
            r->initialAng = get4(b.initial_ang);
            r->inertialLin = get3(b.inertial_lin);
            r->inertialAng = get4(b.inertial_ang);
            r->velocityLin = get3(b.vel);
            r->velocityAng = get3(b.ang_vel);
            r->prevVelocityLin = get3(b.prev_vel);
        }
        // joints: update survivors, delete the ones removed at initialisation
        std::unordered_map<int64_t, const avbd_joint *> jointOut;
        for (const avbd_joint &j : oj)
            jointOut[j.serial] = &j;
        std::unordered_map<int64_t, const avbd_manifold *> manOut;
        for (const avbd_manifold &m : om)
            manOut[m.serial] = &m;
        for (size_t k = 0; k < created.size(); ++k)
        {
            Force *f = created[k];
            if (f->type == SIM_CONSTRAINT_JOINT)
            {
                auto it = jointOut.find((int64_t)k);
                if (it == jointOut.end())
                {
                    delete f;
                    continue;
                }
                Joint *j = (Joint *)f;
                const avbd_joint &o = *it->second;
                j->C0Lin = get3(o.c0_lin);
                j->C0Ang = get3(o.c0_ang);
                j->penaltyLin = get3(o.penalty_lin);
                j->penaltyAng = get3(o.penalty_ang);
                j->lambdaLin = get3(o.lambda_lin);
                j->lambdaAng = get3(o.lambda_ang);
                j->broken = o.broken != 0;
            }
            else if (f->type == SIM_CONSTRAINT_MANIFOLD)
            {
                auto it = manOut.find((int64_t)k);
                if (it == manOut.end())
                {
                    delete f;
                    continue;
                }
                const avbd_manifold &o = *it->second;
                Manifold *m = (Manifold *)f;
                m->numContacts = o.num_contacts;
                m->friction = o.friction;
                for (int r = 0; r < 3; ++r)
                    m->basis[r] = get3(o.basis + 3 * r);
                for (int c = 0; c < o.num_contacts; ++c)
                {
                    const avbd_contact &ci = o.contacts[c];
                    Manifold::Contact &mc = m->contacts[c];
                    mc.feature.key = ci.feature;
                    mc.rA = get3(ci.r_a);
                    mc.rB = get3(ci.r_b);
                    mc.C0 = get3(ci.c0);
                    mc.penalty = get3(ci.penalty);
                    mc.lambda = get3(ci.lambda);
                    mc.stick = ci.stick != 0;
                }
            }
        }
        // manifolds created by this step's broadphase, oldest first (prepending keeps the order)
        for (const avbd_manifold &o : om)
        {
            if (o.serial < (int64_t)created.size())
                continue;
            Manifold *m = new Manifold(&solver, bodies[o.body_a], bodies[o.body_b]);
            m->numContacts = o.num_contacts;
            m->friction = o.friction;
            for (int r = 0; r < 3; ++r)
                m->basis[r] = get3(o.basis + 3 * r);
            for (int c = 0; c < o.num_contacts; ++c)
            {
                const avbd_contact &ci = o.contacts[c];
                Manifold::Contact &mc = m->contacts[c];
                mc.feature.key = ci.feature;
                mc.rA = get3(ci.r_a);
                mc.rB = get3(ci.r_b);
                mc.C0 = get3(ci.c0);
                mc.penalty = get3(ci.penalty);
                mc.lambda = get3(ci.lambda);
                mc.stick = ci.stick != 0;
            }
        }
        solver.world.syncFromLegacy(solver);
        g_stats.applyMs = ms_since(t2);
    }
};

} // namespace

std::unique_ptr<PhysicsBackend> makeHipAvbdPhysicsBackend(int device)
{
    return std::unique_ptr<PhysicsBackend>(new HipAvbdBackend(device));
}

const HipAvbdSolverStats &hipAvbdSolverStats() { return g_stats; }
