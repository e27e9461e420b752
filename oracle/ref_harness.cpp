/*
 * ref_harness.cpp -- drives the compiled reference solver (oracle/_ref).
 *
 * TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
 * the reference's own solver sources, where they lie under the reference tree
 * (solver.cpp rigid.cpp force.cpp joint.cpp spring.cpp manifold.cpp
 * collide.cpp sim_world.cpp), into oracle/_ref/libavbdref.so. Nothing of the
 * reference is copied into this repository; this file only calls its public
 * API (Solver, Rigid, Joint, Spring, Manifold, solver.h:197-415).
 *
 * It exchanges state through the POD tables of include/avbd_hip.h so that the
 * golden fixtures, the oracle and the HIP library all use one format.
 *
 * Coloured stepping: the reference's GPU lane colours bodies greedily in list
 * order (avbd_gpu_solver.cpp:1812-1865, compiled only with Dawn, so the rule is
 * applied here over the reference's own object graph). ref_step(..., colored=1)
 * runs prepareStep, relinks Rigid::next by (colour, list index), runs
 * iteratePrimalDualCpu, restores the list and runs finishStep. Bodies of one
 * colour share no force, so this is the Gauss-Seidel order a coloured GPU sweep
 * computes.
 */
#include "solver.h"
#include "scenes.h"
#include "../include/avbd_hip.h"
#ifdef AVBD_WITH_HIP_ADAPTER
#include "hip_avbd_backend.h"
#endif

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace
{

struct RefWorld
{
    Solver solver;
    std::vector<int> colors;
    int colorCount = 0;

    RefWorld()
    {
        solver.clear();
        solver.broadphaseMode = BROADPHASE_SWEEP_AND_PRUNE;
    }
};

float3 f3(const float *p) { return float3{p[0], p[1], p[2]}; }
quat q4(const float *p) { return quat{p[0], p[1], p[2], p[3]}; }
void put3(float *o, float3 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
void put4(float *o, quat q) { o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; }

void applyParams(Solver &s, const avbd_params *p)
{
    s.dt = p->dt;
    s.gravity = p->gravity;
    s.iterations = p->iterations;
    s.alpha = p->alpha;
    s.betaLin = p->beta_lin;
    s.betaAng = p->beta_ang;
    s.gamma = p->gamma;
}

std::vector<Rigid *> bodyList(Solver &s)
{
    std::vector<Rigid *> v;
    for (Rigid *b = s.bodies; b; b = b->next)
        v.push_back(b);
    return v;
}

// Forces in creation order (the solver list is newest-first).
std::vector<Force *> forceCreationOrder(Solver &s)
{
    std::vector<Force *> v;
    for (Force *f = s.forces; f; f = f->next)
        v.push_back(f);
    std::reverse(v.begin(), v.end());
    return v;
}

void setBodyState(Rigid *r, const avbd_body &b)
{
    r->size = f3(b.size);
    r->shape.size = r->size;
    r->mass = b.mass;
    r->moment = f3(b.moment);
    r->friction = b.friction;
    r->radius = b.radius;
    r->shape.type = (RigidShapeType)b.shape;
    r->shape.radius = b.shape_radius;
    r->shape.halfLength = b.half_length;
    r->positionLin = f3(b.pos);
    r->positionAng = q4(b.rot);
    r->initialLin = f3(b.initial_lin);
    r->initialAng = q4(b.initial_ang);
    r->inertialLin = f3(b.inertial_lin);
    r->inertialAng = q4(b.inertial_ang);
    r->velocityLin = f3(b.vel);
    r->velocityAng = f3(b.ang_vel);
    r->prevVelocityLin = f3(b.prev_vel);
}

void getBodyState(const Rigid *r, avbd_body &b)
{
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
    b.shape = (int32_t)r->shape.type;
    b.shape_radius = r->shape.radius;
    b.half_length = r->shape.halfLength;
    b.dense_id = (int32_t)r->denseId;
}

// Greedy colouring of solved bodies in list order over the solver's own
// per-body force lists (rule of avbd_gpu_solver.cpp:1812-1865).
void colorBodies(RefWorld &w)
{
    std::vector<Rigid *> bodies = bodyList(w.solver);
    std::unordered_map<Rigid *, int> index;
    for (int i = 0; i < (int)bodies.size(); ++i)
        index[bodies[i]] = i;
    w.colors.assign(bodies.size(), -1);
    int maxColor = -1;
    std::vector<char> used;
    for (int i = 0; i < (int)bodies.size(); ++i)
    {
        Rigid *b = bodies[i];
        if (b->mass <= 0.0f || b->forces == 0)
            continue;
        used.assign((size_t)maxColor + 2, 0);
        for (Force *f = b->forces; f; f = (f->bodyA == b) ? f->nextA : f->nextB)
        {
            if (f->type == SIM_CONSTRAINT_IGNORE_COLLISION)
                continue;
            Rigid *o = f->bodyA == b ? f->bodyB : f->bodyA;
            if (!o)
                continue;
            int c = w.colors[index[o]];
            if (c >= 0 && c < (int)used.size())
                used[c] = 1;
        }
        int c = 0;
        while (c < (int)used.size() && used[c])
            ++c;
        w.colors[i] = c;
        maxColor = std::max(maxColor, c);
    }
    w.colorCount = maxColor + 1;
}

void coloredStep(RefWorld &w)
{
    Solver &s = w.solver;
    s.prepareStep();
    colorBodies(w);
    std::vector<Rigid *> original = bodyList(s);
    std::vector<int> order(original.size());
    for (int i = 0; i < (int)order.size(); ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        int cx = w.colors[x] < 0 ? INT_MAX : w.colors[x];
        int cy = w.colors[y] < 0 ? INT_MAX : w.colors[y];
        return cx < cy;
    });
    auto link = [&](const std::vector<Rigid *> &v) {
        s.bodies = v.empty() ? nullptr : v[0];
        for (size_t k = 0; k < v.size(); ++k)
            v[k]->next = k + 1 < v.size() ? v[k + 1] : nullptr;
    };
    std::vector<Rigid *> colored;
    for (int k : order)
        colored.push_back(original[k]);
    link(colored);
    s.iteratePrimalDualCpu();
    link(original);
    s.finishStep();
}

RefWorld *fromState(const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
                    const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm)
{
    RefWorld *w = new RefWorld();
    applyParams(w->solver, p);
    // bodies: create in reverse list order so that list index i is bodies[i]
    for (int i = nb - 1; i >= 0; --i)
    {
        const avbd_body &b = bodies[i];
        Rigid *r = new Rigid(&w->solver, f3(b.size), b.mass > 0 ? 1.0f : 0.0f, b.friction, f3(b.pos), f3(b.vel));
        setBodyState(r, b);
    }
    std::vector<Rigid *> list = bodyList(w->solver);
    struct Item
    {
        int64_t serial;
        int kind;
        int idx;
    };
    std::vector<Item> items;
    for (int i = 0; i < nj; ++i)
        items.push_back({joints[i].serial, 1, i});
    for (int i = 0; i < ns; ++i)
        items.push_back({springs[i].serial, 2, i});
    for (int i = 0; i < nm; ++i)
        items.push_back({manifolds[i].serial, 0, i});
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) { return a.serial < b.serial; });
    for (const Item &it : items)
    {
        if (it.kind == 1)
        {
            const avbd_joint &j = joints[it.idx];
            Rigid *a = j.body_a >= 0 ? list[j.body_a] : nullptr;
            Joint *o = new Joint(&w->solver, a, list[j.body_b], f3(j.r_a), f3(j.r_b), j.stiffness_lin, j.stiffness_ang, j.fracture);
            o->C0Lin = f3(j.c0_lin);
            o->C0Ang = f3(j.c0_ang);
            o->penaltyLin = f3(j.penalty_lin);
            o->penaltyAng = f3(j.penalty_ang);
            o->lambdaLin = f3(j.lambda_lin);
            o->lambdaAng = f3(j.lambda_ang);
            o->torqueArm = j.torque_arm;
            o->broken = j.broken != 0;
        }
        else if (it.kind == 2)
        {
            const avbd_spring &s = springs[it.idx];
            new Spring(&w->solver, list[s.body_a], list[s.body_b], f3(s.r_a), f3(s.r_b), s.stiffness, s.rest);
        }
        else
        {
            const avbd_manifold &m = manifolds[it.idx];
            Manifold *o = new Manifold(&w->solver, list[m.body_a], list[m.body_b]);
            o->numContacts = m.num_contacts;
            o->friction = m.friction;
            for (int r = 0; r < 3; ++r)
                o->basis[r] = f3(m.basis + 3 * r);
            for (int c = 0; c < 8; ++c)
            {
                const avbd_contact &sc = m.contacts[c];
                Manifold::Contact &oc = o->contacts[c];
                oc.feature.key = sc.feature;
                oc.rA = f3(sc.r_a);
                oc.rB = f3(sc.r_b);
                oc.C0 = f3(sc.c0);
                oc.penalty = f3(sc.penalty);
                oc.lambda = f3(sc.lambda);
                oc.stick = sc.stick != 0;
            }
        }
    }
    w->colors.assign(nb, -1);
    return w;
}

} // namespace

extern "C" {

/* Build a scene through the reference constructors, in creation order.
 * Bodies: size3 density friction pos3 rot4 vel3 angvel3 (rot/angvel set after
 * construction, as a scene would); shape/radius/halfLength (may be NULL: all
 * boxes) select Rigid::makeSphere / makeCapsule / makeCylinder. Forces: type
 * 1 = Joint(a,b,rA,rB,p0=kLin, p1=kAng,p2=fracture), 2 = Spring(a,b,rA,rB,
 * p0=k,p1=rest), 3 = IgnoreCollision(a,b); a = -1 is the world. */
void *ref_create(const avbd_params *p, int nb, const float *size, const float *density, const float *friction,
                 const float *pos, const float *rot, const float *vel, const float *angvel, const int *shape,
                 const float *shapeRadius, const float *halfLength, int nf, const int *ftype, const int *fa,
                 const int *fb, const float *rA, const float *rB, const float *p0, const float *p1, const float *p2)
{
    RefWorld *w = new RefWorld();
    applyParams(w->solver, p);
    std::vector<Rigid *> created(nb);
    for (int i = 0; i < nb; ++i)
    {
        int kind = shape ? shape[i] : AVBD_SHAPE_BOX;
        Rigid *r;
        if (kind == AVBD_SHAPE_SPHERE)
            r = Rigid::makeSphere(&w->solver, shapeRadius[i], density[i], friction[i], f3(pos + 3 * i), f3(vel + 3 * i));
        else if (kind == AVBD_SHAPE_CAPSULE)
            r = Rigid::makeCapsule(&w->solver, shapeRadius[i], halfLength[i], density[i], friction[i], f3(pos + 3 * i),
                                   f3(vel + 3 * i));
        else if (kind == AVBD_SHAPE_CYLINDER)
            r = Rigid::makeCylinder(&w->solver, shapeRadius[i], halfLength[i], density[i], friction[i], f3(pos + 3 * i),
                                    f3(vel + 3 * i));
        else
            r = new Rigid(&w->solver, f3(size + 3 * i), density[i], friction[i], f3(pos + 3 * i), f3(vel + 3 * i));
        r->positionAng = q4(rot + 4 * i);
        r->velocityAng = f3(angvel + 3 * i);
        created[i] = r;
    }
    for (int k = 0; k < nf; ++k)
    {
        Rigid *a = fa[k] >= 0 ? created[fa[k]] : nullptr;
        Rigid *b = created[fb[k]];
        if (ftype[k] == 1)
            new Joint(&w->solver, a, b, f3(rA + 3 * k), f3(rB + 3 * k), p0[k], p1[k], p2[k]);
        else if (ftype[k] == 2)
            new Spring(&w->solver, a, b, f3(rA + 3 * k), f3(rB + 3 * k), p0[k], p1[k]);
        else if (ftype[k] == 3)
            new IgnoreCollision(&w->solver, a, b);
    }
    w->colors.assign(nb, -1);
    return w;
}

void *ref_create_from_state(const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
                            const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm)
{
    return fromState(p, bodies, nb, joints, nj, springs, ns, manifolds, nm);
}

void ref_destroy(void *h) { delete (RefWorld *)h; }

/* The reference's own scenes (scenes.h:1146-1245), built by its scene functions into a
 * solver with the given parameters. */
int ref_scene_count(void) { return sceneCount; }
const char *ref_scene_name(int index) { return index >= 0 && index < sceneCount ? sceneNames[index] : ""; }
void *ref_load_scene(const avbd_params *p, int index)
{
    if (index < 0 || index >= sceneCount)
        return nullptr;
    RefWorld *w = new RefWorld();
    applyParams(w->solver, p);
    scenes[index](&w->solver);
    applyParams(w->solver, p);
    w->colors.assign(bodyList(w->solver).size(), -1);
    return w;
}

/* IgnoreCollision pairs of the solver as list-index pairs, in creation order; returns the count. */
int ref_ignore_pairs(void *h, int32_t *out, int cap)
{
    RefWorld *w = (RefWorld *)h;
    std::vector<Rigid *> list = bodyList(w->solver);
    std::unordered_map<Rigid *, int> index;
    for (int i = 0; i < (int)list.size(); ++i)
        index[list[i]] = i;
    int n = 0;
    for (Force *f : forceCreationOrder(w->solver))
    {
        if (f->type != SIM_CONSTRAINT_IGNORE_COLLISION)
            continue;
        if (n < cap && out)
        {
            out[2 * n] = index[f->bodyA];
            out[2 * n + 1] = index[f->bodyB];
        }
        ++n;
    }
    return n;
}

/* IgnoreCollision forces between list-index body pairs (force.cpp:75-79). */
void ref_add_ignores(void *h, const int32_t *pairs, int n)
{
    RefWorld *w = (RefWorld *)h;
    std::vector<Rigid *> list = bodyList(w->solver);
    for (int k = 0; k < n; ++k)
        new IgnoreCollision(&w->solver, list[pairs[2 * k]], list[pairs[2 * k + 1]]);
}

/* colored = 1: reference step in colour order (see header); 0: Solver::step as is. */
void ref_step(void *h, int n, int colored)
{
    RefWorld *w = (RefWorld *)h;
    for (int i = 0; i < n; ++i)
    {
        if (colored)
            coloredStep(*w);
        else
            w->solver.step();
    }
}

double ref_time_steps(void *h, int n, int colored)
{
    auto t0 = std::chrono::steady_clock::now();
    ref_step(h, n, colored);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void ref_counts(void *h, int32_t *nb, int32_t *nj, int32_t *ns, int32_t *nm)
{
    RefWorld *w = (RefWorld *)h;
    int b = 0, j = 0, s = 0, m = 0;
    for (Rigid *r = w->solver.bodies; r; r = r->next)
        ++b;
    for (Force *f = w->solver.forces; f; f = f->next)
    {
        if (f->type == SIM_CONSTRAINT_JOINT)
            ++j;
        else if (f->type == SIM_CONSTRAINT_SPRING)
            ++s;
        else if (f->type == SIM_CONSTRAINT_MANIFOLD)
            ++m;
    }
    *nb = b;
    *nj = j;
    *ns = s;
    *nm = m;
}

/* Tables in list order (bodies) and creation order (forces); serial = rank in
 * the creation order of the forces alive now. */
void ref_get_state(void *h, avbd_body *bo, avbd_joint *jo, avbd_spring *so, avbd_manifold *mo)
{
    RefWorld *w = (RefWorld *)h;
    std::vector<Rigid *> list = bodyList(w->solver);
    std::unordered_map<const Rigid *, int> index;
    for (int i = 0; i < (int)list.size(); ++i)
        index[list[i]] = i;
    if (bo)
        for (size_t i = 0; i < list.size(); ++i)
            getBodyState(list[i], bo[i]);
    std::vector<Force *> forces = forceCreationOrder(w->solver);
    int nj = 0, ns = 0, nm = 0;
    for (size_t k = 0; k < forces.size(); ++k)
    {
        Force *f = forces[k];
        int ia = f->bodyA ? index[f->bodyA] : -1;
        int ib = f->bodyB ? index[f->bodyB] : -1;
        if (f->type == SIM_CONSTRAINT_JOINT)
        {
            Joint *j = (Joint *)f;
            if (jo)
            {
                avbd_joint &o = jo[nj];
                std::memset(&o, 0, sizeof(o));
                o.body_a = ia;
                o.body_b = ib;
                o.serial = (int64_t)k;
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
            }
            ++nj;
        }
        else if (f->type == SIM_CONSTRAINT_SPRING)
        {
            Spring *s = (Spring *)f;
            if (so)
            {
                avbd_spring &o = so[ns];
                std::memset(&o, 0, sizeof(o));
                o.body_a = ia;
                o.body_b = ib;
                o.serial = (int64_t)k;
                put3(o.r_a, s->rA);
                put3(o.r_b, s->rB);
                o.rest = s->rest;
                o.stiffness = s->stiffness;
            }
            ++ns;
        }
        else if (f->type == SIM_CONSTRAINT_MANIFOLD)
        {
            Manifold *m = (Manifold *)f;
            if (mo)
            {
                avbd_manifold &o = mo[nm];
                std::memset(&o, 0, sizeof(o));
                o.body_a = ia;
                o.body_b = ib;
                o.serial = (int64_t)k;
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
            }
            ++nm;
        }
    }
}

int ref_get_colors(void *h, int32_t *out, int cap)
{
    RefWorld *w = (RefWorld *)h;
    int n = (int)w->colors.size();
    for (int i = 0; i < n && i < cap; ++i)
        out[i] = w->colors[i];
    return n;
}

int ref_color_count(void *h) { return ((RefWorld *)h)->colorCount; }

/* Solver::broadphaseMode (solver.h:39-45): 0 all pairs, 1 uniform grid, 2 sweep and prune (the
 * harness default), or -1 for the mode a default-constructed Solver has (solver.cpp:676-680).
 * Returns the mode now set. */
int ref_set_broadphase(void *h, int mode)
{
    RefWorld *w = (RefWorld *)h;
    if (mode < 0)
    {
        Solver fresh;
        w->solver.broadphaseMode = fresh.broadphaseMode;
    }
    else if (mode < BROADPHASE_COUNT)
        w->solver.broadphaseMode = (BroadphaseMode)mode;
    return (int)w->solver.broadphaseMode;
}

/* n steps with external pair and contact sources, as Solver::stepCpuReferenceWithExternalBroadphase
 * (solver.cpp:939-957) sets them around each step: `pairs` (np (a, b) pairs) and the contact table
 * entries by list index, turned into the bodies' dense ids (an index outside the list becomes
 * INVALID_BODY_ID). mode 0: Solver::stepCpuReference; 1: in colour order like ref_step; 2:
 * Solver::step, the installed PhysicsBackend (ref_use_hip). pairs_on = 0 sets only the contact table
 * (Solver::setExternalManifoldContacts) and runs the world's own broadphase. */
void ref_step_external(void *h, int n, int mode, int pairs_on, const int32_t *pairs, int np,
                       const avbd_external_contact *contacts, int nc)
{
    RefWorld *w = (RefWorld *)h;
    Solver &s = w->solver;
    std::vector<Rigid *> list = bodyList(s);
    auto id = [&](int32_t i) -> BodyId {
        return i >= 0 && i < (int32_t)list.size() ? list[i]->denseId : INVALID_BODY_ID;
    };
    std::vector<BroadphasePair> bp;
    for (int k = 0; k < np; ++k)
    {
        BroadphasePair q{};
        q.bodyA = id(pairs[2 * k]);
        q.bodyB = id(pairs[2 * k + 1]);
        bp.push_back(q);
    }
    std::vector<ExternalManifoldContact> ec;
    for (int k = 0; k < nc; ++k)
    {
        const avbd_external_contact &e = contacts[k];
        ExternalManifoldContact x{};
        x.bodyA = id(e.body_a);
        x.bodyB = id(e.body_b);
        x.numContacts = e.num_contacts;
        for (int r = 0; r < 3; ++r)
            x.basis[r] = f3(e.basis + 3 * r);
        for (int c = 0; c < 8; ++c)
        {
            const avbd_contact &sc = e.contacts[c];
            Manifold::Contact &oc = x.contacts[c];
            oc.feature.key = sc.feature;
            oc.rA = f3(sc.r_a);
            oc.rB = f3(sc.r_b);
            oc.C0 = f3(sc.c0);
            oc.penalty = f3(sc.penalty);
            oc.lambda = f3(sc.lambda);
            oc.stick = sc.stick != 0;
        }
        ec.push_back(x);
    }
    for (int i = 0; i < n; ++i)
    {
        if (pairs_on)
        {
            s.setExternalManifoldContacts(ec);
            s.externalBroadphasePairs = bp;
            s.useExternalBroadphasePairs = true;
        }
        else
            s.setExternalManifoldContacts(ec);
        if (mode == 1)
            coloredStep(*w);
        else if (mode == 2)
            s.step();
        else
            s.stepCpuReference();
        s.useExternalBroadphasePairs = false;
        s.externalBroadphasePairs.clear();
        s.clearExternalManifoldContacts();
    }
}

/* The last step's broadphase counters (SolverStats, solver.h:120-130): pair checks, sphere hits,
 * constrained hits, manifolds created. */
void ref_step_counts(void *h, int32_t out[4])
{
    const SolverStats &st = ((RefWorld *)h)->solver.stats;
    out[0] = st.pairChecks;
    out[1] = st.sphereHits;
    out[2] = st.constrainedHits;
    out[3] = st.manifoldsCreated;
}

/* Pair emission of the current state in the world's broadphase mode: a clone runs
 * Solver::benchmarkBroadphaseOnly (solver.cpp:1010-1034), which drops the
 * manifolds and creates one per unconstrained bounding-sphere hit in sweep
 * order. Returns the number of (a, b) pairs written (list indices). */
int ref_broadphase_pairs(void *h, int32_t *out, int cap, int32_t *checks, int32_t *hits)
{
    RefWorld *w = (RefWorld *)h;
    int32_t nb, nj, ns, nm;
    ref_counts(h, &nb, &nj, &ns, &nm);
    std::vector<avbd_body> b(nb);
    std::vector<avbd_joint> j(nj);
    std::vector<avbd_spring> s(ns);
    std::vector<avbd_manifold> m(nm);
    ref_get_state(h, b.data(), j.data(), s.data(), m.data());
    avbd_params p{w->solver.dt, w->solver.gravity, w->solver.iterations, w->solver.alpha, w->solver.betaLin,
                  w->solver.betaAng, w->solver.gamma};
    RefWorld *c = fromState(&p, b.data(), nb, j.data(), nj, s.data(), ns, nullptr, 0);
    c->solver.broadphaseMode = w->solver.broadphaseMode;
    c->solver.benchmarkBroadphaseOnly();
    if (checks)
        *checks = c->solver.stats.pairChecks;
    if (hits)
        *hits = c->solver.stats.sphereHits;
    std::vector<Rigid *> list = bodyList(c->solver);
    std::unordered_map<const Rigid *, int> index;
    for (int i = 0; i < (int)list.size(); ++i)
        index[list[i]] = i;
    int n = 0;
    for (Force *f : forceCreationOrder(c->solver))
    {
        if (f->type != SIM_CONSTRAINT_MANIFOLD)
            continue;
        if (n < cap)
        {
            out[2 * n] = index[f->bodyA];
            out[2 * n + 1] = index[f->bodyB];
        }
        ++n;
    }
    delete c;
    return n;
}

/* Manifold::collide on two box bodies (collide.cpp:1154-1181 -> collideBoxes). */
int ref_collide(const avbd_body *a, const avbd_body *b, avbd_contact *out, float basis[9])
{
    Solver s;
    s.clear();
    Rigid *ra = new Rigid(&s, f3(a->size), 1.0f, 0.5f, f3(a->pos));
    Rigid *rb = new Rigid(&s, f3(b->size), 1.0f, 0.5f, f3(b->pos));
    ra->positionAng = q4(a->rot);
    rb->positionAng = q4(b->rot);
    Manifold::Contact c[8];
    std::memset((void *)c, 0, sizeof(c));
    float3x3 bs{};
    int n = Manifold::collide(ra, rb, c, bs);
    for (int r = 0; r < 3; ++r)
        put3(basis + 3 * r, bs[r]);
    for (int i = 0; i < n; ++i)
    {
        std::memset(&out[i], 0, sizeof(avbd_contact));
        out[i].feature = c[i].feature.key;
        put3(out[i].r_a, c[i].rA);
        put3(out[i].r_b, c[i].rB);
    }
    return n;
}

/* Manifold::collide of two full body records (any shapes; velocities and the solver dt feed
 * the swept sphere test). */
int ref_collide_pair(const avbd_body *a, const avbd_body *b, float dt, avbd_contact *out, float basis[9])
{
    Solver s;
    s.clear();
    s.dt = dt;
    Rigid *ra = new Rigid(&s, f3(a->size), 1.0f, 0.5f, f3(a->pos));
    Rigid *rb = new Rigid(&s, f3(b->size), 1.0f, 0.5f, f3(b->pos));
    setBodyState(ra, *a);
    setBodyState(rb, *b);
    Manifold::Contact c[8];
    std::memset((void *)c, 0, sizeof(c));
    float3x3 bs{};
    int n = Manifold::collide(ra, rb, c, bs);
    for (int r = 0; r < 3; ++r)
        put3(basis + 3 * r, bs[r]);
    for (int i = 0; i < n; ++i)
    {
        std::memset(&out[i], 0, sizeof(avbd_contact));
        out[i].feature = c[i].feature.key;
        put3(out[i].r_a, c[i].rA);
        put3(out[i].r_b, c[i].rB);
    }
    return n;
}

/* Rigid::Rigid for a box: mass properties as the reference computes them. */
void ref_make_box_body(const float size[3], float density, float friction, const float pos[3], const float vel[3],
                       avbd_body *out)
{
    Solver s;
    s.clear();
    Rigid *r = new Rigid(&s, f3(size), density, friction, f3(pos), f3(vel));
    getBodyState(r, *out);
}

/* Rigid::makeSphere / makeCapsule / makeCylinder (rigid.cpp:44-86) record. */
int ref_make_round_body(int shape, float radius, float halfLength, float density, float friction, const float pos[3],
                        const float vel[3], avbd_body *out)
{
    Solver s;
    s.clear();
    Rigid *r;
    if (shape == AVBD_SHAPE_SPHERE)
        r = Rigid::makeSphere(&s, radius, density, friction, f3(pos), f3(vel));
    else if (shape == AVBD_SHAPE_CAPSULE)
        r = Rigid::makeCapsule(&s, radius, halfLength, density, friction, f3(pos), f3(vel));
    else if (shape == AVBD_SHAPE_CYLINDER)
        r = Rigid::makeCylinder(&s, radius, halfLength, density, friction, f3(pos), f3(vel));
    else
        return -1;
    getBodyState(r, *out);
    return 0;
}

/* Joint/Spring constructors' derived fields: torque arm (joint.cpp:32) and
 * automatic rest length (spring.cpp:19-24). */
void ref_joint_spring_derived(const avbd_body *a, const avbd_body *b, const float rA[3], const float rB[3],
                              float *torqueArm, float *rest)
{
    Solver s;
    s.clear();
    Rigid *ra = a ? new Rigid(&s, f3(a->size), 1.0f, 0.5f, f3(a->pos)) : nullptr;
    Rigid *rb = new Rigid(&s, f3(b->size), 1.0f, 0.5f, f3(b->pos));
    if (ra)
        ra->positionAng = q4(a->rot);
    rb->positionAng = q4(b->rot);
    Joint *j = new Joint(&s, ra, rb, f3(rA), f3(rB));
    *torqueArm = j->torqueArm;
    if (ra)
    {
        Spring *sp = new Spring(&s, ra, rb, f3(rA), f3(rB), 1.0f, -1.0f);
        *rest = sp->rest;
    }
    else
        *rest = -1.0f;
}

} // extern "C"

extern "C" {

/* Install the "hip-avbd" PhysicsBackend (avbd-demo3d_amd/host) on this solver, so that
 * ref_step(h, n, 0) runs Solver::step through the GPU drop-in. Returns 0, or -1 when the
 * harness was built without the adapter. */
int ref_use_hip(void *h, int device)
{
#ifdef AVBD_WITH_HIP_ADAPTER
    ((RefWorld *)h)->solver.physicsBackend = makeHipAvbdPhysicsBackend(device);
    return 0;
#else
    (void)h;
    (void)device;
    return -1;
#endif
}

int ref_hip_fallbacks(void)
{
#ifdef AVBD_WITH_HIP_ADAPTER
    return hipAvbdSolverStats().cpuIterateFallbacks;
#else
    return -1;
#endif
}

} // extern "C"
