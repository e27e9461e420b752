/*
 * oracle_capi.cpp -- extern "C" entry points of the CPU oracle (liboracle.so).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py, never by the product path.
 *
 * State goes in and out through the POD records of include/avbd_hip.h, so the
 * oracle, the reference harness (oracle/_ref) and the HIP library exchange the
 * same tables. A world is either FP32 (reproduces the reference arithmetic) or
 * FP64 (same algorithm, double throughout; used to state FP32 errors).
 */
#include "avbd_oracle.hpp"
#include "../include/avbd_hip.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

using namespace avbd_oracle;

namespace
{

template <class R> V3<R> v3(const float *p) { return V3<R>{(R)p[0], (R)p[1], (R)p[2]}; }
template <class R> Q<R> q4(const float *p) { return Q<R>{(R)p[0], (R)p[1], (R)p[2], (R)p[3]}; }
template <class R> void put3(float *o, V3<R> v) { o[0] = (float)v.x; o[1] = (float)v.y; o[2] = (float)v.z; }
template <class R> void put4(float *o, Q<R> q) { o[0] = (float)q.x; o[1] = (float)q.y; o[2] = (float)q.z; o[3] = (float)q.w; }

struct WorldBase
{
    virtual ~WorldBase() {}
    virtual void step(int n) = 0;
    virtual void prepare() = 0;
    virtual void iterate() = 0;
    virtual void finish() = 0;
    virtual void setOrder(int mode) = 0;
    virtual void setBroadphase(int mode) = 0;
    virtual void setIgnores(const int32_t *pairs, int n) = 0;
    virtual void setExternal(int pairsOn, const int32_t *pairs, int np, const avbd_external_contact *contacts,
                             int nc) = 0;
    virtual void counts(int32_t *nb, int32_t *nj, int32_t *ns, int32_t *nm, int64_t *serial) = 0;
    virtual void get(avbd_body *b, avbd_joint *j, avbd_spring *s, avbd_manifold *m) = 0;
    virtual void getBodiesF64(double *out) = 0;
    virtual void getManifoldsF64(double *out) = 0;
    virtual int colors(int32_t *out, int cap) = 0;
    virtual int hits(int32_t *out, int cap) = 0;
    virtual int axis() = 0;
    virtual int pairChecks() = 0;
    virtual int colorCount() = 0;
};

template <class R>
Contact<R> contact_of(const avbd_contact &sc)
{
    Contact<R> ci;
    ci.feature = sc.feature;
    ci.rA = v3<R>(sc.r_a);
    ci.rB = v3<R>(sc.r_b);
    ci.C0 = v3<R>(sc.c0);
    ci.penalty = v3<R>(sc.penalty);
    ci.lambda = v3<R>(sc.lambda);
    ci.stick = sc.stick != 0;
    return ci;
}

template <class R>
struct WorldImpl : WorldBase
{
    World<R> w;

    void step(int n) override
    {
        for (int i = 0; i < n; ++i)
            w.step();
    }
    void prepare() override { w.prepare(); }
    void iterate() override { w.iterate(); }
    void finish() override { w.finish(); }
    void setOrder(int mode) override { w.orderMode = mode; }
    void setBroadphase(int mode) override { w.broadphaseMode = mode; }
    void setExternal(int pairsOn, const int32_t *pairs, int np, const avbd_external_contact *contacts,
                     int nc) override
    {
        w.externalPairsOn = pairsOn != 0;
        w.externalPairs.clear();
        for (int k = 0; k < np; ++k)
            w.externalPairs.push_back({pairs[2 * k], pairs[2 * k + 1]});
        w.externalContacts.clear();
        for (int k = 0; k < nc; ++k)
        {
            const avbd_external_contact &e = contacts[k];
            ExternalContact<R> x{};
            x.a = e.body_a;
            x.b = e.body_b;
            x.n = e.num_contacts;
            for (int r = 0; r < 3; ++r)
                x.basis[r] = v3<R>(e.basis + 3 * r);
            for (int c = 0; c < 8; ++c)
                x.c[c] = contact_of<R>(e.contacts[c]);
            w.externalContacts.push_back(x);
        }
    }
    void setIgnores(const int32_t *pairs, int n) override
    {
        w.ignores.clear();
        for (int k = 0; k < n; ++k)
            w.ignores.push_back({pairs[2 * k], pairs[2 * k + 1]});
    }
    void counts(int32_t *nb, int32_t *nj, int32_t *ns, int32_t *nm, int64_t *serial) override
    {
        if (nb) *nb = (int32_t)w.bodies.size();
        if (nj) *nj = (int32_t)w.joints.size();
        if (ns) *ns = (int32_t)w.springs.size();
        if (nm) *nm = (int32_t)w.manifolds.size();
        if (serial) *serial = w.nextSerial;
    }
    void get(avbd_body *bo, avbd_joint *jo, avbd_spring *so, avbd_manifold *mo) override
    {
        if (bo)
            for (size_t i = 0; i < w.bodies.size(); ++i)
            {
                const Body<R> &b = w.bodies[i];
                avbd_body &o = bo[i];
                std::memset(&o, 0, sizeof(o));
                put3(o.size, b.size);
                o.mass = (float)b.mass;
                put3(o.moment, b.moment);
                o.friction = (float)b.friction;
                o.radius = (float)b.radius;
                put3(o.pos, b.pos);
                put4(o.rot, b.rot);
                put3(o.initial_lin, b.initLin);
                put4(o.initial_ang, b.initAng);
                put3(o.inertial_lin, b.inerLin);
                put4(o.inertial_ang, b.inerAng);
                put3(o.vel, b.vel);
                put3(o.ang_vel, b.angVel);
                put3(o.prev_vel, b.prevVel);
                o.shape_radius = (float)b.shapeRadius;
                o.half_length = (float)b.halfLength;
                o.shape = b.shape;
                o.dense_id = b.denseId;
            }
        if (jo)
            for (size_t i = 0; i < w.joints.size(); ++i)
            {
                const Joint<R> &j = w.joints[i];
                avbd_joint &o = jo[i];
                std::memset(&o, 0, sizeof(o));
                o.body_a = j.a;
                o.body_b = j.b;
                o.serial = j.serial;
                put3(o.r_a, j.rA);
                put3(o.r_b, j.rB);
                put3(o.c0_lin, j.C0Lin);
                put3(o.c0_ang, j.C0Ang);
                put3(o.penalty_lin, j.penLin);
                put3(o.penalty_ang, j.penAng);
                put3(o.lambda_lin, j.lamLin);
                put3(o.lambda_ang, j.lamAng);
                o.stiffness_lin = (float)j.kLin;
                o.stiffness_ang = (float)j.kAng;
                o.fracture = (float)j.fracture;
                o.torque_arm = (float)j.torqueArm;
                o.broken = j.broken ? 1 : 0;
            }
        if (so)
            for (size_t i = 0; i < w.springs.size(); ++i)
            {
                const Spring<R> &s = w.springs[i];
                avbd_spring &o = so[i];
                std::memset(&o, 0, sizeof(o));
                o.body_a = s.a;
                o.body_b = s.b;
                o.serial = s.serial;
                put3(o.r_a, s.rA);
                put3(o.r_b, s.rB);
                o.rest = (float)s.rest;
                o.stiffness = (float)s.k;
            }
        if (mo)
            for (size_t i = 0; i < w.manifolds.size(); ++i)
            {
                const Manifold<R> &m = w.manifolds[i];
                avbd_manifold &o = mo[i];
                std::memset(&o, 0, sizeof(o));
                o.body_a = m.a;
                o.body_b = m.b;
                o.serial = m.serial;
                o.num_contacts = m.n;
                o.friction = (float)m.friction;
                for (int r = 0; r < 3; ++r)
                    put3(o.basis + 3 * r, m.basis[r]);
                for (int c = 0; c < m.n; ++c)
                {
                    const Contact<R> &ci = m.c[c];
                    avbd_contact &co = o.contacts[c];
                    co.feature = ci.feature;
                    put3(co.r_a, ci.rA);
                    put3(co.r_b, ci.rB);
                    put3(co.c0, ci.C0);
                    put3(co.penalty, ci.penalty);
                    put3(co.lambda, ci.lambda);
                    co.stick = ci.stick ? 1 : 0;
                }
            }
    }
    // pos3 rot4 vel3 angvel3 per body
    void getBodiesF64(double *out) override
    {
        for (size_t i = 0; i < w.bodies.size(); ++i)
        {
            const Body<R> &b = w.bodies[i];
            double *o = out + 13 * i;
            o[0] = b.pos.x; o[1] = b.pos.y; o[2] = b.pos.z;
            o[3] = b.rot.x; o[4] = b.rot.y; o[5] = b.rot.z; o[6] = b.rot.w;
            o[7] = b.vel.x; o[8] = b.vel.y; o[9] = b.vel.z;
            o[10] = b.angVel.x; o[11] = b.angVel.y; o[12] = b.angVel.z;
        }
    }
    // per manifold, 8 contacts x (lambda3 penalty3)
    void getManifoldsF64(double *out) override
    {
        for (size_t i = 0; i < w.manifolds.size(); ++i)
        {
            double *o = out + 48 * i;
            for (int c = 0; c < 8; ++c)
            {
                const Contact<R> &ci = w.manifolds[i].c[c];
                bool live = c < w.manifolds[i].n;
                o[6 * c + 0] = live ? (double)ci.lambda.x : 0.0;
                o[6 * c + 1] = live ? (double)ci.lambda.y : 0.0;
                o[6 * c + 2] = live ? (double)ci.lambda.z : 0.0;
                o[6 * c + 3] = live ? (double)ci.penalty.x : 0.0;
                o[6 * c + 4] = live ? (double)ci.penalty.y : 0.0;
                o[6 * c + 5] = live ? (double)ci.penalty.z : 0.0;
            }
        }
    }
    int colors(int32_t *out, int cap) override
    {
        int n = (int)w.colors.size();
        for (int i = 0; i < n && i < cap; ++i)
            out[i] = w.colors[i];
        return n;
    }
    int hits(int32_t *out, int cap) override
    {
        int n = (int)w.hits.size();
        for (int i = 0; i < n && i < cap; ++i)
        {
            out[3 * i] = w.hits[i].a;
            out[3 * i + 1] = w.hits[i].b;
            out[3 * i + 2] = w.hits[i].created;
        }
        return n;
    }
    int axis() override { return w.sweepAxis; }
    int pairChecks() override { return w.pairChecks; }
    int colorCount() override { return w.colorCount; }
};

template <class R>
void fill(World<R> &w, const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
          const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm, int64_t nextSerial)
{
    w.p = Params<R>{(R)p->dt, (R)p->gravity, p->iterations, (R)p->alpha, (R)p->beta_lin, (R)p->beta_ang, (R)p->gamma};
    w.bodies.resize(nb);
    for (int i = 0; i < nb; ++i)
    {
        const avbd_body &s = bodies[i];
        Body<R> &b = w.bodies[i];
        b.size = v3<R>(s.size);
        b.mass = s.mass;
        b.moment = v3<R>(s.moment);
        b.friction = s.friction;
        b.radius = s.radius;
        b.pos = v3<R>(s.pos);
        b.rot = q4<R>(s.rot);
        b.initLin = v3<R>(s.initial_lin);
        b.initAng = q4<R>(s.initial_ang);
        b.inerLin = v3<R>(s.inertial_lin);
        b.inerAng = q4<R>(s.inertial_ang);
        b.vel = v3<R>(s.vel);
        b.angVel = v3<R>(s.ang_vel);
        b.prevVel = v3<R>(s.prev_vel);
        b.shapeRadius = s.shape_radius;
        b.halfLength = s.half_length;
        b.shape = s.shape;
        b.denseId = s.dense_id;
    }
    for (int i = 0; i < nj; ++i)
    {
        const avbd_joint &s = joints[i];
        Joint<R> j;
        j.a = s.body_a;
        j.b = s.body_b;
        j.serial = s.serial;
        j.rA = v3<R>(s.r_a);
        j.rB = v3<R>(s.r_b);
        j.C0Lin = v3<R>(s.c0_lin);
        j.C0Ang = v3<R>(s.c0_ang);
        j.penLin = v3<R>(s.penalty_lin);
        j.penAng = v3<R>(s.penalty_ang);
        j.lamLin = v3<R>(s.lambda_lin);
        j.lamAng = v3<R>(s.lambda_ang);
        j.kLin = s.stiffness_lin;
        j.kAng = s.stiffness_ang;
        j.fracture = s.fracture;
        j.torqueArm = s.torque_arm;
        j.broken = s.broken != 0;
        w.joints.push_back(j);
    }
    for (int i = 0; i < ns; ++i)
    {
        const avbd_spring &s = springs[i];
        Spring<R> sp;
        sp.a = s.body_a;
        sp.b = s.body_b;
        sp.serial = s.serial;
        sp.rA = v3<R>(s.r_a);
        sp.rB = v3<R>(s.r_b);
        sp.rest = s.rest;
        sp.k = s.stiffness;
        w.springs.push_back(sp);
    }
    for (int i = 0; i < nm; ++i)
    {
        const avbd_manifold &s = manifolds[i];
        Manifold<R> m{};
        m.a = s.body_a;
        m.b = s.body_b;
        m.serial = s.serial;
        m.n = s.num_contacts;
        m.friction = s.friction;
        for (int r = 0; r < 3; ++r)
            m.basis[r] = v3<R>(s.basis + 3 * r);
        for (int c = 0; c < 8; ++c)
            m.c[c] = contact_of<R>(s.contacts[c]);
        w.manifolds.push_back(m);
    }
    auto bySerial = [](const auto &x, const auto &y) { return x.serial < y.serial; };
    std::stable_sort(w.joints.begin(), w.joints.end(), bySerial);
    std::stable_sort(w.springs.begin(), w.springs.end(), bySerial);
    std::stable_sort(w.manifolds.begin(), w.manifolds.end(), bySerial);
    w.nextSerial = nextSerial;
    w.colors.assign(nb, -1);
}

template <class R>
WorldBase *build(const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
                 const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm, int64_t nextSerial)
{
    auto *impl = new WorldImpl<R>();
    fill(impl->w, p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial);
    return impl;
}

// Forward-mode reference gradient. Seeds FP64 directions on the initial linear and angular
// velocities, the masses (each moment scales with its body's mass) and the friction
// coefficients (any seed pointer may be NULL), steps a World<DualFD> in colour order and returns
// per body the values and tangents of pos(3) rot(4) vel(3) angvel(3).
template <class D>
void tangent_run(const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
                 const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm, int64_t nextSerial,
                 const int32_t *ignores, int ni, int steps, const double *dvel, const double *dangvel,
                 const double *dmass, const double *dfriction, double *val13, double *tan13)
{
    World<D> w;
    fill(w, p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial);
    for (int k = 0; k < ni; ++k)
        w.ignores.push_back({ignores[2 * k], ignores[2 * k + 1]});
    for (int i = 0; i < nb; ++i)
    {
        Body<D> &b = w.bodies[i];
        for (int k = 0; k < 3; ++k)
        {
            if (dvel)
                b.vel[k].d = dvel[3 * i + k];
            if (dangvel)
                b.angVel[k].d = dangvel[3 * i + k];
        }
        if (dmass && b.mass.v > 0.0f)
        {
            b.mass.d = dmass[i];
            for (int k = 0; k < 3; ++k)
                b.moment[k].d = dmass[i] * (double)b.moment[k].v / (double)b.mass.v;
        }
        if (dfriction)
            b.friction.d = dfriction[i];
    }
    for (int s = 0; s < steps; ++s)
        w.step();
    for (int i = 0; i < nb; ++i)
    {
        const Body<D> &b = w.bodies[i];
        const D f[13] = {b.pos.x, b.pos.y, b.pos.z, b.rot.x, b.rot.y, b.rot.z, b.rot.w,
                              b.vel.x, b.vel.y, b.vel.z, b.angVel.x, b.angVel.y, b.angVel.z};
        for (int k = 0; k < 13; ++k)
        {
            val13[13 * i + k] = f[k].v;
            tan13[13 * i + k] = f[k].d;
        }
    }
}

// FP64 run with FP64 overrides of the initial velocities, angular velocities, masses (moments
// scale with the mass) and friction coefficients (NULL keeps the table's value): the central
// finite differences of the adjoint checks perturb these. Output per body as orc_get_bodies_f64.
} // namespace

extern "C" {

void *orc_world_create(int fp64, const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
                       const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm, int64_t nextSerial)
{
    if (fp64)
        return build<double>(p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial);
    return build<float>(p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial);
}

void orc_world_destroy(void *w) { delete (WorldBase *)w; }

void orc_tangent_run(int fp64, const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints,
                     int nj, const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm,
                     int64_t nextSerial, const int32_t *ignores, int ni, int steps, const double *dvel,
                     const double *dangvel, const double *dmass, const double *dfriction, double *val13,
                     double *tan13)
{
    if (fp64)
        tangent_run<DualDD>(p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial, ignores, ni, steps,
                            dvel, dangvel, dmass, dfriction, val13, tan13);
    else
        tangent_run<DualFD>(p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial, ignores, ni, steps,
                            dvel, dangvel, dmass, dfriction, val13, tan13);
}

void orc_f64_run(const avbd_params *p, const avbd_body *bodies, int nb, const avbd_joint *joints, int nj,
                 const avbd_spring *springs, int ns, const avbd_manifold *manifolds, int nm, int64_t nextSerial,
                 const int32_t *ignores, int ni, int steps, const double *vel, const double *angvel,
                 const double *mass, const double *friction, double *out13)
{
    WorldImpl<double> impl;
    World<double> &w = impl.w;
    fill(w, p, bodies, nb, joints, nj, springs, ns, manifolds, nm, nextSerial);
    for (int k = 0; k < ni; ++k)
        w.ignores.push_back({ignores[2 * k], ignores[2 * k + 1]});
    for (int i = 0; i < nb; ++i)
    {
        Body<double> &b = w.bodies[i];
        for (int k = 0; k < 3; ++k)
        {
            if (vel)
                b.vel[k] = vel[3 * i + k];
            if (angvel)
                b.angVel[k] = angvel[3 * i + k];
        }
        if (mass && bodies[i].mass > 0.0f)
        {
            b.mass = mass[i];
            for (int k = 0; k < 3; ++k)
                b.moment[k] = mass[i] * ((double)bodies[i].moment[k] / (double)bodies[i].mass);
        }
        if (friction)
            b.friction = friction[i];
    }
    impl.step(steps);
    impl.getBodiesF64(out13);
}
/* Frozen-branch run: a World<ShadowT> from the state, with FP64 overrides of the initial
 * velocities, angular velocities, masses (moments scale with the mass) and friction applied to the
 * FP64 values only. Its shadows step the FP32 trajectory of the unperturbed state and take every
 * decision; out13 receives the FP64 values (as orc_f64_run), shadow13 (may be NULL) the shadows. */
void orc_shadow_run(const avbd_params *p, const avbd_body *bodies, int nb, const avbd_manifold *manifolds, int nm,
                    int64_t nextSerial, int steps, const double *vel, const double *angvel, const double *mass,
                    const double *friction, double *out13, double *shadow13)
{
    World<ShadowT> w;
    fill(w, p, bodies, nb, nullptr, 0, nullptr, 0, manifolds, nm, nextSerial);
    for (int i = 0; i < nb; ++i)
    {
        Body<ShadowT> &b = w.bodies[i];
        for (int k = 0; k < 3; ++k)
        {
            if (vel)
                b.vel[k].v = vel[3 * i + k];
            if (angvel)
                b.angVel[k].v = angvel[3 * i + k];
        }
        if (mass && bodies[i].mass > 0.0f)
        {
            b.mass.v = mass[i];
            for (int k = 0; k < 3; ++k)
                b.moment[k].v = mass[i] * ((double)bodies[i].moment[k] / (double)bodies[i].mass);
        }
        if (friction)
            b.friction.v = friction[i];
    }
    for (int s = 0; s < steps; ++s)
        w.step();
    for (int i = 0; i < nb; ++i)
    {
        const Body<ShadowT> &b = w.bodies[i];
        const ShadowT f[13] = {b.pos.x, b.pos.y, b.pos.z, b.rot.x, b.rot.y, b.rot.z, b.rot.w,
                               b.vel.x, b.vel.y, b.vel.z, b.angVel.x, b.angVel.y, b.angVel.z};
        for (int k = 0; k < 13; ++k)
        {
            out13[13 * i + k] = f[k].v;
            if (shadow13)
                shadow13[13 * i + k] = (double)f[k].s;
        }
    }
}

void orc_set_order(void *w, int mode) { ((WorldBase *)w)->setOrder(mode); }
/* Solver::broadphaseMode (solver.h:39-45): 0 all pairs, 1 uniform grid, 2 sweep and prune (default). */
void orc_set_broadphase(void *w, int mode) { ((WorldBase *)w)->setBroadphase(mode); }
/* External pair and contact sources (Solver::externalBroadphasePairs, useExternalBroadphasePairs and
 * the external contact table, solver.h:361-365): pairs_on selects broadphaseExternalPairs, `pairs`
 * holds np (a, b) list-index pairs, `contacts` nc table entries; replaces the previous sources. */
void orc_set_external(void *w, int pairs_on, const int32_t *pairs, int np, const avbd_external_contact *contacts, int nc)
{
    ((WorldBase *)w)->setExternal(pairs_on, pairs, np, contacts, nc);
}
/* IgnoreCollision pairs (list indices), replacing the previous set. */
void orc_set_ignores(void *w, const int32_t *pairs, int n) { ((WorldBase *)w)->setIgnores(pairs, n); }
void orc_step(void *w, int n) { ((WorldBase *)w)->step(n); }
void orc_prepare(void *w) { ((WorldBase *)w)->prepare(); }
void orc_iterate(void *w) { ((WorldBase *)w)->iterate(); }
void orc_finish(void *w) { ((WorldBase *)w)->finish(); }
void orc_counts(void *w, int32_t *nb, int32_t *nj, int32_t *ns, int32_t *nm, int64_t *serial)
{
    ((WorldBase *)w)->counts(nb, nj, ns, nm, serial);
}
void orc_get_state(void *w, avbd_body *b, avbd_joint *j, avbd_spring *s, avbd_manifold *m)
{
    ((WorldBase *)w)->get(b, j, s, m);
}
void orc_get_bodies_f64(void *w, double *out) { ((WorldBase *)w)->getBodiesF64(out); }
void orc_get_manifolds_f64(void *w, double *out) { ((WorldBase *)w)->getManifoldsF64(out); }
int orc_get_colors(void *w, int32_t *out, int cap) { return ((WorldBase *)w)->colors(out, cap); }
int orc_get_hits(void *w, int32_t *out, int cap) { return ((WorldBase *)w)->hits(out, cap); }
int orc_sweep_axis(void *w) { return ((WorldBase *)w)->axis(); }
int orc_pair_checks(void *w) { return ((WorldBase *)w)->pairChecks(); }
int orc_color_count(void *w) { return ((WorldBase *)w)->colorCount(); }

/* Timed CPU run: n steps, returns wall seconds (the oracle's own loop is single-threaded). */
double orc_time_steps(void *w, int n)
{
    auto t0 = std::chrono::steady_clock::now();
    ((WorldBase *)w)->step(n);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

/* Rigid::Rigid mass properties for a box (rigid.cpp:24-42). */
void orc_make_box_body(const float size[3], float density, float friction, const float pos[3], const float rot[4],
                       const float vel[3], const float angVel[3], avbd_body *out)
{
    std::memset(out, 0, sizeof(*out));
    V3<float> s = v3<float>(size);
    float mass = s.x * s.y * s.z * density;
    V3<float> moment{(s.y * s.y + s.z * s.z) / 12.0f * mass, (s.x * s.x + s.z * s.z) / 12.0f * mass,
                     (s.x * s.x + s.y * s.y) / 12.0f * mass};
    put3(out->size, s);
    out->mass = mass;
    put3(out->moment, moment);
    out->friction = friction;
    out->radius = length(s * 0.5f);
    std::memcpy(out->pos, pos, 12);
    std::memcpy(out->rot, rot, 16);
    std::memcpy(out->vel, vel, 12);
    std::memcpy(out->ang_vel, angVel, 12);
    std::memcpy(out->prev_vel, vel, 12);
    std::memcpy(out->initial_lin, pos, 12);
    std::memcpy(out->initial_ang, rot, 16);
    std::memcpy(out->inertial_lin, pos, 12);
    std::memcpy(out->inertial_ang, rot, 16);
    out->shape = AVBD_SHAPE_BOX;
    out->shape_radius = out->radius;
}

/* Rigid::makeSphere / makeCapsule / makeCylinder mass properties (rigid.cpp:44-86): the box
 * constructor's record of the bounding size, then the shape's own mass, moments and radii. */
int orc_make_round_body(int shape, float radius, float halfLength, float density, float friction, const float pos[3],
                        const float rot[4], const float vel[3], const float angVel[3], avbd_body *out)
{
    const float pi = 3.14159265358979323846f;
    float d = radius * 2.0f;
    if (shape == AVBD_SHAPE_SPHERE)
        halfLength = 0.0f;
    float len = halfLength * 2.0f;
    float size[3] = {d, d, shape == AVBD_SHAPE_CAPSULE ? len + d : shape == AVBD_SHAPE_CYLINDER ? len : d};
    orc_make_box_body(size, density, friction, pos, rot, vel, angVel, out);
    auto ball = [&](float r) { return 4.0f / 3.0f * pi * r * r * r * density; };
    if (shape == AVBD_SHAPE_SPHERE)
    {
        out->mass = ball(radius);
        float in = 2.0f / 5.0f * out->mass * radius * radius;
        out->moment[0] = out->moment[1] = out->moment[2] = in;
        out->radius = radius;
    }
    else if (shape == AVBD_SHAPE_CAPSULE)
    {
        float mc = pi * radius * radius * len * density;
        float mb = ball(radius);
        out->mass = mc + mb;
        float axial = 0.5f * mc * radius * radius + 2.0f / 5.0f * mb * radius * radius;
        float across = (3.0f * radius * radius + len * len) / 12.0f * mc + 2.0f / 5.0f * mb * radius * radius +
                       mb * halfLength * halfLength;
        out->moment[0] = out->moment[1] = across;
        out->moment[2] = axial;
        out->radius = halfLength + radius;
    }
    else if (shape == AVBD_SHAPE_CYLINDER)
    {
        out->mass = pi * radius * radius * len * density;
        float axial = 0.5f * out->mass * radius * radius;
        float across = (3.0f * radius * radius + len * len) / 12.0f * out->mass;
        out->moment[0] = out->moment[1] = across;
        out->moment[2] = axial;
        out->radius = sqrtf(radius * radius + halfLength * halfLength);
    }
    else
        return -1;
    out->shape = shape;
    out->shape_radius = radius;
    out->half_length = halfLength;
    return 0;
}

/* Round-shape branch-hit counters (avbd_oracle::RoundBranch order); reset = 1 clears them. */
int orc_round_branch_hits(long long *out, int reset)
{
    for (int i = 0; i < avbd_oracle::RB_COUNT; ++i)
    {
        if (out)
            out[i] = avbd_oracle::roundBranchHits[i];
        if (reset)
            avbd_oracle::roundBranchHits[i] = 0;
    }
    return avbd_oracle::RB_COUNT;
}

/* Narrowphase of two bodies of any shapes (Manifold::collide, collide.cpp:1154-1181); dt is the
 * solver step the swept sphere test reads. Returns the contact count. */
int orc_collide_pair(const avbd_body *a, const avbd_body *b, float dt, avbd_contact *out, float basis[9])
{
    avbd_oracle::World<float> w;
    for (const avbd_body *s : {a, b})
    {
        Body<float> o{};
        o.size = v3<float>(s->size);
        o.mass = s->mass;
        o.friction = s->friction;
        o.radius = s->radius;
        o.pos = v3<float>(s->pos);
        o.rot = q4<float>(s->rot);
        o.vel = v3<float>(s->vel);
        o.angVel = v3<float>(s->ang_vel);
        o.shapeRadius = s->shape_radius;
        o.halfLength = s->half_length;
        o.shape = s->shape;
        w.bodies.push_back(o);
    }
    Contact<float> c[8];
    std::memset((void *)c, 0, sizeof(c));
    M3<float> bs{};
    int n = collidePair(w.bodies[0], w.bodies[1], dt, c, bs);
    for (int r = 0; r < 3; ++r)
        put3(basis + 3 * r, bs[r]);
    for (int i = 0; i < n; ++i)
    {
        std::memset(&out[i], 0, sizeof(avbd_contact));
        out[i].feature = c[i].feature;
        put3(out[i].r_a, c[i].rA);
        put3(out[i].r_b, c[i].rB);
    }
    return n;
}

/* Box-box narrowphase of two bodies (collide.cpp:898-955). Returns the contact count. */
int orc_collide(int fp64, const avbd_body *a, const avbd_body *b, avbd_contact *out, float basis[9])
{
    auto run = [&](auto tag) {
        using R = decltype(tag);
        Body<R> A{}, B{};
        A.pos = v3<R>(a->pos);
        A.rot = q4<R>(a->rot);
        A.size = v3<R>(a->size);
        B.pos = v3<R>(b->pos);
        B.rot = q4<R>(b->rot);
        B.size = v3<R>(b->size);
        Contact<R> c[8];
        std::memset((void *)c, 0, sizeof(c));
        M3<R> bs{};
        int n = collideBoxes(A, B, c, bs);
        for (int r = 0; r < 3; ++r)
            put3(basis + 3 * r, bs[r]);
        for (int i = 0; i < n; ++i)
        {
            std::memset(&out[i], 0, sizeof(avbd_contact));
            out[i].feature = c[i].feature;
            put3(out[i].r_a, c[i].rA);
            put3(out[i].r_b, c[i].rB);
        }
        return n;
    };
    return fp64 ? run(0.0) : run(0.0f);
}

/* 6x6 LDL^T solve (maths.h:481-547); A given as lin/ang/cross 3x3 row-major blocks. */
void orc_solve6(int fp64, const double *lin9, const double *ang9, const double *cross9, const double *b6, double *x6)
{
    auto run = [&](auto tag) {
        using R = decltype(tag);
        M3<R> L, A, C;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
            {
                L[r][c] = (R)lin9[3 * r + c];
                A[r][c] = (R)ang9[3 * r + c];
                C[r][c] = (R)cross9[3 * r + c];
            }
        V3<R> xl, xa;
        solve6(L, A, C, V3<R>{(R)b6[0], (R)b6[1], (R)b6[2]}, V3<R>{(R)b6[3], (R)b6[4], (R)b6[5]}, xl, xa);
        x6[0] = xl.x; x6[1] = xl.y; x6[2] = xl.z;
        x6[3] = xa.x; x6[4] = xa.y; x6[5] = xa.z;
    };
    if (fp64)
        run(0.0);
    else
        run(0.0f);
}

/* Quaternion helpers for identity tests (maths.h:269-277, 441-446). */
void orc_quat_ops(const float *q, const float *p, const float *w, const float *v, float *diff3, float *integ4, float *rot3)
{
    Q<float> a = q4<float>(q), b = q4<float>(p);
    put3(diff3, a - b);
    put4(integ4, a + v3<float>(w));
    put3(rot3, rotate(a, v3<float>(v)));
}

int orc_sizeof(int which)
{
    switch (which)
    {
    case 0: return (int)sizeof(avbd_body);
    case 1: return (int)sizeof(avbd_contact);
    case 2: return (int)sizeof(avbd_manifold);
    case 3: return (int)sizeof(avbd_joint);
    case 4: return (int)sizeof(avbd_spring);
    case 5: return (int)sizeof(avbd_params);
    case 9: return (int)sizeof(avbd_external_contact);
    default: return -1;
    }
}

} // extern "C"

extern "C" {

/* Joint::Joint / Spring::Spring derived fields (joint.cpp:25-33, spring.cpp:14-25):
 * torque arm = |sizeA + sizeB|^2 (sizeA = 0 for a world anchor); a negative
 * rest length is replaced by the current anchor distance. */
void orc_joint_spring_derived(const avbd_body *a, const avbd_body *b, const float rA[3], const float rB[3],
                              float *torqueArm, float *rest)
{
    V3<float> sa = a ? v3<float>(a->size) : V3<float>{0, 0, 0};
    *torqueArm = lengthSq(sa + v3<float>(b->size));
    if (a)
    {
        V3<float> pa = transform(v3<float>(a->pos), q4<float>(a->rot), v3<float>(rA));
        V3<float> pb = transform(v3<float>(b->pos), q4<float>(b->rot), v3<float>(rB));
        *rest = length(pa - pb);
    }
    else
        *rest = -1.0f;
}

} // extern "C"
