// avbd_round.hpp -- sphere, capsule and cylinder collision (collide.cpp:435-1181) and the
// shape-dependent contact offsets (manifold.cpp:14-90), generic over the scalar as
// avbd_collide.hpp is: T = float is the step's narrowphase, T = DualF the adjoint's Jacobian of
// the basis and anchors (its value half repeats the FP32 operations, so it takes the step's
// branches).
//
// Operand order follows the reference operation by operation (no FMA contraction), so every
// branch and every contact point carries the reference's FP32 bits. Component access by a
// runtime axis index goes through ternaries (compv / setc) so nothing is spilled to scratch.
// Shape constants (radii, half lengths, masses) stay float; poses and velocities are T.
#pragma once

#include "avbd_collide.hpp"

namespace avbd
{

// RigidShapeType (solver.h:47-53)
constexpr int kShapeBox = 0;
constexpr int kShapeSphere = 1;
constexpr int kShapeCapsule = 2;
constexpr int kShapeCylinder = 3;
constexpr float kMarginC = 0.01f; // COLLISION_MARGIN

// One body as the narrowphase sees it.
template <class T> struct ShapeBodyT
{
    V3T<T> pos, size, vel;
    Q4T<T> rot;
    float mass, radius, halfLength; // radius / halfLength: RigidShape::radius / halfLength
    int shape;
};
using ShapeBody = ShapeBodyT<float>;

template <class T> AVBD_HD void setc(V3T<T> &v, int i, T s)
{
    if (i == 0)
        v.x = s;
    else if (i == 1)
        v.y = s;
    else
        v.z = s;
}
template <class T> AVBD_HD V3T<T> c3(float x, float y, float z) { return V3T<T>{T(x), T(y), T(z)}; }
template <class T> AVBD_HD V3T<T> local_z(const ShapeBodyT<T> &b) { return rotate(b.rot, c3<T>(0.0f, 0.0f, 1.0f)); }
AVBD_HD bool sphere_or_capsule(int s) { return s == kShapeSphere || s == kShapeCapsule; }

// Core segment of a capsule (a sphere's is its centre twice), collide.cpp:468-481.
template <class T> AVBD_HD void round_core(const ShapeBodyT<T> &b, V3T<T> &s0, V3T<T> &s1)
{
    if (b.shape == kShapeCapsule)
    {
        V3T<T> z = local_z(b);
        s0 = b.pos - z * b.halfLength;
        s1 = b.pos + z * b.halfLength;
        return;
    }
    s0 = b.pos;
    s1 = b.pos;
}
// collide.cpp:483-488
template <class T> AVBD_HD void cylinder_core(const ShapeBodyT<T> &b, V3T<T> &s0, V3T<T> &s1)
{
    V3T<T> z = local_z(b);
    s0 = b.pos - z * b.halfLength;
    s1 = b.pos + z * b.halfLength;
}

// Box-frame helpers (collide.cpp:435-456).
template <class T> AVBD_HD V3T<T> obb_local(const ObbT<T> &o, V3T<T> p)
{
    V3T<T> d = p - o.c;
    return V3T<T>{dot(d, o.ax0), dot(d, o.ax1), dot(d, o.ax2)};
}
template <class T> AVBD_HD V3T<T> obb_world(const ObbT<T> &o, V3T<T> l)
{
    return o.c + o.ax0 * l.x + o.ax1 * l.y + o.ax2 * l.z;
}
template <class T> AVBD_HD V3T<T> obb_closest(const ObbT<T> &o, V3T<T> p)
{
    V3T<T> d = p - o.c;
    return o.c + o.ax0 * clampv(dot(d, o.ax0), -o.h.x, o.h.x) + o.ax1 * clampv(dot(d, o.ax1), -o.h.y, o.h.y) +
           o.ax2 * clampv(dot(d, o.ax2), -o.h.z, o.h.z);
}

// collide.cpp:490-498
template <class T> AVBD_HD V3T<T> seg_closest(V3T<T> a, V3T<T> b, V3T<T> p)
{
    V3T<T> ab = b - a;
    T den = lensq(ab);
    if (den < kSatEps)
        return a;
    T t = clampv(dot(p - a, ab) / den, T(0.0f), T(1.0f));
    return a + ab * t;
}

// Thin static box: smallest half extent (ties to the lower axis) at most a quarter of the other
// two, and its face pointing up (+z); collide.cpp:957-976.
template <class T> AVBD_HD bool thin_top(const ObbT<T> &o, float boxMass, int &axis, float &sign, V3T<T> &normal)
{
    if (boxMass > 0.0f)
        return false;
    axis = 0;
    if (o.h.y < compv(o.h, axis))
        axis = 1;
    if (o.h.z < compv(o.h, axis))
        axis = 2;
    int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    if (compv(o.h, axis) * 4.0f > minv(compv(o.h, a1), compv(o.h, a2)))
        return false;
    V3T<T> ax = obb_ax(o, axis);
    sign = dot(ax, c3<T>(0.0f, 0.0f, 1.0f)) >= 0.0f ? 1.0f : -1.0f;
    normal = ax * sign;
    return true;
}

// In-plane reach test against the face grown by `grow`, and the clamp onto the face.
template <class T> AVBD_HD bool thin_top_reach(const ObbT<T> &o, int axis, V3T<T> local, float grow)
{
    for (int k = 0; k < 3; ++k)
    {
        if (k == axis)
            continue;
        T l = compv(local, k), h = compv(o.h, k);
        if (l < -h - grow || l > h + grow)
            return false;
    }
    return true;
}
template <class T> AVBD_HD void clamp_others(const ObbT<T> &o, int axis, V3T<T> &local)
{
    for (int k = 0; k < 3; ++k)
        if (k != axis)
            setc(local, k, clampv(compv(local, k), -compv(o.h, k), compv(o.h, k)));
}

// Spheres centred on cA and cB (collide.cpp:500-517); A and B are the manifold's bodies.
template <class T>
AVBD_HD int round_pair(const ShapeBodyT<T> &A, const ShapeBodyT<T> &B, V3T<T> cA, float rA, V3T<T> cB, float rB,
                       int key, NewContactsT<T> &nc, V3T<T> &b0, V3T<T> &b1, V3T<T> &b2)
{
    V3T<T> delta = cB - cA;
    float rs = rA + rB;
    T d2 = lensq(delta);
    if (d2 > rs * rs)
        return 0;
    V3T<T> n = d2 > kSatEps ? delta / sqrtv(d2) : c3<T>(1.0f, 0.0f, 0.0f);
    V3T<T> xA = cA + n * rA;
    V3T<T> xB = cB - n * rB;
    orthonormal(-n, b0, b1, b2);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, xA, xB, key);
    return nc.count;
}

// Slab test of the sphere's step-start motion against the box grown by its radius
// (collide.cpp:519-587).
template <class T>
AVBD_HD bool swept_sphere(const ObbT<T> &o, const ShapeBodyT<T> &sph, const ShapeBodyT<T> &box, float dt, V3T<T> &n,
                          V3T<T> &xBox)
{
    V3T<T> rel = sph.vel - box.vel;
    if (dt <= 0.0f || lensq(rel) < kSatEps)
        return false;
    V3T<T> cur = sph.pos;
    V3T<T> prev = cur - rel * dt;
    V3T<T> l0 = obb_local(o, prev), l1 = obb_local(o, cur);
    V3T<T> dl = l1 - l0;
    V3T<T> grown = o.h + c3<T>(sph.radius, sph.radius, sph.radius);
    T tIn = T(0.0f), tOut = T(1.0f);
    int hit = -1;
    float hitSign = 1.0f;
    for (int k = 0; k < 3; ++k)
    {
        T d = compv(dl, k), a = compv(l0, k), g = compv(grown, k);
        if (absv(d) < kSatEps)
        {
            if (a < -g || a > g)
                return false;
            continue;
        }
        T inv = 1.0f / d;
        T t0 = (-g - a) * inv;
        T t1 = (g - a) * inv;
        float s = -1.0f;
        if (t0 > t1)
        {
            T tmp = t0;
            t0 = t1;
            t1 = tmp;
            s = 1.0f;
        }
        if (t0 > tIn)
        {
            tIn = t0;
            hit = k;
            hitSign = s;
        }
        tOut = minv(tOut, t1);
        if (tIn > tOut)
            return false;
    }
    if (hit < 0 || tIn < 0.0f || tIn > 1.0f)
        return false;
    V3T<T> surf = l0 + dl * tIn;
    setc(surf, hit, compv(o.h, hit) * hitSign);
    clamp_others(o, hit, surf);
    n = obb_ax(o, hit) * hitSign;
    xBox = obb_world(o, surf);
    return true;
}

// Nearest point of a sphere/capsule to a thin static top and its foot (collide.cpp:632-685).
template <class T>
AVBD_HD bool thin_top_round_point(const ObbT<T> &o, const ShapeBodyT<T> &rb, const ShapeBodyT<T> &box, V3T<T> &centre,
                                  V3T<T> &n, V3T<T> &xBox)
{
    int axis;
    float sign;
    V3T<T> normal;
    if (!thin_top(o, box.mass, axis, sign, normal))
        return false;
    V3T<T> s0, s1;
    round_core(rb, s0, s1);
    T d0 = dot(s0, normal), d1 = dot(s1, normal);
    if (rb.shape == kShapeCapsule && absv(d0 - d1) < 1.0e-4f)
        centre = seg_closest(s0, s1, rb.pos);
    else
        centre = d0 <= d1 ? s0 : s1;
    V3T<T> local = obb_local(o, centre);
    if (compv(local, axis) * sign - compv(o.h, axis) > rb.radius)
        return false;
    if (!thin_top_reach(o, axis, local, rb.radius))
        return false;
    setc(local, axis, compv(o.h, axis) * sign);
    clamp_others(o, axis, local);
    n = normal;
    xBox = obb_world(o, local);
    return true;
}

// One sphere-cap contact on a thin static top (collide.cpp:687-715).
template <class T>
AVBD_HD void thin_top_round_contact(NewContactsT<T> &nc, const ShapeBodyT<T> &A, const ShapeBodyT<T> &B,
                                    const ObbT<T> &o, const ShapeBodyT<T> &rb, bool roundIsA, int axis, float sign,
                                    V3T<T> normal, V3T<T> centre, int key)
{
    V3T<T> local = obb_local(o, centre);
    if (compv(local, axis) * sign - compv(o.h, axis) > rb.radius)
        return;
    if (!thin_top_reach(o, axis, local, rb.radius))
        return;
    setc(local, axis, compv(o.h, axis) * sign);
    clamp_others(o, axis, local);
    V3T<T> xRound = centre - normal * rb.radius;
    V3T<T> xBox = obb_world(o, local);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, roundIsA ? xRound : xBox, roundIsA ? xBox : xRound, key);
}

// collide.cpp:717-754: -1 when the box is no thin static top or nothing reaches it.
template <class T>
AVBD_HD int thin_top_round(const ObbT<T> &o, const ShapeBodyT<T> &rb, const ShapeBodyT<T> &box, bool roundIsA,
                           const ShapeBodyT<T> &A, const ShapeBodyT<T> &B, NewContactsT<T> &nc, V3T<T> &b0,
                           V3T<T> &b1, V3T<T> &b2)
{
    int axis;
    float sign;
    V3T<T> normal;
    if (!thin_top(o, box.mass, axis, sign, normal))
        return -1;
    V3T<T> nAB = roundIsA ? -normal : normal;
    orthonormal(-nAB, b0, b1, b2);
    if (rb.shape == kShapeCapsule)
    {
        V3T<T> s0, s1;
        round_core(rb, s0, s1);
        thin_top_round_contact(nc, A, B, o, rb, roundIsA, axis, sign, normal, s0, 0x04000000);
        thin_top_round_contact(nc, A, B, o, rb, roundIsA, axis, sign, normal, s1, 0x04000001);
    }
    else
        thin_top_round_contact(nc, A, B, o, rb, roundIsA, axis, sign, normal, rb.pos, 0x02000000);
    return nc.count > 0 ? nc.count : -1;
}

// Sphere/capsule against a box (collide.cpp:767-896).
template <class T>
AVBD_HD int collide_round_box(const ShapeBodyT<T> &rb, const ShapeBodyT<T> &box, bool roundIsA, float dt,
                              NewContactsT<T> &nc, V3T<T> &b0, V3T<T> &b1, V3T<T> &b2)
{
    const ShapeBodyT<T> &A = roundIsA ? rb : box;
    const ShapeBodyT<T> &B = roundIsA ? box : rb;
    ObbT<T> o = make_obb(box.pos, box.rot, box.size);
    int top = thin_top_round(o, rb, box, roundIsA, A, B, nc, b0, b1, b2);
    if (top >= 0)
        return top;
    V3T<T> centre = rb.pos;
    if (rb.shape == kShapeCapsule)
    {
        V3T<T> s0, s1;
        round_core(rb, s0, s1);
        V3T<T> c0 = obb_closest(o, s0), c1 = obb_closest(o, s1);
        centre = lensq(s0 - c0) <= lensq(s1 - c1) ? s0 : s1;
    }
    float r = rb.radius;
    V3T<T> closest = obb_closest(o, centre);
    V3T<T> toRound = centre - closest;
    T d2 = lensq(toRound);
    V3T<T> tc, tn, tp;
    if (thin_top_round_point(o, rb, box, tc, tn, tp))
    {
        centre = tc;
        toRound = tn;
        closest = tp;
        d2 = T(1.0f);
    }
    else
    {
        V3T<T> local = obb_local(o, centre);
        bool inside = absv(local.x) <= o.h.x && absv(local.y) <= o.h.y && absv(local.z) <= o.h.z;
        V3T<T> sn, sp;
        if (inside)
        {
            if (rb.shape == kShapeSphere && swept_sphere(o, rb, box, dt, sn, sp))
            {
                toRound = sn;
                closest = sp;
                d2 = T(1.0f);
            }
            else
            {
                // push-out face: a static box's up face, else the face most against the relative
                // motion, else the nearest face (collide.cpp:827-873)
                int axis = 0;
                float sign = local.x >= 0.0f ? 1.0f : -1.0f;
                T best = o.h.x - absv(local.x);
                V3T<T> rel = rb.vel - box.vel;
                if (box.mass <= 0.0f)
                {
                    axis = 0;
                    if (o.h.y < compv(o.h, axis))
                        axis = 1;
                    if (o.h.z < compv(o.h, axis))
                        axis = 2;
                    sign = dot(obb_ax(o, axis), c3<T>(0.0f, 0.0f, 1.0f)) >= 0.0f ? 1.0f : -1.0f;
                }
                else if (lensq(rel) > kSatEps)
                {
                    V3T<T> against = -rel;
                    T bestDot = T(-FLT_MAX);
                    for (int k = 0; k < 3; ++k)
                        for (int si = 0; si < 2; ++si)
                        {
                            float s = si == 0 ? -1.0f : 1.0f;
                            T d = dot(obb_ax(o, k) * s, against);
                            if (d > bestDot)
                            {
                                bestDot = d;
                                axis = k;
                                sign = s;
                            }
                        }
                }
                else
                {
                    T dy = o.h.y - absv(local.y);
                    if (dy < best)
                    {
                        axis = 1;
                        sign = local.y >= 0.0f ? 1.0f : -1.0f;
                        best = dy;
                    }
                    T dz = o.h.z - absv(local.z);
                    if (dz < best)
                    {
                        axis = 2;
                        sign = local.z >= 0.0f ? 1.0f : -1.0f;
                    }
                }
                V3T<T> surf = local;
                setc(surf, axis, compv(o.h, axis) * sign);
                closest = obb_world(o, surf);
                toRound = obb_ax(o, axis) * sign;
                d2 = T(1.0f);
            }
        }
        else if (d2 > r * r)
        {
            if (rb.shape != kShapeSphere || !swept_sphere(o, rb, box, dt, sn, sp))
                return 0;
            toRound = sn;
            closest = sp;
            d2 = T(1.0f);
        }
    }
    V3T<T> n = d2 > kSatEps ? toRound / sqrtv(d2) : c3<T>(0.0f, 0.0f, 1.0f);
    V3T<T> nAB = roundIsA ? -n : n;
    V3T<T> xRound = centre - n * r;
    orthonormal(-nAB, b0, b1, b2);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, roundIsA ? xRound : closest, roundIsA ? closest : xRound,
                rb.shape == kShapeCapsule ? 0x04000000 : 0x02000000);
    return nc.count;
}

// Rim point of a cylinder against a thin static top (collide.cpp:978-1007).
template <class T>
AVBD_HD void thin_top_cylinder_contact(NewContactsT<T> &nc, const ShapeBodyT<T> &A, const ShapeBodyT<T> &B,
                                       const ObbT<T> &o, const ShapeBodyT<T> &cyl, bool cylIsA, int axis, float sign,
                                       V3T<T> xCyl, int key)
{
    V3T<T> local = obb_local(o, xCyl);
    if (compv(local, axis) * sign - compv(o.h, axis) > kMarginC)
        return;
    if (!thin_top_reach(o, axis, local, cyl.radius))
        return;
    setc(local, axis, (compv(o.h, axis) + kMarginC) * sign);
    clamp_others(o, axis, local);
    V3T<T> xBox = obb_world(o, local);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, cylIsA ? xCyl : xBox, cylIsA ? xBox : xCyl, key);
}

// Unit vector perpendicular to n: orthonormal's first tangent, normalised on its own.
template <class T> AVBD_HD V3T<T> perp_unit(V3T<T> n)
{
    V3T<T> t = absv(n.x) > absv(n.z) ? V3T<T>{-n.y, n.x, T(0.0f)} : V3T<T>{T(0.0f), -n.z, n.y};
    return t / sqrtv(lensq(t));
}
template <class T> AVBD_HD V3T<T> unit(V3T<T> v) { return v / sqrtv(lensq(v)); }

// Standing cylinder's cap rim (4 points) or a lying one's two base rims on a thin static top;
// -1 when that does not apply (collide.cpp:1009-1053).
template <class T>
AVBD_HD int thin_top_cylinder(const ObbT<T> &o, const ShapeBodyT<T> &cyl, const ShapeBodyT<T> &box, bool cylIsA,
                              NewContactsT<T> &nc, V3T<T> &b0, V3T<T> &b1, V3T<T> &b2)
{
    int axis;
    float sign;
    V3T<T> normal;
    if (!thin_top(o, box.mass, axis, sign, normal))
        return -1;
    V3T<T> nAB = cylIsA ? -normal : normal;
    orthonormal(-nAB, b0, b1, b2);
    const ShapeBodyT<T> &A = cylIsA ? cyl : box;
    const ShapeBodyT<T> &B = cylIsA ? box : cyl;
    V3T<T> z = local_z(cyl);
    T zd = dot(z, normal);
    float r = cyl.radius;
    if (absv(zd) > 0.65f)
    {
        float zs = zd >= 0.0f ? 1.0f : -1.0f;
        V3T<T> cap = cyl.pos - z * (cyl.halfLength * zs);
        V3T<T> u = perp_unit(z);
        V3T<T> v = unit(cross(z, u));
        V3T<T> rim[4] = {u * r, -u * r, v * r, -v * r};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            thin_top_cylinder_contact(nc, A, B, o, cyl, cylIsA, axis, sign, cap + rim[i], 0x05000000 | i);
    }
    else
    {
        V3T<T> radial = normal - z * zd;
        if (lensq(radial) < kSatEps)
            radial = normal;
        radial = unit(radial);
        V3T<T> e0 = cyl.pos - z * cyl.halfLength;
        V3T<T> e1 = cyl.pos + z * cyl.halfLength;
        thin_top_cylinder_contact(nc, A, B, o, cyl, cylIsA, axis, sign, e0 - radial * r, 0x05000100);
        thin_top_cylinder_contact(nc, A, B, o, cyl, cylIsA, axis, sign, e1 - radial * r, 0x05000101);
    }
    return nc.count > 0 ? nc.count : -1;
}

// Two cylinders (collide.cpp:1080-1152).
template <class T>
AVBD_HD int collide_cylinders(const ShapeBodyT<T> &A, const ShapeBodyT<T> &B, NewContactsT<T> &nc, V3T<T> &b0,
                              V3T<T> &b1, V3T<T> &b2)
{
    V3T<T> zA = local_z(A);
    V3T<T> zB = local_z(B);
    T par = dot(zA, zB);
    V3T<T> delta = B.pos - A.pos;
    if (absv(par) > 0.95f)
    {
        T axial = dot(delta, zA);
        V3T<T> radial = delta - zA * axial;
        T rl2 = lensq(radial);
        float rs = A.radius + B.radius;
        if (rl2 > rs * rs)
            return 0;
        if (A.halfLength + B.halfLength - absv(axial) < -kMarginC)
            return 0;
        if (absv(axial) > (A.halfLength + B.halfLength) * 0.5f)
        {
            // end caps facing each other: four rim points of the smaller cap
            float side = axial >= 0.0f ? 1.0f : -1.0f;
            V3T<T> n = zA * side;
            V3T<T> capA = A.pos + n * A.halfLength;
            V3T<T> capB = B.pos - n * B.halfLength;
            float rr = minv(A.radius, B.radius) * 0.85f;
            orthonormal(-n, b0, b1, b2);
            V3T<T> u = perp_unit(n);
            V3T<T> v = unit(cross(n, u));
            V3T<T> rim[4] = {u * rr, -u * rr, v * rr, -v * rr};
#pragma unroll
            for (int i = 0; i < 4; ++i)
                add_contact(nc, A.pos, A.rot, B.pos, B.rot, capA + rim[i], capB + rim[i], 0x06000000 | i);
            return nc.count;
        }
        // side by side: the two ends of the axial overlap
        V3T<T> n;
        if (rl2 > kSatEps)
            n = radial / sqrtv(rl2);
        else
        {
            V3T<T> o0, o1, o2;
            orthonormal(zA, o0, o1, o2);
            n = o1;
        }
        orthonormal(-n, b0, b1, b2);
        T ends[2] = {maxv(T(-A.halfLength), axial - B.halfLength), minv(T(A.halfLength), axial + B.halfLength)};
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            T tA = ends[i];
            T tB = tA - axial;
            V3T<T> xA = A.pos + zA * tA + n * A.radius;
            V3T<T> xB = B.pos + zA * tB - n * B.radius;
            add_contact(nc, A.pos, A.rot, B.pos, B.rot, xA, xB, 0x06000100 | i);
        }
        return nc.count;
    }
    V3T<T> a0, a1, q0, q1, cA, cB;
    cylinder_core(A, a0, a1);
    cylinder_core(B, q0, q1);
    closest_seg(a0, a1, q0, q1, cA, cB);
    return round_pair(A, B, cA, A.radius, cB, B.radius, 0x06000200, nc, b0, b1, b2);
}

// Manifold::collide (collide.cpp:1154-1181): the pair dispatch. A cylinder against a box that is
// no thin static top is collided as the box of its bounding size.
template <class T>
AVBD_HD int collide_pair(const ShapeBodyT<T> &A, const ShapeBodyT<T> &B, float dt, NewContactsT<T> &nc, V3T<T> &b0,
                         V3T<T> &b1, V3T<T> &b2)
{
    nc.count = 0;
    bool cylA = A.shape == kShapeCylinder, cylB = B.shape == kShapeCylinder;
    bool rndA = sphere_or_capsule(A.shape), rndB = sphere_or_capsule(B.shape);
    if (cylA && cylB)
        return collide_cylinders(A, B, nc, b0, b1, b2);
    if ((cylA && rndB) || (rndA && cylB))
    {
        // collide.cpp:1065-1078
        const ShapeBodyT<T> &cyl = cylA ? A : B;
        const ShapeBodyT<T> &rb = cylA ? B : A;
        V3T<T> c0, c1, r0, r1, cc, cr;
        cylinder_core(cyl, c0, c1);
        round_core(rb, r0, r1);
        closest_seg(c0, c1, r0, r1, cc, cr);
        if (cylA)
            return round_pair(A, B, cc, cyl.radius, cr, rb.radius, 0x07000000, nc, b0, b1, b2);
        return round_pair(A, B, cr, rb.radius, cc, cyl.radius, 0x07000000, nc, b0, b1, b2);
    }
    if ((cylA && B.shape == kShapeBox) || (A.shape == kShapeBox && cylB))
    {
        const ShapeBodyT<T> &cyl = cylA ? A : B;
        const ShapeBodyT<T> &box = cylA ? B : A;
        ObbT<T> o = make_obb(box.pos, box.rot, box.size);
        int top = thin_top_cylinder(o, cyl, box, cylA, nc, b0, b1, b2);
        if (top >= 0)
            return top;
        return collide_boxes<T>(A.pos, A.rot, A.size, B.pos, B.rot, B.size, nc, b0, b1, b2);
    }
    if (rndA && rndB)
    {
        // collide.cpp:756-765
        V3T<T> a0, a1, q0, q1, cA, cB;
        round_core(A, a0, a1);
        round_core(B, q0, q1);
        closest_seg(a0, a1, q0, q1, cA, cB);
        return round_pair(A, B, cA, A.radius, cB, B.radius, 0x03000000, nc, b0, b1, b2);
    }
    if (rndA && B.shape == kShapeBox)
        return collide_round_box(A, B, true, dt, nc, b0, b1, b2);
    if (A.shape == kShapeBox && rndB)
        return collide_round_box(B, A, false, dt, nc, b0, b1, b2);
    return collide_boxes<T>(A.pos, A.rot, A.size, B.pos, B.rot, B.size, nc, b0, b1, b2);
}

// ---- contact offsets (manifold.cpp:14-90) ----

// A manifold's shape code as the contact records carry it: shape of A in bits 0-1, of B in
// bits 2-3, bit 4 set for cylinder cap contacts (feature keys 0x050000xx / 0x060000xx). Zero
// for box-box contacts, whose offsets are the rotated local anchors.
constexpr int kCodeCap = 16;
AVBD_HD int shape_code(int shapeA, int shapeB, int key)
{
    bool cap = (key & (int)0xFFFFFF00) == 0x05000000 || (key & (int)0xFFFFFF00) == 0x06000000;
    return shapeA | (shapeB << 2) | (cap ? kCodeCap : 0);
}
// manifold.cpp:14-27: every round shape (cylinders included) refreshes its contact anchors.
AVBD_HD bool any_round(int s) { return s == kShapeSphere || s == kShapeCapsule || s == kShapeCylinder; }

// World-frame offset of a contact point from the body centre (manifold.cpp:29-90); n is the
// manifold normal (basis row 0), isA selects the side.
template <class T>
AVBD_HD V3T<T> contact_offset_t(int shape, float radius, float halfLength, bool cap, Q4T<T> q, V3T<T> lo, V3T<T> n,
                                bool isA)
{
    if (shape == kShapeSphere)
        return n * (isA ? -radius : radius);
    if (shape == kShapeCapsule)
    {
        V3T<T> z = rotate(q, c3<T>(0.0f, 0.0f, 1.0f));
        V3T<T> nl = rotate(qconj(q), n);
        T cz = isA ? lo.z + nl.z * radius : lo.z - nl.z * radius;
        cz = clampv(cz, T(-halfLength), T(halfLength));
        V3T<T> along = z * cz;
        return along + n * (isA ? -radius : radius);
    }
    if (shape == kShapeCylinder)
    {
        V3T<T> dirW = n * (isA ? -1.0f : 1.0f);
        V3T<T> dirL = rotate(qconj(q), dirW);
        T storedRad = sqrtv(lo.x * lo.x + lo.y * lo.y);
        if (cap)
        {
            V3T<T> l = c3<T>(0.0f, 0.0f, 0.0f);
            if (storedRad > 1.0e-6f)
            {
                l.x = lo.x / storedRad * radius;
                l.y = lo.y / storedRad * radius;
            }
            float capSign = absv(lo.z) > 1.0e-6f ? (lo.z >= 0.0f ? 1.0f : -1.0f) : (dirL.z >= 0.0f ? 1.0f : -1.0f);
            l.z = T(halfLength * capSign);
            return rotate(q, l);
        }
        T dirRad = sqrtv(dirL.x * dirL.x + dirL.y * dirL.y);
        T z = absv(dirL.z) > 0.15f ? T(dirL.z >= 0.0f ? halfLength : -halfLength)
                                   : clampv(lo.z, T(-halfLength), T(halfLength));
        V3T<T> l{T(0.0f), T(0.0f), z};
        if (dirRad > 1.0e-6f)
        {
            l.x = dirL.x / dirRad * radius;
            l.y = dirL.y / dirRad * radius;
        }
        else if (storedRad > 1.0e-6f)
        {
            l.x = lo.x / storedRad * radius;
            l.y = lo.y / storedRad * radius;
        }
        return rotate(q, l);
    }
    return rotate(q, lo);
}
AVBD_HD V3 contact_offset(int shape, float radius, float halfLength, bool cap, Q4 q, V3 lo, V3 n, bool isA)
{
    return fv(contact_offset_t<float>(shape, radius, halfLength, cap, tq(q), tv(lo), tv(n), isA));
}

} // namespace avbd
