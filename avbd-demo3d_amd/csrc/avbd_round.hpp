// avbd_round.hpp -- sphere, capsule and cylinder collision (collide.cpp:435-1181) and the
// shape-dependent contact offsets (manifold.cpp:14-90), FP32.
//
// Operand order follows the reference operation by operation (no FMA contraction), so every
// branch and every contact point carries the reference's FP32 bits. Component access by a
// runtime axis index goes through ternaries (compv / setc) so nothing is spilled to scratch.
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
struct ShapeBody
{
    V3T<float> pos, size, vel;
    Q4T<float> rot;
    float mass, radius, halfLength; // radius / halfLength: RigidShape::radius / halfLength
    int shape;
};

using F3 = V3T<float>;

AVBD_HD void setc(F3 &v, int i, float s)
{
    if (i == 0)
        v.x = s;
    else if (i == 1)
        v.y = s;
    else
        v.z = s;
}
AVBD_HD F3 f3(float x, float y, float z) { return F3{x, y, z}; }
AVBD_HD F3 local_z(const ShapeBody &b) { return rotate(b.rot, f3(0.0f, 0.0f, 1.0f)); }
AVBD_HD bool sphere_or_capsule(int s) { return s == kShapeSphere || s == kShapeCapsule; }

// Core segment of a capsule (a sphere's is its centre twice), collide.cpp:468-481.
AVBD_HD void round_core(const ShapeBody &b, F3 &s0, F3 &s1)
{
    if (b.shape == kShapeCapsule)
    {
        F3 z = local_z(b);
        s0 = b.pos - z * b.halfLength;
        s1 = b.pos + z * b.halfLength;
        return;
    }
    s0 = b.pos;
    s1 = b.pos;
}
// collide.cpp:483-488
AVBD_HD void cylinder_core(const ShapeBody &b, F3 &s0, F3 &s1)
{
    F3 z = local_z(b);
    s0 = b.pos - z * b.halfLength;
    s1 = b.pos + z * b.halfLength;
}

// Box-frame helpers (collide.cpp:435-456).
AVBD_HD F3 obb_local(const ObbT<float> &o, F3 p)
{
    F3 d = p - o.c;
    return f3(dot(d, o.ax0), dot(d, o.ax1), dot(d, o.ax2));
}
AVBD_HD F3 obb_world(const ObbT<float> &o, F3 l) { return o.c + o.ax0 * l.x + o.ax1 * l.y + o.ax2 * l.z; }
AVBD_HD F3 obb_closest(const ObbT<float> &o, F3 p)
{
    F3 d = p - o.c;
    return o.c + o.ax0 * clampv(dot(d, o.ax0), -o.h.x, o.h.x) + o.ax1 * clampv(dot(d, o.ax1), -o.h.y, o.h.y) +
           o.ax2 * clampv(dot(d, o.ax2), -o.h.z, o.h.z);
}

// collide.cpp:490-498
AVBD_HD F3 seg_closest(F3 a, F3 b, F3 p)
{
    F3 ab = b - a;
    float den = lensq(ab);
    if (den < kSatEps)
        return a;
    float t = clampv(dot(p - a, ab) / den, 0.0f, 1.0f);
    return a + ab * t;
}

// Thin static box: smallest half extent (ties to the lower axis) at most a quarter of the other
// two, and its face pointing up (+z); collide.cpp:957-976.
AVBD_HD bool thin_top(const ObbT<float> &o, float boxMass, int &axis, float &sign, F3 &normal)
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
    F3 ax = obb_ax(o, axis);
    sign = dot(ax, f3(0.0f, 0.0f, 1.0f)) >= 0.0f ? 1.0f : -1.0f;
    normal = ax * sign;
    return true;
}

// In-plane reach test against the face grown by `grow`, and the clamp onto the face.
AVBD_HD bool thin_top_reach(const ObbT<float> &o, int axis, F3 local, float grow)
{
    for (int k = 0; k < 3; ++k)
    {
        if (k == axis)
            continue;
        float l = compv(local, k), h = compv(o.h, k);
        if (l < -h - grow || l > h + grow)
            return false;
    }
    return true;
}
AVBD_HD void clamp_others(const ObbT<float> &o, int axis, F3 &local)
{
    for (int k = 0; k < 3; ++k)
        if (k != axis)
            setc(local, k, clampv(compv(local, k), -compv(o.h, k), compv(o.h, k)));
}

// Spheres centred on cA and cB (collide.cpp:500-517); A and B are the manifold's bodies.
AVBD_HD int round_pair(const ShapeBody &A, const ShapeBody &B, F3 cA, float rA, F3 cB, float rB, int key,
                       NewContactsT<float> &nc, F3 &b0, F3 &b1, F3 &b2)
{
    F3 delta = cB - cA;
    float rs = rA + rB;
    float d2 = lensq(delta);
    if (d2 > rs * rs)
        return 0;
    F3 n = d2 > kSatEps ? delta / sqrtf(d2) : f3(1.0f, 0.0f, 0.0f);
    F3 xA = cA + n * rA;
    F3 xB = cB - n * rB;
    orthonormal(-n, b0, b1, b2);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, xA, xB, key);
    return nc.count;
}

// Slab test of the sphere's step-start motion against the box grown by its radius
// (collide.cpp:519-587).
AVBD_HD bool swept_sphere(const ObbT<float> &o, const ShapeBody &sph, const ShapeBody &box, float dt, F3 &n, F3 &xBox)
{
    F3 rel = sph.vel - box.vel;
    if (dt <= 0.0f || lensq(rel) < kSatEps)
        return false;
    F3 cur = sph.pos;
    F3 prev = cur - rel * dt;
    F3 l0 = obb_local(o, prev), l1 = obb_local(o, cur);
    F3 dl = l1 - l0;
    F3 grown = o.h + f3(sph.radius, sph.radius, sph.radius);
    float tIn = 0.0f, tOut = 1.0f;
    int hit = -1;
    float hitSign = 1.0f;
    for (int k = 0; k < 3; ++k)
    {
        float d = compv(dl, k), a = compv(l0, k), g = compv(grown, k);
        if (fabsf(d) < kSatEps)
        {
            if (a < -g || a > g)
                return false;
            continue;
        }
        float inv = 1.0f / d;
        float t0 = (-g - a) * inv;
        float t1 = (g - a) * inv;
        float s = -1.0f;
        if (t0 > t1)
        {
            float tmp = t0;
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
    F3 surf = l0 + dl * tIn;
    setc(surf, hit, compv(o.h, hit) * hitSign);
    clamp_others(o, hit, surf);
    n = obb_ax(o, hit) * hitSign;
    xBox = obb_world(o, surf);
    return true;
}

// Nearest point of a sphere/capsule to a thin static top and its foot (collide.cpp:632-685).
AVBD_HD bool thin_top_round_point(const ObbT<float> &o, const ShapeBody &rb, const ShapeBody &box, F3 &centre, F3 &n,
                                  F3 &xBox)
{
    int axis;
    float sign;
    F3 normal;
    if (!thin_top(o, box.mass, axis, sign, normal))
        return false;
    F3 s0, s1;
    round_core(rb, s0, s1);
    float d0 = dot(s0, normal), d1 = dot(s1, normal);
    if (rb.shape == kShapeCapsule && fabsf(d0 - d1) < 1.0e-4f)
        centre = seg_closest(s0, s1, rb.pos);
    else
        centre = d0 <= d1 ? s0 : s1;
    F3 local = obb_local(o, centre);
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
AVBD_HD void thin_top_round_contact(NewContactsT<float> &nc, const ShapeBody &A, const ShapeBody &B,
                                    const ObbT<float> &o, const ShapeBody &rb, bool roundIsA, int axis, float sign,
                                    F3 normal, F3 centre, int key)
{
    F3 local = obb_local(o, centre);
    if (compv(local, axis) * sign - compv(o.h, axis) > rb.radius)
        return;
    if (!thin_top_reach(o, axis, local, rb.radius))
        return;
    setc(local, axis, compv(o.h, axis) * sign);
    clamp_others(o, axis, local);
    F3 xRound = centre - normal * rb.radius;
    F3 xBox = obb_world(o, local);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, roundIsA ? xRound : xBox, roundIsA ? xBox : xRound, key);
}

// collide.cpp:717-754: -1 when the box is no thin static top or nothing reaches it.
AVBD_HD int thin_top_round(const ObbT<float> &o, const ShapeBody &rb, const ShapeBody &box, bool roundIsA,
                           const ShapeBody &A, const ShapeBody &B, NewContactsT<float> &nc, F3 &b0, F3 &b1, F3 &b2)
{
    int axis;
    float sign;
    F3 normal;
    if (!thin_top(o, box.mass, axis, sign, normal))
        return -1;
    F3 nAB = roundIsA ? -normal : normal;
    orthonormal(-nAB, b0, b1, b2);
    if (rb.shape == kShapeCapsule)
    {
        F3 s0, s1;
        round_core(rb, s0, s1);
        thin_top_round_contact(nc, A, B, o, rb, roundIsA, axis, sign, normal, s0, 0x04000000);
        thin_top_round_contact(nc, A, B, o, rb, roundIsA, axis, sign, normal, s1, 0x04000001);
    }
    else
        thin_top_round_contact(nc, A, B, o, rb, roundIsA, axis, sign, normal, rb.pos, 0x02000000);
    return nc.count > 0 ? nc.count : -1;
}

// Sphere/capsule against a box (collide.cpp:767-896).
AVBD_HD int collide_round_box(const ShapeBody &rb, const ShapeBody &box, bool roundIsA, float dt,
                              NewContactsT<float> &nc, F3 &b0, F3 &b1, F3 &b2)
{
    const ShapeBody &A = roundIsA ? rb : box;
    const ShapeBody &B = roundIsA ? box : rb;
    ObbT<float> o = make_obb(box.pos, box.rot, box.size);
    int top = thin_top_round(o, rb, box, roundIsA, A, B, nc, b0, b1, b2);
    if (top >= 0)
        return top;
    F3 centre = rb.pos;
    if (rb.shape == kShapeCapsule)
    {
        F3 s0, s1;
        round_core(rb, s0, s1);
        F3 c0 = obb_closest(o, s0), c1 = obb_closest(o, s1);
        centre = lensq(s0 - c0) <= lensq(s1 - c1) ? s0 : s1;
    }
    float r = rb.radius;
    F3 closest = obb_closest(o, centre);
    F3 toRound = centre - closest;
    float d2 = lensq(toRound);
    F3 tc, tn, tp;
    if (thin_top_round_point(o, rb, box, tc, tn, tp))
    {
        centre = tc;
        toRound = tn;
        closest = tp;
        d2 = 1.0f;
    }
    else
    {
        F3 local = obb_local(o, centre);
        bool inside = fabsf(local.x) <= o.h.x && fabsf(local.y) <= o.h.y && fabsf(local.z) <= o.h.z;
        F3 sn, sp;
        if (inside)
        {
            if (rb.shape == kShapeSphere && swept_sphere(o, rb, box, dt, sn, sp))
            {
                toRound = sn;
                closest = sp;
                d2 = 1.0f;
            }
            else
            {
                // push-out face: a static box's up face, else the face most against the relative
                // motion, else the nearest face (collide.cpp:827-873)
                int axis = 0;
                float sign = local.x >= 0.0f ? 1.0f : -1.0f;
                float best = o.h.x - fabsf(local.x);
                F3 rel = rb.vel - box.vel;
                if (box.mass <= 0.0f)
                {
                    axis = 0;
                    if (o.h.y < compv(o.h, axis))
                        axis = 1;
                    if (o.h.z < compv(o.h, axis))
                        axis = 2;
                    sign = dot(obb_ax(o, axis), f3(0.0f, 0.0f, 1.0f)) >= 0.0f ? 1.0f : -1.0f;
                }
                else if (lensq(rel) > kSatEps)
                {
                    F3 against = -rel;
                    float bestDot = -FLT_MAX;
                    for (int k = 0; k < 3; ++k)
                        for (int si = 0; si < 2; ++si)
                        {
                            float s = si == 0 ? -1.0f : 1.0f;
                            float d = dot(obb_ax(o, k) * s, against);
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
                    float dy = o.h.y - fabsf(local.y);
                    if (dy < best)
                    {
                        axis = 1;
                        sign = local.y >= 0.0f ? 1.0f : -1.0f;
                        best = dy;
                    }
                    float dz = o.h.z - fabsf(local.z);
                    if (dz < best)
                    {
                        axis = 2;
                        sign = local.z >= 0.0f ? 1.0f : -1.0f;
                    }
                }
                F3 surf = local;
                setc(surf, axis, compv(o.h, axis) * sign);
                closest = obb_world(o, surf);
                toRound = obb_ax(o, axis) * sign;
                d2 = 1.0f;
            }
        }
        else if (d2 > r * r)
        {
            if (rb.shape != kShapeSphere || !swept_sphere(o, rb, box, dt, sn, sp))
                return 0;
            toRound = sn;
            closest = sp;
            d2 = 1.0f;
        }
    }
    F3 n = d2 > kSatEps ? toRound / sqrtf(d2) : f3(0.0f, 0.0f, 1.0f);
    F3 nAB = roundIsA ? -n : n;
    F3 xRound = centre - n * r;
    orthonormal(-nAB, b0, b1, b2);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, roundIsA ? xRound : closest, roundIsA ? closest : xRound,
                rb.shape == kShapeCapsule ? 0x04000000 : 0x02000000);
    return nc.count;
}

// Rim point of a cylinder against a thin static top (collide.cpp:978-1007).
AVBD_HD void thin_top_cylinder_contact(NewContactsT<float> &nc, const ShapeBody &A, const ShapeBody &B,
                                       const ObbT<float> &o, const ShapeBody &cyl, bool cylIsA, int axis, float sign,
                                       F3 xCyl, int key)
{
    F3 local = obb_local(o, xCyl);
    if (compv(local, axis) * sign - compv(o.h, axis) > kMarginC)
        return;
    if (!thin_top_reach(o, axis, local, cyl.radius))
        return;
    setc(local, axis, (compv(o.h, axis) + kMarginC) * sign);
    clamp_others(o, axis, local);
    F3 xBox = obb_world(o, local);
    add_contact(nc, A.pos, A.rot, B.pos, B.rot, cylIsA ? xCyl : xBox, cylIsA ? xBox : xCyl, key);
}

// Unit vector perpendicular to n: orthonormal's first tangent, normalised on its own.
AVBD_HD F3 perp_unit(F3 n)
{
    F3 t = fabsf(n.x) > fabsf(n.z) ? f3(-n.y, n.x, 0.0f) : f3(0.0f, -n.z, n.y);
    return t / sqrtf(lensq(t));
}
AVBD_HD F3 unit(F3 v) { return v / sqrtf(lensq(v)); }

// Standing cylinder's cap rim (4 points) or a lying one's two base rims on a thin static top;
// -1 when that does not apply (collide.cpp:1009-1053).
AVBD_HD int thin_top_cylinder(const ObbT<float> &o, const ShapeBody &cyl, const ShapeBody &box, bool cylIsA,
                              NewContactsT<float> &nc, F3 &b0, F3 &b1, F3 &b2)
{
    int axis;
    float sign;
    F3 normal;
    if (!thin_top(o, box.mass, axis, sign, normal))
        return -1;
    F3 nAB = cylIsA ? -normal : normal;
    orthonormal(-nAB, b0, b1, b2);
    const ShapeBody &A = cylIsA ? cyl : box;
    const ShapeBody &B = cylIsA ? box : cyl;
    F3 z = local_z(cyl);
    float zd = dot(z, normal);
    float r = cyl.radius;
    if (fabsf(zd) > 0.65f)
    {
        float zs = zd >= 0.0f ? 1.0f : -1.0f;
        F3 cap = cyl.pos - z * (cyl.halfLength * zs);
        F3 u = perp_unit(z);
        F3 v = unit(cross(z, u));
        F3 rim[4] = {u * r, -u * r, v * r, -v * r};
#pragma unroll
        for (int i = 0; i < 4; ++i)
            thin_top_cylinder_contact(nc, A, B, o, cyl, cylIsA, axis, sign, cap + rim[i], 0x05000000 | i);
    }
    else
    {
        F3 radial = normal - z * zd;
        if (lensq(radial) < kSatEps)
            radial = normal;
        radial = unit(radial);
        F3 e0 = cyl.pos - z * cyl.halfLength;
        F3 e1 = cyl.pos + z * cyl.halfLength;
        thin_top_cylinder_contact(nc, A, B, o, cyl, cylIsA, axis, sign, e0 - radial * r, 0x05000100);
        thin_top_cylinder_contact(nc, A, B, o, cyl, cylIsA, axis, sign, e1 - radial * r, 0x05000101);
    }
    return nc.count > 0 ? nc.count : -1;
}

// Two cylinders (collide.cpp:1080-1152).
AVBD_HD int collide_cylinders(const ShapeBody &A, const ShapeBody &B, NewContactsT<float> &nc, F3 &b0, F3 &b1, F3 &b2)
{
    F3 zA = local_z(A);
    F3 zB = local_z(B);
    float par = dot(zA, zB);
    F3 delta = B.pos - A.pos;
    if (fabsf(par) > 0.95f)
    {
        float axial = dot(delta, zA);
        F3 radial = delta - zA * axial;
        float rl2 = lensq(radial);
        float rs = A.radius + B.radius;
        if (rl2 > rs * rs)
            return 0;
        if (A.halfLength + B.halfLength - fabsf(axial) < -kMarginC)
            return 0;
        if (fabsf(axial) > (A.halfLength + B.halfLength) * 0.5f)
        {
            // end caps facing each other: four rim points of the smaller cap
            float side = axial >= 0.0f ? 1.0f : -1.0f;
            F3 n = zA * side;
            F3 capA = A.pos + n * A.halfLength;
            F3 capB = B.pos - n * B.halfLength;
            float rr = minv(A.radius, B.radius) * 0.85f;
            orthonormal(-n, b0, b1, b2);
            F3 u = perp_unit(n);
            F3 v = unit(cross(n, u));
            F3 rim[4] = {u * rr, -u * rr, v * rr, -v * rr};
#pragma unroll
            for (int i = 0; i < 4; ++i)
                add_contact(nc, A.pos, A.rot, B.pos, B.rot, capA + rim[i], capB + rim[i], 0x06000000 | i);
            return nc.count;
        }
        // side by side: the two ends of the axial overlap
        F3 n;
        if (rl2 > kSatEps)
            n = radial / sqrtf(rl2);
        else
        {
            F3 o0, o1, o2;
            orthonormal(zA, o0, o1, o2);
            n = o1;
        }
        orthonormal(-n, b0, b1, b2);
        float ends[2] = {maxv(-A.halfLength, axial - B.halfLength), minv(A.halfLength, axial + B.halfLength)};
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            float tA = ends[i];
            float tB = tA - axial;
            F3 xA = A.pos + zA * tA + n * A.radius;
            F3 xB = B.pos + zA * tB - n * B.radius;
            add_contact(nc, A.pos, A.rot, B.pos, B.rot, xA, xB, 0x06000100 | i);
        }
        return nc.count;
    }
    F3 a0, a1, q0, q1, cA, cB;
    cylinder_core(A, a0, a1);
    cylinder_core(B, q0, q1);
    closest_seg(a0, a1, q0, q1, cA, cB);
    return round_pair(A, B, cA, A.radius, cB, B.radius, 0x06000200, nc, b0, b1, b2);
}

// Manifold::collide (collide.cpp:1154-1181): the pair dispatch. A cylinder against a box that is
// no thin static top is collided as the box of its bounding size.
AVBD_HD int collide_pair(const ShapeBody &A, const ShapeBody &B, float dt, NewContactsT<float> &nc, F3 &b0, F3 &b1,
                         F3 &b2)
{
    nc.count = 0;
    bool cylA = A.shape == kShapeCylinder, cylB = B.shape == kShapeCylinder;
    bool rndA = sphere_or_capsule(A.shape), rndB = sphere_or_capsule(B.shape);
    if (cylA && cylB)
        return collide_cylinders(A, B, nc, b0, b1, b2);
    if ((cylA && rndB) || (rndA && cylB))
    {
        // collide.cpp:1065-1078
        const ShapeBody &cyl = cylA ? A : B;
        const ShapeBody &rb = cylA ? B : A;
        F3 c0, c1, r0, r1, cc, cr;
        cylinder_core(cyl, c0, c1);
        round_core(rb, r0, r1);
        closest_seg(c0, c1, r0, r1, cc, cr);
        if (cylA)
            return round_pair(A, B, cc, cyl.radius, cr, rb.radius, 0x07000000, nc, b0, b1, b2);
        return round_pair(A, B, cr, rb.radius, cc, cyl.radius, 0x07000000, nc, b0, b1, b2);
    }
    if ((cylA && B.shape == kShapeBox) || (A.shape == kShapeBox && cylB))
    {
        const ShapeBody &cyl = cylA ? A : B;
        const ShapeBody &box = cylA ? B : A;
        ObbT<float> o = make_obb(box.pos, box.rot, box.size);
        int top = thin_top_cylinder(o, cyl, box, cylA, nc, b0, b1, b2);
        if (top >= 0)
            return top;
        return collide_boxes<float>(A.pos, A.rot, A.size, B.pos, B.rot, B.size, nc, b0, b1, b2);
    }
    if (rndA && rndB)
    {
        // collide.cpp:756-765
        F3 a0, a1, q0, q1, cA, cB;
        round_core(A, a0, a1);
        round_core(B, q0, q1);
        closest_seg(a0, a1, q0, q1, cA, cB);
        return round_pair(A, B, cA, A.radius, cB, B.radius, 0x03000000, nc, b0, b1, b2);
    }
    if (rndA && B.shape == kShapeBox)
        return collide_round_box(A, B, true, dt, nc, b0, b1, b2);
    if (A.shape == kShapeBox && rndB)
        return collide_round_box(B, A, false, dt, nc, b0, b1, b2);
    return collide_boxes<float>(A.pos, A.rot, A.size, B.pos, B.rot, B.size, nc, b0, b1, b2);
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
AVBD_HD V3 contact_offset(int shape, float radius, float halfLength, bool cap, Q4 q, V3 lo, V3 n, bool isA)
{
    if (shape == kShapeSphere)
        return n * (isA ? -radius : radius);
    if (shape == kShapeCapsule)
    {
        V3 z = rotate(q, V3{0.0f, 0.0f, 1.0f});
        V3 nl = rotate_inv(q, n);
        float cz = isA ? lo.z + nl.z * radius : lo.z - nl.z * radius;
        cz = clampr(cz, -halfLength, halfLength);
        V3 along = z * cz;
        return along + n * (isA ? -radius : radius);
    }
    if (shape == kShapeCylinder)
    {
        V3 dirW = n * (isA ? -1.0f : 1.0f);
        V3 dirL = rotate_inv(q, dirW);
        float storedRad = sqrtf(lo.x * lo.x + lo.y * lo.y);
        if (cap)
        {
            V3 l{0.0f, 0.0f, 0.0f};
            if (storedRad > 1.0e-6f)
            {
                l.x = lo.x / storedRad * radius;
                l.y = lo.y / storedRad * radius;
            }
            float capSign = fabsf(lo.z) > 1.0e-6f ? (lo.z >= 0.0f ? 1.0f : -1.0f) : (dirL.z >= 0.0f ? 1.0f : -1.0f);
            l.z = halfLength * capSign;
            return rotate(q, l);
        }
        float dirRad = sqrtf(dirL.x * dirL.x + dirL.y * dirL.y);
        float z = fabsf(dirL.z) > 0.15f ? (dirL.z >= 0.0f ? halfLength : -halfLength) : clampr(lo.z, -halfLength, halfLength);
        V3 l{0.0f, 0.0f, z};
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

} // namespace avbd
