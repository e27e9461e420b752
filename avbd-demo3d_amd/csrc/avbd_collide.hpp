// avbd_collide.hpp -- box-box collision (collide.cpp:60-433, 898-955), generic over the scalar.
//
// One source serves two uses:
//   * T = float: the step's narrowphase. Operand order follows the reference, so with
//     -ffp-contract=off it produces the reference's FP32 bits.
//   * T = DualF (FP32 value, FP64 tangent): the adjoint's pose Jacobian of a manifold's basis and
//     contact offsets. The value half runs exactly the FP32 operations of the float instance, so
//     every branch (SAT axis, reference/incident face, clipping, depth test, merge) is the one the
//     step took, and the tangent is the derivative of that branch.
#pragma once

#include "avbd_math.hpp"

#include <cfloat>

namespace avbd
{

constexpr float kSatEps = 1.0e-6f;
constexpr float kPlaneEps = 1.0e-5f;
constexpr float kMergeDistSq = 1.0e-6f;
constexpr int kMaxNewContacts = 8;

// ---- FP32 value with an FP64 tangent ----
struct DualF
{
    float v;
    double d;
    DualF() = default;
    AVBD_HD DualF(float x) : v(x), d(0.0) {}
    AVBD_HD DualF(float x, double dx) : v(x), d(dx) {}
};

AVBD_HD DualF operator+(DualF a, DualF b) { return DualF(a.v + b.v, a.d + b.d); }
AVBD_HD DualF operator-(DualF a, DualF b) { return DualF(a.v - b.v, a.d - b.d); }
AVBD_HD DualF operator-(DualF a) { return DualF(-a.v, -a.d); }
AVBD_HD DualF operator*(DualF a, DualF b) { return DualF(a.v * b.v, a.d * (double)b.v + (double)a.v * b.d); }
AVBD_HD DualF operator/(DualF a, DualF b)
{
    double q = (double)a.v / (double)b.v;
    return DualF(a.v / b.v, (a.d - q * b.d) / (double)b.v);
}
AVBD_HD DualF operator*(DualF a, float s) { return DualF(a.v * s, a.d * (double)s); }
AVBD_HD DualF operator*(float s, DualF a) { return DualF(s * a.v, (double)s * a.d); }
AVBD_HD DualF operator/(DualF a, float s) { return DualF(a.v / s, a.d / (double)s); }
AVBD_HD DualF operator/(float s, DualF a)
{
    double q = (double)s / (double)a.v;
    return DualF(s / a.v, -q * a.d / (double)a.v);
}
AVBD_HD DualF operator+(DualF a, float s) { return DualF(a.v + s, a.d); }
AVBD_HD DualF operator-(DualF a, float s) { return DualF(a.v - s, a.d); }
AVBD_HD bool operator<(DualF a, DualF b) { return a.v < b.v; }
AVBD_HD bool operator>(DualF a, DualF b) { return a.v > b.v; }
AVBD_HD bool operator<=(DualF a, DualF b) { return a.v <= b.v; }
AVBD_HD bool operator>=(DualF a, DualF b) { return a.v >= b.v; }
AVBD_HD float val(float x) { return x; }
AVBD_HD float val(DualF x) { return x.v; }
AVBD_HD float absv(float x) { return fabsf(x); }
AVBD_HD DualF absv(DualF x) { return x.v < 0.0f ? -x : x; }
AVBD_HD float sqrtv(float x) { return sqrtf(x); }
// the derivative at 0 (a zero vector's length, a kink) is taken as 0
AVBD_HD DualF sqrtv(DualF x)
{
    float r = sqrtf(x.v);
    return DualF(r, x.v == 0.0f ? 0.0 : x.d / (2.0 * sqrt((double)x.v)));
}
template <class T> AVBD_HD T minv(T a, T b) { return a < b ? a : b; }
template <class T> AVBD_HD T maxv(T a, T b) { return a > b ? a : b; }
template <class T> AVBD_HD T clampv(T x, T a, T b) { return maxv(a, minv(b, x)); }

// ---- generic vectors and quaternions (the operand order of avbd_math.hpp) ----
template <class T> struct V3T
{
    T x, y, z;
};
template <class T> struct Q4T
{
    T x, y, z, w;
};

template <class T> AVBD_HD V3T<T> operator+(V3T<T> a, V3T<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> AVBD_HD V3T<T> operator-(V3T<T> a, V3T<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> AVBD_HD V3T<T> operator-(V3T<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T, class S> AVBD_HD V3T<T> operator*(V3T<T> a, S s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T, class S> AVBD_HD V3T<T> operator/(V3T<T> a, S s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> AVBD_HD T dot(V3T<T> a, V3T<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> AVBD_HD T lensq(V3T<T> a) { return dot(a, a); }
template <class T> AVBD_HD V3T<T> cross(V3T<T> a, V3T<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> AVBD_HD T compv(V3T<T> v, int i) { return i == 0 ? v.x : i == 1 ? v.y : v.z; }
template <class T> AVBD_HD Q4T<T> qconj(Q4T<T> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class T> AVBD_HD V3T<T> rotate(Q4T<T> q, V3T<T> v)
{
    V3T<T> u{q.x, q.y, q.z};
    V3T<T> t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}
// orthonormal basis with the normal in row 0 (maths.h:468-474)
template <class T> AVBD_HD void orthonormal(V3T<T> n, V3T<T> &r0, V3T<T> &r1, V3T<T> &r2)
{
    V3T<T> t1 = absv(n.x) > absv(n.z) ? V3T<T>{-n.y, n.x, T(0.0f)} : V3T<T>{T(0.0f), -n.z, n.y};
    t1 = t1 / sqrtv(lensq(t1));
    r0 = n;
    r1 = t1;
    r2 = cross(n, t1);
}

AVBD_HD V3T<float> tv(V3 v) { return {v.x, v.y, v.z}; }
AVBD_HD Q4T<float> tq(Q4 q) { return {q.x, q.y, q.z, q.w}; }
AVBD_HD V3 fv(V3T<float> v) { return V3{v.x, v.y, v.z}; }

// ---- oriented boxes ----
template <class T> struct ObbT
{
    V3T<T> c, h, ax0, ax1, ax2;
};

// Component-wise value selects (a conditional of lvalues selects an address, which forces the
// operands into private memory).
template <class T> AVBD_HD V3T<T> sel3(bool c, V3T<T> a, V3T<T> b)
{
    return V3T<T>{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z};
}

template <class T> AVBD_HD V3T<T> obb_ax(const ObbT<T> &o, int k) { return sel3(k == 0, o.ax0, sel3(k == 1, o.ax1, o.ax2)); }

template <class T> AVBD_HD ObbT<T> make_obb(V3T<T> pos, Q4T<T> rot, V3T<T> size)
{
    ObbT<T> o;
    o.c = pos;
    o.h = size * 0.5f;
    o.ax0 = rotate(rot, V3T<T>{T(1.0f), T(0.0f), T(0.0f)});
    o.ax1 = rotate(rot, V3T<T>{T(0.0f), T(1.0f), T(0.0f)});
    o.ax2 = rotate(rot, V3T<T>{T(0.0f), T(0.0f), T(1.0f)});
    return o;
}

template <class T> AVBD_HD V3T<T> support_pt(const ObbT<T> &o, V3T<T> d)
{
    float sx = dot(d, o.ax0) >= 0.0f ? 1.0f : -1.0f;
    float sy = dot(d, o.ax1) >= 0.0f ? 1.0f : -1.0f;
    float sz = dot(d, o.ax2) >= 0.0f ? 1.0f : -1.0f;
    return o.c + o.ax0 * (o.h.x * sx) + o.ax1 * (o.h.y * sy) + o.ax2 * (o.h.z * sz);
}

template <class T> struct SatBestT
{
    int type, ia, ib;
    T sep;
    V3T<T> n;
    bool valid;
};

template <class T>
AVBD_HD bool sat_axis(const ObbT<T> &A, const ObbT<T> &B, V3T<T> delta, V3T<T> axis, int type, int ia, int ib,
                      SatBestT<T> &best)
{
    T l2 = lensq(axis);
    if (l2 < kSatEps)
        return true;
    T inv = 1.0f / sqrtv(l2);
    V3T<T> n = axis * inv;
    if (dot(n, delta) < 0.0f)
        n = -n;
    T dist = absv(dot(delta, n));
    T rA = A.h.x * absv(dot(n, A.ax0)) + A.h.y * absv(dot(n, A.ax1)) + A.h.z * absv(dot(n, A.ax2));
    T rB = B.h.x * absv(dot(n, B.ax0)) + B.h.y * absv(dot(n, B.ax1)) + B.h.z * absv(dot(n, B.ax2));
    T sep = dist - (rA + rB);
    if (sep > 0.0f)
        return false;
    if (!best.valid || sep > best.sep)
    {
        best.valid = true;
        best.type = type;
        best.ia = ia;
        best.ib = ib;
        best.sep = sep;
        best.n = n;
    }
    return true;
}

// Contacts of one collide call: feature key and body-local offsets (ContactBuilder, collide.cpp:30-58).
template <class T> struct NewContactsT
{
    int count;
    int feature[kMaxNewContacts];
    V3T<T> rA[kMaxNewContacts];
    V3T<T> rB[kMaxNewContacts];
    V3T<T> mid[kMaxNewContacts];
};

// Clip polygon of up to 16 vertices in the lane's private memory (the generic storage).
template <class T> struct LocalPoly
{
    V3T<T> v[16];
    AVBD_HD V3T<T> get(int i) const { return v[i]; }
    AVBD_HD void set(int i, V3T<T> x) { v[i] = x; }
};

// The clip polygon storage a contact sink brings: private arrays for NewContactsT.
template <class T> AVBD_HD LocalPoly<T> make_poly(NewContactsT<T> &, int) { return LocalPoly<T>{}; }

template <class T>
AVBD_HD void add_contact(NewContactsT<T> &nc, V3T<T> posA, Q4T<T> rotA, V3T<T> posB, Q4T<T> rotB, V3T<T> xA,
                         V3T<T> xB, int key)
{
    V3T<T> m = (xA + xB) * 0.5f;
    for (int i = 0; i < nc.count; ++i)
        if (lensq(m - nc.mid[i]) < kMergeDistSq)
            return;
    if (nc.count >= kMaxNewContacts)
        return;
    int c = nc.count;
    nc.feature[c] = key;
    nc.rA[c] = rotate(qconj(rotA), xA - posA);
    nc.rB[c] = rotate(qconj(rotB), xB - posB);
    nc.mid[c] = m;
    nc.count = c + 1;
}

template <class T, class PI, class PO> AVBD_HD int clip_plane(const PI &in, int n, V3T<T> pn, T off, PO &out)
{
    if (n <= 0)
        return 0;
    int m = 0;
    V3T<T> a = in.get(n - 1);
    T da = dot(pn, a) - off;
    for (int i = 0; i < n; ++i)
    {
        V3T<T> b = in.get(i);
        T db = dot(pn, b) - off;
        bool ain = da <= kPlaneEps;
        bool bin = db <= kPlaneEps;
        if (ain != bin)
        {
            T t = T(0.0f);
            T den = da - db;
            if (absv(den) > kSatEps)
                t = clampv(da / den, T(0.0f), T(1.0f));
            if (m < 16)
                out.set(m++, a + (b - a) * t);
        }
        if (bin && m < 16)
            out.set(m++, b);
        a = b;
        da = db;
    }
    return m;
}

template <class T> AVBD_HD void face_axes(const ObbT<T> &o, int k, V3T<T> &u, V3T<T> &v, T &eu, T &ev)
{
    // k = 0: (ax1, ax2, h.y, h.z); k = 1: (ax0, ax2, h.x, h.z); k = 2: (ax0, ax1, h.x, h.y)
    const T hx = o.h.x, hy = o.h.y, hz = o.h.z;
    u = sel3(k == 0, o.ax1, o.ax0);
    v = sel3(k == 2, o.ax1, o.ax2);
    eu = k == 0 ? hy : hx;
    ev = k == 2 ? hy : hz;
}

// Component-wise select of two boxes (a select of references, or of whole aggregates, would
// force both boxes into private memory).
template <class T> AVBD_HD ObbT<T> obb_sel(bool c, const ObbT<T> &a, const ObbT<T> &b)
{
    ObbT<T> o;
    o.c = sel3(c, a.c, b.c);
    o.h = sel3(c, a.h, b.h);
    o.ax0 = sel3(c, a.ax0, b.ax0);
    o.ax1 = sel3(c, a.ax1, b.ax1);
    o.ax2 = sel3(c, a.ax2, b.ax2);
    return o;
}

template <class T, class NC>
AVBD_HD void face_manifold(NC &nc, V3T<T> posA, Q4T<T> rotA, V3T<T> posB, Q4T<T> rotB, const ObbT<T> &A,
                           const ObbT<T> &B, bool refA, int refAxis, V3T<T> nAB)
{
    const ObbT<T> ref = obb_sel(refA, A, B);
    const ObbT<T> inc = obb_sel(refA, B, A);
    V3T<T> outward = refA ? nAB : -nAB;
    V3T<T> rax = obb_ax(ref, refAxis);
    float fs = dot(outward, rax) >= 0.0f ? 1.0f : -1.0f;
    V3T<T> fn = rax * fs;
    V3T<T> fc = ref.c + fn * compv(ref.h, refAxis);
    V3T<T> fu, fv;
    T eu, ev;
    face_axes(ref, refAxis, fu, fv, eu, ev);

    int incAxis = 0;
    float best = -FLT_MAX;
    for (int i = 0; i < 3; ++i)
    {
        float d = val(absv(dot(obb_ax(inc, i), fn)));
        if (d > best)
        {
            best = d;
            incAxis = i;
        }
    }
    auto p0 = make_poly(nc, 0);
    auto p1 = make_poly(nc, 1);
    {
        V3T<T> iax = obb_ax(inc, incAxis);
        float s = dot(iax, fn) > 0.0f ? -1.0f : 1.0f;
        V3T<T> n = iax * s;
        V3T<T> c = inc.c + n * compv(inc.h, incAxis);
        V3T<T> u, v;
        T ue, ve;
        face_axes(inc, incAxis, u, v, ue, ve);
        p0.set(0, c + u * ue + v * ve);
        p0.set(1, c - u * ue + v * ve);
        p0.set(2, c - u * ue - v * ve);
        p0.set(3, c + u * ue - v * ve);
    }
    int n = 4;
    n = clip_plane(p0, n, fu, dot(fu, fc) + eu, p1);
    if (!n)
        return;
    V3T<T> nu = -fu;
    n = clip_plane(p1, n, nu, dot(nu, fc) + eu, p0);
    if (!n)
        return;
    n = clip_plane(p0, n, fv, dot(fv, fc) + ev, p1);
    if (!n)
        return;
    V3T<T> nv = -fv;
    n = clip_plane(p1, n, nv, dot(nv, fc) + ev, p0);
    if (!n)
        return;
    int prefix = ((refA ? 0 : 1) << 24) | ((refAxis & 0xFF) << 16) | ((incAxis & 0xFF) << 8);
    for (int i = 0; i < n && nc.count < kMaxNewContacts; ++i)
    {
        V3T<T> pi = p0.get(i);
        T depth = dot(pi - fc, fn);
        if (depth > kPlaneEps)
            continue;
        V3T<T> pr = pi - fn * depth;
        add_contact(nc, posA, rotA, posB, rotB, refA ? pr : pi, refA ? pi : pr, prefix | (i & 0xFF));
    }
    if (!nc.count)
        add_contact(nc, posA, rotA, posB, rotB, support_pt(A, nAB), support_pt(B, -nAB), prefix);
}

template <class T> AVBD_HD void edge_support(const ObbT<T> &o, int k, V3T<T> d, V3T<T> &e0, V3T<T> &e1)
{
    int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    V3T<T> a1 = obb_ax(o, k1), a2 = obb_ax(o, k2), a0 = obb_ax(o, k);
    float s1 = dot(d, a1) >= 0.0f ? 1.0f : -1.0f;
    float s2 = dot(d, a2) >= 0.0f ? 1.0f : -1.0f;
    V3T<T> mid = o.c + a1 * (compv(o.h, k1) * s1) + a2 * (compv(o.h, k2) * s2);
    e0 = mid - a0 * compv(o.h, k);
    e1 = mid + a0 * compv(o.h, k);
}

template <class T> AVBD_HD void closest_seg(V3T<T> p0, V3T<T> p1, V3T<T> q0, V3T<T> q1, V3T<T> &c0, V3T<T> &c1)
{
    V3T<T> d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    T s = T(0.0f), t = T(0.0f);
    if (a <= kSatEps && e <= kSatEps)
    {
        c0 = p0;
        c1 = q0;
        return;
    }
    if (a <= kSatEps)
        t = clampv(f / e, T(0.0f), T(1.0f));
    else
    {
        T c = dot(d1, r);
        if (e <= kSatEps)
            s = clampv(-c / a, T(0.0f), T(1.0f));
        else
        {
            T b = dot(d1, d2);
            T den = a * e - b * b;
            if (absv(den) > kSatEps)
                s = clampv((b * f - c * e) / den, T(0.0f), T(1.0f));
            t = (b * s + f) / e;
            if (t < 0.0f)
            {
                t = T(0.0f);
                s = clampv(-c / a, T(0.0f), T(1.0f));
            }
            else if (t > 1.0f)
            {
                t = T(1.0f);
                s = clampv((b - c) / a, T(0.0f), T(1.0f));
            }
        }
    }
    c0 = p0 + d1 * s;
    c1 = q0 + d2 * t;
}

// Box-box collide: returns the contact count; the basis rows are written only when the boxes
// overlap on all 15 axes.
// NC: the contact sink (NewContactsT<T>, or a sink that takes each contact as it is found and
// supplies add_contact / make_poly overloads of its own).
template <class T, class NC>
AVBD_HD int collide_boxes(V3T<T> posA, Q4T<T> rotA, V3T<T> sizeA, V3T<T> posB, Q4T<T> rotB, V3T<T> sizeB,
                          NC &nc, V3T<T> &b0, V3T<T> &b1, V3T<T> &b2)
{
    nc.count = 0;
    ObbT<T> A = make_obb(posA, rotA, sizeA), B = make_obb(posB, rotB, sizeB);
    V3T<T> delta = B.c - A.c;
    SatBestT<T> face, edge;
    face.valid = false;
    face.sep = T(-FLT_MAX);
    face.type = face.ia = face.ib = 0;
    face.n = V3T<T>{T(0.0f), T(0.0f), T(0.0f)};
    edge = face;
    for (int i = 0; i < 3; ++i)
        if (!sat_axis(A, B, delta, obb_ax(A, i), 0, i, -1, face))
            return 0;
    for (int i = 0; i < 3; ++i)
        if (!sat_axis(A, B, delta, obb_ax(B, i), 1, -1, i, face))
            return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (!sat_axis(A, B, delta, cross(obb_ax(A, i), obb_ax(B, j)), 2, i, j, edge))
                return 0;
    if (!face.valid)
        return 0;
    SatBestT<T> best = face;
    if (edge.valid && 0.95f * val(edge.sep) > val(face.sep) + 0.01f)
        best = edge;
    orthonormal(-best.n, b0, b1, b2);
    if (best.type == 2)
    {
        V3T<T> a0, a1, q0, q1, xA, xB;
        edge_support(A, best.ia, best.n, a0, a1);
        edge_support(B, best.ib, -best.n, q0, q1);
        closest_seg(a0, a1, q0, q1, xA, xB);
        int key = (2 << 24) | ((best.ia & 0xFF) << 8) | (best.ib & 0xFF);
        add_contact(nc, posA, rotA, posB, rotB, xA, xB, key);
        if (!nc.count)
            add_contact(nc, posA, rotA, posB, rotB, support_pt(A, best.n), support_pt(B, -best.n), key);
        return nc.count;
    }
    if (best.type == 0)
        face_manifold(nc, posA, rotA, posB, rotB, A, B, true, best.ia, best.n);
    else
        face_manifold(nc, posA, rotA, posB, rotB, A, B, false, best.ib, best.n);
    return nc.count;
}

} // namespace avbd
