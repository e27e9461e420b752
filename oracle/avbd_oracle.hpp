/*
 * avbd_oracle.hpp -- CPU oracle for the AVBD rigid-body step.
 *
 * TEST INFRASTRUCTURE ONLY. This restatement is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it. The product path never calls it.
 *
 * It restates, in its own code, the reference CPU solver of ts-jls/avbd-demo3d:
 *   maths.h           vector/quaternion algebra (maths.h:255-277, 431-479), 6x6 LDL^T (maths.h:481-547)
 *   rigid.cpp         box mass properties (rigid.cpp:24-42)
 *   joint.cpp         ball-socket joint init/primal/dual (joint.cpp:14-182)
 *   spring.cpp        penalty spring primal (spring.cpp:14-69)
 *   manifold.cpp      contact merge, warm start, primal, dual (manifold.cpp:100-280)
 *   collide.cpp       box-box SAT + clipping manifolds (collide.cpp:18-433, 898-955)
 *   solver.cpp        sweep-and-prune broadphase (solver.cpp:488-577), hit classification
 *                     (solver.cpp:131-164, 202-248), prepareStep (solver.cpp:1043-1170),
 *                     iteratePrimalDualCpu (solver.cpp:1172-1273), finishStep (solver.cpp:1275-1298)
 *   avbd_gpu_solver.cpp  greedy body coloring (avbd_gpu_solver.cpp:1812-1865)
 *
 * The scalar type is a template parameter: Real = float reproduces the
 * reference's arithmetic operation by operation (build with -ffp-contract=off),
 * Real = double gives the FP64 run used to state each FP32 result's error.
 *
 * Iteration order: mode ORDER_COLOR visits solved bodies color by color (list
 * order inside a color), which is what a colored GPU sweep computes; mode
 * ORDER_LIST visits them in body-list order exactly like solver.cpp:1181.
 *
 * Pinning (see tests/test_oracle_*.py): every function here is checked either
 * bit-exactly against golden vectors written by the compiled reference
 * (oracle/_ref) or against closed forms/invariants.
 */
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace avbd_oracle
{

template <class R>
struct V3
{
    R x, y, z;
    R &operator[](int i) { return (&x)[i]; }
    const R &operator[](int i) const { return (&x)[i]; }
};

template <class R>
struct Q
{
    R x, y, z, w;
};

template <class R>
struct M3
{
    V3<R> r[3];
    V3<R> &operator[](int i) { return r[i]; }
    const V3<R> &operator[](int i) const { return r[i]; }
    V3<R> col(int i) const { return V3<R>{r[0][i], r[1][i], r[2][i]}; }
};

// ---- scalar helpers (maths.h:281-315) ----
template <class R> inline R fmin_(R a, R b) { return a < b ? a : b; }
template <class R> inline R fmax_(R a, R b) { return a > b ? a : b; }
template <class R> inline R clampv(R x, R a, R b) { return fmax_(a, fmin_(b, x)); }
template <class R> inline R signv(R x) { return x < 0 ? R(-1) : x > 0 ? R(1) : R(0); }
inline float sqrt_(float v) { return sqrtf(v); }
inline double sqrt_(double v) { return std::sqrt(v); }
inline float abs_(float v) { return fabsf(v); }
inline double abs_(double v) { return std::fabs(v); }
inline bool isfinite_(float v) { return std::isfinite(v); }
inline bool isfinite_(double v) { return std::isfinite(v); }
inline bool isinf_(float v) { return std::isinf(v); }
inline bool isinf_(double v) { return std::isinf(v); }

// ---- forward-mode scalars: a value with an FP64 tangent ----
// With an FP32 value (DualFD) the value half performs exactly the FP32 operations of Real =
// float, so a World<DualFD> follows the FP32 trajectory bit for bit (every branch included)
// while the tangent carries the FP64 directional derivative of that trajectory: the
// forward-mode reference for the adjoint. With an FP64 value (DualDD) it follows the FP64
// trajectory, the one central finite differences in FP64 see.
template <class V>
struct DualT
{
    V v;
    double d;
    DualT() = default;
    constexpr DualT(float x) : v((V)x), d(0.0) {}
    constexpr DualT(double x) : v((V)x), d(0.0) {}
    constexpr DualT(int x) : v((V)x), d(0.0) {}
    constexpr DualT(size_t x) : v((V)x), d(0.0) {}
    constexpr DualT(V x, double dx, int) : v(x), d(dx) {}
};
using DualFD = DualT<float>;
using DualDD = DualT<double>;
template <class V> inline DualT<V> mk(V v, double d) { return DualT<V>(v, d, 0); }
template <class V> inline DualT<V> operator+(DualT<V> a, DualT<V> b) { return mk<V>(a.v + b.v, a.d + b.d); }
template <class V> inline DualT<V> operator-(DualT<V> a, DualT<V> b) { return mk<V>(a.v - b.v, a.d - b.d); }
template <class V> inline DualT<V> operator-(DualT<V> a) { return mk<V>(-a.v, -a.d); }
template <class V> inline DualT<V> operator*(DualT<V> a, DualT<V> b)
{
    return mk<V>(a.v * b.v, a.d * (double)b.v + (double)a.v * b.d);
}
template <class V> inline DualT<V> operator/(DualT<V> a, DualT<V> b)
{
    double q = (double)a.v / (double)b.v;
    return mk<V>(a.v / b.v, (a.d - q * b.d) / (double)b.v);
}
template <class V> inline DualT<V> &operator+=(DualT<V> &a, DualT<V> b) { return a = a + b; }
template <class V> inline DualT<V> &operator-=(DualT<V> &a, DualT<V> b) { return a = a - b; }
template <class V> inline DualT<V> &operator*=(DualT<V> &a, DualT<V> b) { return a = a * b; }
template <class V> inline DualT<V> &operator/=(DualT<V> &a, DualT<V> b) { return a = a / b; }
template <class V> inline bool operator<(DualT<V> a, DualT<V> b) { return a.v < b.v; }
template <class V> inline bool operator>(DualT<V> a, DualT<V> b) { return a.v > b.v; }
template <class V> inline bool operator<=(DualT<V> a, DualT<V> b) { return a.v <= b.v; }
template <class V> inline bool operator>=(DualT<V> a, DualT<V> b) { return a.v >= b.v; }
template <class V> inline bool operator==(DualT<V> a, DualT<V> b) { return a.v == b.v; }
template <class V> inline bool operator!=(DualT<V> a, DualT<V> b) { return a.v != b.v; }
// mixed forms with plain literals (the templates above do not convert their operands)
#define AVBD_DUAL_MIXED(OP)                                                                                \
    template <class V, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>         \
    inline auto operator OP(DualT<V> a, S b)->decltype(a OP DualT<V>(b))                                   \
    {                                                                                                      \
        return a OP DualT<V>(b);                                                                           \
    }                                                                                                      \
    template <class V, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>         \
    inline auto operator OP(S a, DualT<V> b)->decltype(DualT<V>(a) OP b)                                   \
    {                                                                                                      \
        return DualT<V>(a) OP b;                                                                           \
    }
AVBD_DUAL_MIXED(+)
AVBD_DUAL_MIXED(-)
AVBD_DUAL_MIXED(*)
AVBD_DUAL_MIXED(/)
AVBD_DUAL_MIXED(<)
AVBD_DUAL_MIXED(>)
AVBD_DUAL_MIXED(<=)
AVBD_DUAL_MIXED(>=)
#undef AVBD_DUAL_MIXED
// At 0 the derivative of sqrt is unbounded; the zero arguments met here are lengths of zero
// vectors (a geometric-stiffness column of a joint that carries no force: joint.cpp:93,
// maths.h:476), where |h| has a kink and the derivative is taken as 0, the subgradient
// the adjoint takes too.
template <class V> inline DualT<V> sqrt_(DualT<V> x)
{
    return mk<V>(sqrt_(x.v), x.v == V(0) ? 0.0 : x.d / (2.0 * std::sqrt((double)x.v)));
}
template <class V> inline DualT<V> abs_(DualT<V> x) { return x.v < V(0) ? -x : x; }
// |x| where it scales a penalty's growth (k += beta |C|: manifold.cpp:231-280, joint.cpp:125-182):
// at x == 0 its derivative is taken as 0, the subgradient the adjoint takes there (the
// narrowphase's |x| keeps abs_).
template <class R> inline R absk_(R x) { return abs_(x); }
template <class V> inline DualT<V> absk_(DualT<V> x) { return x.v == V(0) ? mk<V>(abs_(x).v, 0.0) : abs_(x); }
// clamp(x, 0, 1) of the warm start's gravity weight (solver.cpp:1145-1149): on a bound its
// derivative is taken as 0 too, as the adjoint takes it.
template <class R> inline R clampk_(R x, R a, R b) { return clampv(x, a, b); }
template <class V> inline DualT<V> clampk_(DualT<V> x, DualT<V> a, DualT<V> b)
{
    DualT<V> c = clampv(x, a, b);
    return x.v > a.v && x.v < b.v ? c : mk<V>(c.v, 0.0);
}
// The plain value of a scalar: the broadphase's discrete decisions use the value alone.
inline float plain_(float x) { return x; }
inline double plain_(double x) { return x; }
template <class V> inline V plain_(DualT<V> x) { return x.v; }
template <class V> inline bool isfinite_(DualT<V> x) { return std::isfinite(x.v); }
template <class V> inline bool isinf_(DualT<V> x) { return std::isinf(x.v); }

// ---- frozen-branch scalars: an FP32 shadow that decides, an FP64 value that computes ----
// Every arithmetic operation is carried out twice: in FP32 on the shadow s (exactly the operations
// of Real = float) and in FP64 on the value v; every comparison, clamp, sign and stick decision
// reads the shadow. A World<ShadowT> therefore takes the FP32 trajectory's branches at every step
// (contact sets, features, clamps, penalty saturation, stick flags, colours) while v is the FP64
// evaluation of that same piecewise-smooth map. Perturbing only v (the initial velocities, masses
// or friction) leaves s and so every branch unchanged: central differences of v are then
// differences of one smooth map -- the map the adjoint differentiates -- and converge as h -> 0.
struct ShadowT
{
    float s;
    double v;
    ShadowT() = default;
    constexpr ShadowT(float x) : s(x), v((double)x) {}
    constexpr ShadowT(double x) : s((float)x), v(x) {}
    constexpr ShadowT(int x) : s((float)x), v((double)x) {}
    constexpr ShadowT(size_t x) : s((float)x), v((double)x) {}
    constexpr ShadowT(float sx, double vx, int) : s(sx), v(vx) {}
};
inline ShadowT operator+(ShadowT a, ShadowT b) { return ShadowT(a.s + b.s, a.v + b.v, 0); }
inline ShadowT operator-(ShadowT a, ShadowT b) { return ShadowT(a.s - b.s, a.v - b.v, 0); }
inline ShadowT operator-(ShadowT a) { return ShadowT(-a.s, -a.v, 0); }
inline ShadowT operator*(ShadowT a, ShadowT b) { return ShadowT(a.s * b.s, a.v * b.v, 0); }
inline ShadowT operator/(ShadowT a, ShadowT b) { return ShadowT(a.s / b.s, a.v / b.v, 0); }
inline ShadowT &operator+=(ShadowT &a, ShadowT b) { return a = a + b; }
inline ShadowT &operator-=(ShadowT &a, ShadowT b) { return a = a - b; }
inline ShadowT &operator*=(ShadowT &a, ShadowT b) { return a = a * b; }
inline ShadowT &operator/=(ShadowT &a, ShadowT b) { return a = a / b; }
inline bool operator<(ShadowT a, ShadowT b) { return a.s < b.s; }
inline bool operator>(ShadowT a, ShadowT b) { return a.s > b.s; }
inline bool operator<=(ShadowT a, ShadowT b) { return a.s <= b.s; }
inline bool operator>=(ShadowT a, ShadowT b) { return a.s >= b.s; }
inline bool operator==(ShadowT a, ShadowT b) { return a.s == b.s; }
inline bool operator!=(ShadowT a, ShadowT b) { return a.s != b.s; }
#define AVBD_SHADOW_MIXED(OP)                                                                              \
    template <class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>                  \
    inline auto operator OP(ShadowT a, S b)->decltype(a OP ShadowT(b))                                     \
    {                                                                                                      \
        return a OP ShadowT(b);                                                                            \
    }                                                                                                      \
    template <class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type>                  \
    inline auto operator OP(S a, ShadowT b)->decltype(ShadowT(a) OP b)                                     \
    {                                                                                                      \
        return ShadowT(a) OP b;                                                                            \
    }
AVBD_SHADOW_MIXED(+)
AVBD_SHADOW_MIXED(-)
AVBD_SHADOW_MIXED(*)
AVBD_SHADOW_MIXED(/)
AVBD_SHADOW_MIXED(<)
AVBD_SHADOW_MIXED(>)
AVBD_SHADOW_MIXED(<=)
AVBD_SHADOW_MIXED(>=)
#undef AVBD_SHADOW_MIXED
inline ShadowT sqrt_(ShadowT x) { return ShadowT(sqrtf(x.s), std::sqrt(x.v), 0); }
inline ShadowT abs_(ShadowT x) { return x.s < 0.0f ? -x : x; }
inline float plain_(ShadowT x) { return x.s; }
inline bool isfinite_(ShadowT x) { return std::isfinite(x.s); }
inline bool isinf_(ShadowT x) { return std::isinf(x.s); }

// ---- vector algebra (maths.h:117-165, 317-406) ----
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> &operator+=(V3<R> &a, V3<R> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class R> inline V3<R> &operator-=(V3<R> &a, V3<R> b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R lengthSq(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return sqrt_(lengthSq(a)); }
template <class R> inline V3<R> normalize(V3<R> a) { return a / length(a); }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> absv(V3<R> a) { return {abs_(a.x), abs_(a.y), abs_(a.z)}; }
template <class R> inline V3<R> absvk(V3<R> a) { return {absk_(a.x), absk_(a.y), absk_(a.z)}; }
template <class R> inline V3<R> minv(V3<R> a, R b) { return {fmin_(a.x, b), fmin_(a.y, b), fmin_(a.z, b)}; }
template <class R> inline V3<R> clampv(V3<R> v, R a, R b) { return {clampv(v.x, a, b), clampv(v.y, a, b), clampv(v.z, a, b)}; }

// ---- 3x3 algebra (maths.h:196-237, 380-429, 476-479) ----
template <class R> inline V3<R> operator*(const M3<R> &a, V3<R> b) { return {dot(a[0], b), dot(a[1], b), dot(a[2], b)}; }
template <class R> inline M3<R> operator*(const M3<R> &a, const M3<R> &b)
{
    M3<R> o;
    for (int i = 0; i < 3; ++i)
        o.r[i] = V3<R>{dot(a.r[i], b.col(0)), dot(a.r[i], b.col(1)), dot(a.r[i], b.col(2))};
    return o;
}
template <class R> inline M3<R> operator*(const M3<R> &a, R s) { return {a[0] * s, a[1] * s, a[2] * s}; }
template <class R> inline M3<R> operator/(const M3<R> &a, R s) { return {a[0] / s, a[1] / s, a[2] / s}; }
template <class R> inline M3<R> operator+(const M3<R> &a, const M3<R> &b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
template <class R> inline M3<R> operator-(const M3<R> &a) { return {-a[0], -a[1], -a[2]}; }
template <class R> inline M3<R> &operator+=(M3<R> &a, const M3<R> &b) { a[0] += b[0]; a[1] += b[1]; a[2] += b[2]; return a; }
template <class R> inline M3<R> transpose(const M3<R> &a) { return {a.col(0), a.col(1), a.col(2)}; }
template <class R> inline M3<R> diag(R a, R b, R c) { return {V3<R>{a, 0, 0}, V3<R>{0, b, 0}, V3<R>{0, 0, c}}; }
template <class R> inline M3<R> ident(R s) { return diag<R>(s, s, s); }
template <class R> inline M3<R> skew(V3<R> r) { return {V3<R>{0, -r.z, r.y}, V3<R>{r.z, 0, -r.x}, V3<R>{-r.y, r.x, 0}}; }
template <class R> inline M3<R> outer(V3<R> a, V3<R> b) { return {b * a.x, b * a.y, b * a.z}; }
template <class R> inline M3<R> diagonalize(const M3<R> &m) { return diag<R>(length(m.col(0)), length(m.col(1)), length(m.col(2))); }

// ---- quaternions (maths.h:239-277, 431-474) ----
template <class R> inline Q<R> operator*(Q<R> a, R s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class R> inline Q<R> operator/(Q<R> a, R s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }
template <class R> inline Q<R> operator+(Q<R> a, Q<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class R> inline Q<R> operator*(Q<R> a, Q<R> b)
{
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class R> inline R lengthSq(Q<R> q) { return lengthSq(V3<R>{q.x, q.y, q.z}) + q.w * q.w; }
template <class R> inline Q<R> normalize(Q<R> q) { return q / sqrt_(lengthSq(q)); }
template <class R> inline Q<R> conjugate(Q<R> q) { return {-q.x, -q.y, -q.z, q.w}; }
template <class R> inline Q<R> inverse(Q<R> q) { return conjugate(q) / lengthSq(q); }
// q - p = 2 vec(q p^-1)
template <class R> inline V3<R> operator-(Q<R> a, Q<R> b)
{
    Q<R> d = a * inverse(b);
    return V3<R>{d.x, d.y, d.z} * R(2);
}
// q + w = normalize(q + 1/2 (w,0) q)
template <class R> inline Q<R> operator+(Q<R> a, V3<R> b)
{
    return normalize(a + (Q<R>{b.x, b.y, b.z, 0} * a) * R(0.5));
}
template <class R> inline V3<R> rotate(Q<R> q, V3<R> v)
{
    V3<R> u{q.x, q.y, q.z};
    V3<R> t = cross(u, v) * R(2);
    return v + t * q.w + cross(u, t);
}
template <class R> inline V3<R> transform(V3<R> lin, Q<R> ang, V3<R> v) { return rotate(ang, v) + lin; }
template <class R> inline M3<R> orthonormal(V3<R> n)
{
    V3<R> t1 = abs_(n.x) > abs_(n.z) ? V3<R>{-n.y, n.x, 0} : V3<R>{0, -n.z, n.y};
    t1 = normalize(t1);
    V3<R> t2 = cross(n, t1);
    return {n, t1, t2};
}

// ---- 6x6 LDL^T solve of [[Lin, Cross^T],[Cross, Ang]] x = b (maths.h:481-547) ----
template <class R>
inline void solve6(const M3<R> &aLin, const M3<R> &aAng, const M3<R> &aCross, V3<R> bLin, V3<R> bAng, V3<R> &xLin, V3<R> &xAng)
{
    R A11 = aLin[0][0];
    R A21 = aLin[1][0], A22 = aLin[1][1];
    R A31 = aLin[2][0], A32 = aLin[2][1], A33 = aLin[2][2];
    R A41 = aCross[0][0], A42 = aCross[0][1], A43 = aCross[0][2], A44 = aAng[0][0];
    R A51 = aCross[1][0], A52 = aCross[1][1], A53 = aCross[1][2], A54 = aAng[1][0], A55 = aAng[1][1];
    R A61 = aCross[2][0], A62 = aCross[2][1], A63 = aCross[2][2], A64 = aAng[2][0], A65 = aAng[2][1], A66 = aAng[2][2];

    R D1 = A11;
    R L21 = A21 / D1, L31 = A31 / D1, L41 = A41 / D1, L51 = A51 / D1, L61 = A61 / D1;
    R D2 = A22 - L21 * L21 * D1;
    R L32 = (A32 - L21 * L31 * D1) / D2;
    R L42 = (A42 - L21 * L41 * D1) / D2;
    R L52 = (A52 - L21 * L51 * D1) / D2;
    R L62 = (A62 - L21 * L61 * D1) / D2;
    R D3 = A33 - (L31 * L31 * D1 + L32 * L32 * D2);
    R L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3;
    R L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3;
    R L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
    R D4 = A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3);
    R L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4;
    R L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
    R D5 = A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4);
    R L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
    R D6 = A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5);

    R y1 = bLin[0];
    R y2 = bLin[1] - L21 * y1;
    R y3 = bLin[2] - L31 * y1 - L32 * y2;
    R y4 = bAng[0] - L41 * y1 - L42 * y2 - L43 * y3;
    R y5 = bAng[1] - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
    R y6 = bAng[2] - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;

    R z1 = y1 / D1, z2 = y2 / D2, z3 = y3 / D3, z4 = y4 / D4, z5 = y5 / D5, z6 = y6 / D6;

    xAng[2] = z6;
    xAng[1] = z5 - L65 * xAng[2];
    xAng[0] = z4 - L54 * xAng[1] - L64 * xAng[2];
    xLin[2] = z3 - L43 * xAng[0] - L53 * xAng[1] - L63 * xAng[2];
    xLin[1] = z2 - L32 * xLin[2] - L42 * xAng[0] - L52 * xAng[1] - L62 * xAng[2];
    xLin[0] = z1 - L21 * xLin[1] - L31 * xLin[2] - L41 * xAng[0] - L51 * xAng[1] - L61 * xAng[2];
}

// Forward mode through the solve: the values are solve6's own (the scalar's value operations),
// the tangent is that of the exact solution x = H^-1 b at those values, dx = H^-1 (db - dH x),
// evaluated in FP64. Differentiating the FP32 elimination operation by operation instead would
// add the FP32 factors' rounding times the condition number of H (1e6 and more for stiff contacts)
// to the tangent; the adjoint's transposed solve is the exact H^-1 as well.
template <class V>
inline void solve6(const M3<DualT<V>> &aLin, const M3<DualT<V>> &aAng, const M3<DualT<V>> &aCross,
                   V3<DualT<V>> bLin, V3<DualT<V>> bAng, V3<DualT<V>> &xLin, V3<DualT<V>> &xAng)
{
    M3<V> vl, va, vc;
    M3<double> H[3], dH[3];  // lin, ang, cross in FP64: values and tangents
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
            vl[i][j] = aLin[i][j].v, va[i][j] = aAng[i][j].v, vc[i][j] = aCross[i][j].v;
            H[0][i][j] = (double)aLin[i][j].v, H[1][i][j] = (double)aAng[i][j].v, H[2][i][j] = (double)aCross[i][j].v;
            dH[0][i][j] = aLin[i][j].d, dH[1][i][j] = aAng[i][j].d, dH[2][i][j] = aCross[i][j].d;
        }
    V3<V> xl, xa;
    solve6<V>(vl, va, vc, V3<V>{bLin.x.v, bLin.y.v, bLin.z.v}, V3<V>{bAng.x.v, bAng.y.v, bAng.z.v}, xl, xa);
    // the symmetric 6x6 matrix from the lower triangle solve6 reads: rows 0-2 linear, 3-5 angular
    auto entry = [](const M3<double> *m, int r, int c) {
        if (c > r)
        {
            int t = r;
            r = c;
            c = t;
        }
        return r < 3 ? m[0][r][c] : (c < 3 ? m[2][r - 3][c] : m[1][r - 3][c - 3]);
    };
    const double x[6] = {(double)xl.x, (double)xl.y, (double)xl.z, (double)xa.x, (double)xa.y, (double)xa.z};
    const double db[6] = {bLin.x.d, bLin.y.d, bLin.z.d, bAng.x.d, bAng.y.d, bAng.z.d};
    double rhs[6];
    for (int r = 0; r < 6; ++r)
    {
        double acc = db[r];
        for (int c = 0; c < 6; ++c)
            acc -= entry(dH, r, c) * x[c];
        rhs[r] = acc;
    }
    V3<double> dxl, dxa;
    solve6<double>(H[0], H[1], H[2], V3<double>{rhs[0], rhs[1], rhs[2]}, V3<double>{rhs[3], rhs[4], rhs[5]}, dxl, dxa);
    xLin = V3<DualT<V>>{mk<V>(xl.x, dxl.x), mk<V>(xl.y, dxl.y), mk<V>(xl.z, dxl.z)};
    xAng = V3<DualT<V>>{mk<V>(xa.x, dxa.x), mk<V>(xa.y, dxa.y), mk<V>(xa.z, dxa.z)};
}

// ---- constants (solver.h:21-24, collide.cpp:18-22) ----
template <class R> struct K
{
    static constexpr R penaltyMin = R(1.0f);
    static constexpr R penaltyMax = R(10000000000.0f);
    static constexpr R margin = R(0.01f);
    static constexpr R stickThresh = R(0.00001f);
    static constexpr R satEps = R(1.0e-6f);
    static constexpr R planeEps = R(1.0e-5f);
    static constexpr R mergeDistSq = R(1.0e-6f);
};

// ---- state ----
template <class R>
struct Body
{
    V3<R> size;
    R mass;
    V3<R> moment;
    R friction;
    R radius;
    V3<R> pos;
    Q<R> rot;
    V3<R> initLin;
    Q<R> initAng;
    V3<R> inerLin;
    Q<R> inerAng;
    V3<R> vel;
    V3<R> angVel;
    V3<R> prevVel;
    R shapeRadius; // RigidShape::radius
    R halfLength;  // RigidShape::halfLength
    int shape;
    int denseId;
};

template <class R>
struct Contact
{
    int32_t feature;
    V3<R> rA, rB, C0, penalty, lambda;
    bool stick;
};

template <class R>
struct Manifold
{
    int a, b;
    int64_t serial;
    int n;
    R friction;
    M3<R> basis;
    Contact<R> c[8];
};

// One entry of the external contact table (ExternalManifoldContact, solver.h:336-343), by body
// list index (-1: no body).
template <class R>
struct ExternalContact
{
    int a, b, n;
    M3<R> basis;
    Contact<R> c[8];
};

template <class R>
struct Joint
{
    int a, b; // a == -1: world anchor
    int64_t serial;
    V3<R> rA, rB, C0Lin, C0Ang, penLin, penAng, lamLin, lamAng;
    R kLin, kAng, fracture, torqueArm;
    bool broken;
};

template <class R>
struct Spring
{
    int a, b;
    int64_t serial;
    V3<R> rA, rB;
    R rest, k;
};

template <class R>
struct Params
{
    R dt, gravity;
    int iterations;
    R alpha, betaLin, betaAng, gamma;
};

enum ForceKind
{
    KIND_MANIFOLD = 0,
    KIND_JOINT = 1,
    KIND_SPRING = 2
};

struct ForceRef
{
    int kind;
    int index;
    int64_t serial;
};

struct Hit
{
    int a, b;
    int created;
};

// Solver::broadphaseMode (solver.h:39-45): all pairs, uniform grid (the reference's default,
// solver.cpp:676-680) or sweep and prune.
enum BroadphaseMode
{
    BP_ALL_PAIRS = 0,
    BP_GRID = 1,
    BP_SAP = 2
};

enum OrderMode
{
    ORDER_COLOR = 0,
    ORDER_LIST = 1
};

// ---- box-box narrowphase (collide.cpp:18-433, 898-955) ----
template <class R>
struct Obb
{
    V3<R> c;
    V3<R> h;
    V3<R> ax[3];
};

template <class R>
inline Obb<R> makeObb(const Body<R> &b)
{
    Obb<R> o;
    o.c = b.pos;
    o.h = b.size * R(0.5);
    o.ax[0] = rotate(b.rot, V3<R>{1, 0, 0});
    o.ax[1] = rotate(b.rot, V3<R>{0, 1, 0});
    o.ax[2] = rotate(b.rot, V3<R>{0, 0, 1});
    return o;
}

template <class R>
inline V3<R> supportPt(const Obb<R> &o, V3<R> d)
{
    R sx = dot(d, o.ax[0]) >= 0 ? R(1) : R(-1);
    R sy = dot(d, o.ax[1]) >= 0 ? R(1) : R(-1);
    R sz = dot(d, o.ax[2]) >= 0 ? R(1) : R(-1);
    return o.c + o.ax[0] * (o.h.x * sx) + o.ax[1] * (o.h.y * sy) + o.ax[2] * (o.h.z * sz);
}

template <class R>
inline void faceAxes(const Obb<R> &o, int k, V3<R> &u, V3<R> &v, R &eu, R &ev)
{
    int iu = k == 0 ? 1 : 0;
    int iv = k == 2 ? 1 : 2;
    u = o.ax[iu];
    v = o.ax[iv];
    eu = o.h[iu];
    ev = o.h[iv];
}

template <class R>
inline int clipPlane(const V3<R> *in, int n, V3<R> pn, R off, V3<R> *out)
{
    if (n <= 0)
        return 0;
    int m = 0;
    V3<R> a = in[n - 1];
    R da = dot(pn, a) - off;
    for (int i = 0; i < n; ++i)
    {
        V3<R> b = in[i];
        R db = dot(pn, b) - off;
        bool ain = da <= K<R>::planeEps;
        bool bin = db <= K<R>::planeEps;
        if (ain != bin)
        {
            R t = 0;
            R den = da - db;
            if (abs_(den) > K<R>::satEps)
                t = clampv(da / den, R(0), R(1));
            if (m < 16)
                out[m++] = a + (b - a) * t;
        }
        if (bin && m < 16)
            out[m++] = b;
        a = b;
        da = db;
    }
    return m;
}

template <class R>
struct ContactBuilder
{
    const Body<R> &A;
    const Body<R> &B;
    Contact<R> *out;
    int count = 0;
    V3<R> mids[8];

    ContactBuilder(const Body<R> &a, const Body<R> &b, Contact<R> *o) : A(a), B(b), out(o) {}

    void add(V3<R> xA, V3<R> xB, int32_t key)
    {
        V3<R> mid = (xA + xB) * R(0.5);
        for (int i = 0; i < count; ++i)
            if (lengthSq(mid - mids[i]) < K<R>::mergeDistSq)
                return;
        if (count >= 8)
            return;
        Contact<R> &c = out[count];
        c.feature = key;
        c.rA = rotate(conjugate(A.rot), xA - A.pos);
        c.rB = rotate(conjugate(B.rot), xB - B.pos);
        mids[count] = mid;
        ++count;
    }
};

template <class R>
struct SatBest
{
    int type; // 0 face A, 1 face B, 2 edge
    int ia, ib;
    R sep;
    V3<R> n;
    bool valid;
};

// Returns false when the axis separates the boxes.
template <class R>
inline bool satAxis(const Obb<R> &A, const Obb<R> &B, V3<R> delta, V3<R> axis, int type, int ia, int ib, SatBest<R> &best)
{
    R l2 = lengthSq(axis);
    if (l2 < K<R>::satEps)
        return true;
    R inv = R(1) / sqrt_(l2);
    V3<R> n = axis * inv;
    if (dot(n, delta) < 0)
        n = -n;
    R dist = abs_(dot(delta, n));
    R rA = A.h.x * abs_(dot(n, A.ax[0])) + A.h.y * abs_(dot(n, A.ax[1])) + A.h.z * abs_(dot(n, A.ax[2]));
    R rB = B.h.x * abs_(dot(n, B.ax[0])) + B.h.y * abs_(dot(n, B.ax[1])) + B.h.z * abs_(dot(n, B.ax[2]));
    R sep = dist - (rA + rB);
    if (sep > 0)
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

template <class R>
inline void closestSegPts(V3<R> p0, V3<R> p1, V3<R> q0, V3<R> q1, V3<R> &c0, V3<R> &c1)
{
    V3<R> d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
    R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    R s = 0, t = 0;
    if (a <= K<R>::satEps && e <= K<R>::satEps)
    {
        c0 = p0;
        c1 = q0;
        return;
    }
    if (a <= K<R>::satEps)
    {
        t = clampv(f / e, R(0), R(1));
    }
    else
    {
        R c = dot(d1, r);
        if (e <= K<R>::satEps)
        {
            s = clampv(-c / a, R(0), R(1));
        }
        else
        {
            R b = dot(d1, d2);
            R den = a * e - b * b;
            if (abs_(den) > K<R>::satEps)
                s = clampv((b * f - c * e) / den, R(0), R(1));
            t = (b * s + f) / e;
            if (t < 0)
            {
                t = 0;
                s = clampv(-c / a, R(0), R(1));
            }
            else if (t > 1)
            {
                t = 1;
                s = clampv((b - c) / a, R(0), R(1));
            }
        }
    }
    c0 = p0 + d1 * s;
    c1 = q0 + d2 * t;
}

template <class R>
inline void edgeSupport(const Obb<R> &o, int k, V3<R> d, V3<R> &e0, V3<R> &e1)
{
    int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    R s1 = dot(d, o.ax[k1]) >= 0 ? R(1) : R(-1);
    R s2 = dot(d, o.ax[k2]) >= 0 ? R(1) : R(-1);
    V3<R> mid = o.c + o.ax[k1] * (o.h[k1] * s1) + o.ax[k2] * (o.h[k2] * s2);
    e0 = mid - o.ax[k] * o.h[k];
    e1 = mid + o.ax[k] * o.h[k];
}

template <class R>
inline int faceManifold(ContactBuilder<R> &cb, const Obb<R> &A, const Obb<R> &B, bool refA, int refAxis, V3<R> nAB)
{
    const Obb<R> &ref = refA ? A : B;
    const Obb<R> &inc = refA ? B : A;
    V3<R> outward = refA ? nAB : -nAB;

    // reference face frame (collide.cpp:110-117)
    R fs = dot(outward, ref.ax[refAxis]) >= 0 ? R(1) : R(-1);
    V3<R> fn = ref.ax[refAxis] * fs;
    V3<R> fc = ref.c + fn * ref.h[refAxis];
    V3<R> fu, fv;
    R eu, ev;
    faceAxes(ref, refAxis, fu, fv, eu, ev);

    // incident face: most anti-parallel axis (collide.cpp:119-135)
    int incAxis = 0;
    R best = -FLT_MAX;
    for (int i = 0; i < 3; ++i)
    {
        R d = abs_(dot(inc.ax[i], fn));
        if (d > best)
        {
            best = d;
            incAxis = i;
        }
    }
    V3<R> p0[16], p1[16];
    {
        R s = dot(inc.ax[incAxis], fn) > 0 ? R(-1) : R(1);
        V3<R> n = inc.ax[incAxis] * s;
        V3<R> c = inc.c + n * inc.h[incAxis];
        V3<R> u, v;
        R ueu, vev;
        faceAxes(inc, incAxis, u, v, ueu, vev);
        p0[0] = c + u * ueu + v * vev;
        p0[1] = c - u * ueu + v * vev;
        p0[2] = c - u * ueu - v * vev;
        p0[3] = c + u * ueu - v * vev;
    }
    int n = 4;
    n = clipPlane(p0, n, fu, dot(fu, fc) + eu, p1);
    if (!n)
        return 0;
    V3<R> nu = -fu;
    n = clipPlane(p1, n, nu, dot(nu, fc) + eu, p0);
    if (!n)
        return 0;
    n = clipPlane(p0, n, fv, dot(fv, fc) + ev, p1);
    if (!n)
        return 0;
    V3<R> nv = -fv;
    n = clipPlane(p1, n, nv, dot(nv, fc) + ev, p0);
    if (!n)
        return 0;

    int32_t prefix = ((refA ? 0 : 1) << 24) | ((refAxis & 0xFF) << 16) | ((incAxis & 0xFF) << 8);
    for (int i = 0; i < n && cb.count < 8; ++i)
    {
        V3<R> pi = p0[i];
        R depth = dot(pi - fc, fn);
        if (depth > K<R>::planeEps)
            continue;
        V3<R> pr = pi - fn * depth;
        cb.add(refA ? pr : pi, refA ? pi : pr, prefix | (i & 0xFF));
    }
    if (!cb.count)
        cb.add(supportPt(A, nAB), supportPt(B, -nAB), prefix);
    return cb.count;
}

template <class R>
inline int collideBoxes(const Body<R> &bA, const Body<R> &bB, Contact<R> *out, M3<R> &basis)
{
    Obb<R> A = makeObb(bA), B = makeObb(bB);
    V3<R> delta = B.c - A.c;
    SatBest<R> face{}, edge{};
    face.sep = -FLT_MAX;
    face.valid = false;
    edge.sep = -FLT_MAX;
    edge.valid = false;
    for (int i = 0; i < 3; ++i)
        if (!satAxis(A, B, delta, A.ax[i], 0, i, -1, face))
            return 0;
    for (int i = 0; i < 3; ++i)
        if (!satAxis(A, B, delta, B.ax[i], 1, -1, i, face))
            return 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (!satAxis(A, B, delta, cross(A.ax[i], B.ax[j]), 2, i, j, edge))
                return 0;
    if (!face.valid)
        return 0;
    SatBest<R> best = face;
    if (edge.valid && R(0.95f) * edge.sep > face.sep + R(0.01f))
        best = edge;
    basis = orthonormal(-best.n);

    ContactBuilder<R> cb(bA, bB, out);
    if (best.type == 2)
    {
        V3<R> a0, a1, b0, b1, xA, xB;
        edgeSupport(A, best.ia, best.n, a0, a1);
        edgeSupport(B, best.ib, -best.n, b0, b1);
        closestSegPts(a0, a1, b0, b1, xA, xB);
        int32_t key = (2 << 24) | ((best.ia & 0xFF) << 8) | (best.ib & 0xFF);
        cb.add(xA, xB, key);
        if (!cb.count)
            cb.add(supportPt(A, best.n), supportPt(B, -best.n), key);
        return cb.count;
    }
    if (best.type == 0)
        return faceManifold(cb, A, B, true, best.ia, best.n);
    return faceManifold(cb, A, B, false, best.ib, best.n);
}

// ---- round shapes: spheres, capsules, cylinders (collide.cpp:435-1181) ----
// Shape codes of RigidShapeType (solver.h:47-53). A capsule's and a cylinder's core segment
// runs along the body's local z axis, halfLength either side of its centre.
enum ShapeKind
{
    SHAPE_BOX = 0,
    SHAPE_SPHERE = 1,
    SHAPE_CAPSULE = 2,
    SHAPE_CYLINDER = 3
};

// Branch-hit counters of the round-shape narrowphase and the rolling pass (test
// instrumentation: which code paths a scene exercises).
enum RoundBranch
{
    RB_THIN_TOP_ROUND = 0,   // sphere/capsule on a thin static top
    RB_ROUND_BOX_NEAR,       // centre outside the box, within the radius
    RB_ROUND_BOX_TOP_POINT,  // thin-top point after the top contacts missed
    RB_INSIDE_SWEPT,         // centre inside, swept sphere test hit
    RB_INSIDE_STATIC,        // centre inside a static box: its up face
    RB_INSIDE_MOTION,        // centre inside: face against the relative motion
    RB_INSIDE_NEAREST,       // centre inside, no relative motion: nearest face
    RB_OUTSIDE_SWEPT,        // centre beyond the radius, swept sphere test hit
    RB_CYL_CAPS,             // coaxial cylinders, caps facing
    RB_CYL_SIDE,             // parallel cylinders side by side
    RB_CYL_CROSSED,          // non-parallel cylinders
    RB_CYL_TOP_RIM,          // standing cylinder on a thin static top
    RB_CYL_TOP_SIDE,         // lying cylinder on a thin static top
    RB_CYL_BOX,              // cylinder collided as its bounding box
    RB_ROLLING,              // rolling friction applied to a sphere contact
    RB_ROUND_ROUND,          // sphere/capsule pairs and cylinder-round pairs
    RB_COUNT
};
inline long long roundBranchHits[RB_COUNT] = {};
inline void hitBranch(int b) { ++roundBranchHits[b]; }

// collide.cpp:458-466: the narrowphase's "round" family is spheres and capsules; cylinders are
// dispatched on their own.
template <class R> inline bool sphereOrCapsule(const Body<R> &b) { return b.shape == SHAPE_SPHERE || b.shape == SHAPE_CAPSULE; }
template <class R> inline V3<R> localZ(const Body<R> &b) { return rotate(b.rot, V3<R>{0, 0, 1}); }

// Core segment of a capsule (a sphere's is its centre twice), collide.cpp:468-481.
template <class R> inline void roundCore(const Body<R> &b, V3<R> &s0, V3<R> &s1)
{
    if (b.shape == SHAPE_CAPSULE)
    {
        V3<R> z = localZ(b);
        s0 = b.pos - z * b.halfLength;
        s1 = b.pos + z * b.halfLength;
        return;
    }
    s0 = b.pos;
    s1 = b.pos;
}
// collide.cpp:483-488
template <class R> inline void cylinderCore(const Body<R> &b, V3<R> &s0, V3<R> &s1)
{
    V3<R> z = localZ(b);
    s0 = b.pos - z * b.halfLength;
    s1 = b.pos + z * b.halfLength;
}

// Box-frame helpers (collide.cpp:435-456).
template <class R> inline V3<R> obbLocal(const Obb<R> &o, V3<R> p)
{
    V3<R> d = p - o.c;
    return V3<R>{dot(d, o.ax[0]), dot(d, o.ax[1]), dot(d, o.ax[2])};
}
template <class R> inline V3<R> obbWorld(const Obb<R> &o, V3<R> l) { return o.c + o.ax[0] * l.x + o.ax[1] * l.y + o.ax[2] * l.z; }
template <class R> inline V3<R> obbClosest(const Obb<R> &o, V3<R> p)
{
    V3<R> d = p - o.c;
    return o.c + o.ax[0] * clampv(dot(d, o.ax[0]), -o.h.x, o.h.x) + o.ax[1] * clampv(dot(d, o.ax[1]), -o.h.y, o.h.y) +
           o.ax[2] * clampv(dot(d, o.ax[2]), -o.h.z, o.h.z);
}

// collide.cpp:490-498
template <class R> inline V3<R> segClosest(V3<R> a, V3<R> b, V3<R> p)
{
    V3<R> ab = b - a;
    R den = lengthSq(ab);
    if (den < K<R>::satEps)
        return a;
    R t = clampv(dot(p - a, ab) / den, R(0), R(1));
    return a + ab * t;
}

// Smallest half extent (ties to the lower axis) of a static box that is thin enough to count as a
// floor/slab, and the face of that axis pointing up (+z); collide.cpp:957-976.
template <class R> inline bool thinTop(const Obb<R> &o, const Body<R> &box, int &axis, R &sign, V3<R> &normal)
{
    if (box.mass > 0)
        return false;
    axis = 0;
    if (o.h.y < o.h[axis])
        axis = 1;
    if (o.h.z < o.h[axis])
        axis = 2;
    int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    if (o.h[axis] * R(4.0f) > fmin_(o.h[a1], o.h[a2]))
        return false;
    sign = dot(o.ax[axis], V3<R>{0, 0, 1}) >= 0 ? R(1) : R(-1);
    normal = o.ax[axis] * sign;
    return true;
}

// Local point clamped onto the top face (all but the thin axis clamped), after a proximity test
// against the face grown by `grow` in-plane; false if out of reach. collide.cpp:687-715.
template <class R> inline bool thinTopFoot(const Obb<R> &o, int axis, V3<R> &local, R grow)
{
    for (int k = 0; k < 3; ++k)
    {
        if (k == axis)
            continue;
        if (local[k] < -o.h[k] - grow || local[k] > o.h[k] + grow)
            return false;
    }
    return true;
}
template <class R> inline void clampOthers(const Obb<R> &o, int axis, V3<R> &local)
{
    for (int k = 0; k < 3; ++k)
        if (k != axis)
            local[k] = clampv(local[k], -o.h[k], o.h[k]);
}

// collide.cpp:500-517: contact between two spheres centred on the given points.
template <class R>
inline int roundPair(const Body<R> &bA, const Body<R> &bB, V3<R> cA, R rA, V3<R> cB, R rB, int32_t key, Contact<R> *out,
                     M3<R> &basis)
{
    V3<R> delta = cB - cA;
    R rs = rA + rB;
    R d2 = lengthSq(delta);
    if (d2 > rs * rs)
        return 0;
    V3<R> n = d2 > K<R>::satEps ? delta / sqrt_(d2) : V3<R>{1, 0, 0};
    V3<R> xA = cA + n * rA;
    V3<R> xB = cB - n * rB;
    basis = orthonormal(-n);
    ContactBuilder<R> cb(bA, bB, out);
    cb.add(xA, xB, key);
    return cb.count;
}

// Swept test of a sphere's step-start motion against a box grown by its radius (slab method),
// collide.cpp:519-587. On a hit: the face normal and the clamped surface point.
template <class R>
inline bool sweptSphere(const Obb<R> &o, const Body<R> &sph, const Body<R> &box, R dt, V3<R> &n, V3<R> &xBox)
{
    V3<R> rel = sph.vel - box.vel;
    if (dt <= 0 || lengthSq(rel) < K<R>::satEps)
        return false;
    V3<R> cur = sph.pos;
    V3<R> prev = cur - rel * dt;
    V3<R> l0 = obbLocal(o, prev), l1 = obbLocal(o, cur);
    V3<R> dl = l1 - l0;
    V3<R> grown = o.h + V3<R>{sph.shapeRadius, sph.shapeRadius, sph.shapeRadius};
    R tIn = 0, tOut = 1;
    int hit = -1;
    R hitSign = 1;
    for (int k = 0; k < 3; ++k)
    {
        if (abs_(dl[k]) < K<R>::satEps)
        {
            if (l0[k] < -grown[k] || l0[k] > grown[k])
                return false;
            continue;
        }
        R inv = R(1) / dl[k];
        R t0 = (-grown[k] - l0[k]) * inv;
        R t1 = (grown[k] - l0[k]) * inv;
        R s = -1;
        if (t0 > t1)
        {
            R tmp = t0;
            t0 = t1;
            t1 = tmp;
            s = 1;
        }
        if (t0 > tIn)
        {
            tIn = t0;
            hit = k;
            hitSign = s;
        }
        tOut = fmin_(tOut, t1);
        if (tIn > tOut)
            return false;
    }
    if (hit < 0 || tIn < 0 || tIn > 1)
        return false;
    V3<R> surf = l0 + dl * tIn;
    surf[hit] = o.h[hit] * hitSign;
    clampOthers(o, hit, surf);
    n = o.ax[hit] * hitSign;
    xBox = obbWorld(o, surf);
    return true;
}

// The point of a sphere/capsule nearest a thin static box's top, and its foot on the top face;
// collide.cpp:632-685.
template <class R>
inline bool thinTopRoundPoint(const Obb<R> &o, const Body<R> &rb, const Body<R> &box, V3<R> &centre, V3<R> &n, V3<R> &xBox)
{
    int axis;
    R sign;
    V3<R> normal;
    if (!thinTop(o, box, axis, sign, normal))
        return false;
    V3<R> s0, s1;
    roundCore(rb, s0, s1);
    R d0 = dot(s0, normal), d1 = dot(s1, normal);
    if (rb.shape == SHAPE_CAPSULE && abs_(d0 - d1) < R(1.0e-4f))
        centre = segClosest(s0, s1, rb.pos);
    else
        centre = d0 <= d1 ? s0 : s1;
    V3<R> local = obbLocal(o, centre);
    if (local[axis] * sign - o.h[axis] > rb.shapeRadius)
        return false;
    if (!thinTopFoot(o, axis, local, rb.shapeRadius))
        return false;
    local[axis] = o.h[axis] * sign;
    clampOthers(o, axis, local);
    n = normal;
    xBox = obbWorld(o, local);
    return true;
}

// One sphere-cap contact against a thin static box top; collide.cpp:687-715.
template <class R>
inline void thinTopRoundContact(ContactBuilder<R> &cb, const Obb<R> &o, const Body<R> &rb, bool roundIsA, int axis, R sign,
                                V3<R> normal, V3<R> centre, int32_t key)
{
    V3<R> local = obbLocal(o, centre);
    if (local[axis] * sign - o.h[axis] > rb.shapeRadius)
        return;
    if (!thinTopFoot(o, axis, local, rb.shapeRadius))
        return;
    local[axis] = o.h[axis] * sign;
    clampOthers(o, axis, local);
    V3<R> xRound = centre - normal * rb.shapeRadius;
    V3<R> xBox = obbWorld(o, local);
    cb.add(roundIsA ? xRound : xBox, roundIsA ? xBox : xRound, key);
}

// collide.cpp:717-754: -1 when the box is not a thin static top or nothing touches it.
template <class R>
inline int thinTopRound(const Obb<R> &o, const Body<R> &rb, const Body<R> &box, bool roundIsA, const Body<R> &bA,
                        const Body<R> &bB, Contact<R> *out, M3<R> &basis)
{
    int axis;
    R sign;
    V3<R> normal;
    if (!thinTop(o, box, axis, sign, normal))
        return -1;
    V3<R> nAB = roundIsA ? -normal : normal;
    basis = orthonormal(-nAB);
    ContactBuilder<R> cb(bA, bB, out);
    if (rb.shape == SHAPE_CAPSULE)
    {
        V3<R> s0, s1;
        roundCore(rb, s0, s1);
        thinTopRoundContact(cb, o, rb, roundIsA, axis, sign, normal, s0, 0x04000000);
        thinTopRoundContact(cb, o, rb, roundIsA, axis, sign, normal, s1, 0x04000001);
    }
    else
    {
        thinTopRoundContact(cb, o, rb, roundIsA, axis, sign, normal, rb.pos, 0x02000000);
    }
    if (cb.count > 0)
        hitBranch(RB_THIN_TOP_ROUND);
    return cb.count > 0 ? cb.count : -1;
}

// Sphere/capsule against a box; collide.cpp:767-896.
template <class R>
inline int collideRoundBox(const Body<R> &rb, const Body<R> &box, bool roundIsA, R dt, Contact<R> *out, M3<R> &basis)
{
    const Body<R> &bA = roundIsA ? rb : box;
    const Body<R> &bB = roundIsA ? box : rb;
    Obb<R> o = makeObb(box);
    int top = thinTopRound(o, rb, box, roundIsA, bA, bB, out, basis);
    if (top >= 0)
        return top;
    V3<R> centre = rb.pos;
    if (rb.shape == SHAPE_CAPSULE)
    {
        V3<R> s0, s1;
        roundCore(rb, s0, s1);
        V3<R> c0 = obbClosest(o, s0), c1 = obbClosest(o, s1);
        centre = lengthSq(s0 - c0) <= lengthSq(s1 - c1) ? s0 : s1;
    }
    R r = rb.shapeRadius;
    V3<R> closest = obbClosest(o, centre);
    V3<R> toRound = centre - closest;
    R d2 = lengthSq(toRound);
    V3<R> tc, tn, tp;
    if (thinTopRoundPoint(o, rb, box, tc, tn, tp))
    {
        hitBranch(RB_ROUND_BOX_TOP_POINT);
        centre = tc;
        toRound = tn;
        closest = tp;
        d2 = 1;
    }
    else
    {
        V3<R> local = obbLocal(o, centre);
        bool inside = abs_(local.x) <= o.h.x && abs_(local.y) <= o.h.y && abs_(local.z) <= o.h.z;
        V3<R> sn, sp;
        if (inside)
        {
            if (rb.shape == SHAPE_SPHERE && sweptSphere(o, rb, box, dt, sn, sp))
            {
                hitBranch(RB_INSIDE_SWEPT);
                toRound = sn;
                closest = sp;
                d2 = 1;
            }
            else
            {
                // push-out face: the floor's top for a static box, else the face most against the
                // relative motion, else the nearest face (collide.cpp:827-873)
                int axis = 0;
                R sign = local.x >= 0 ? R(1) : R(-1);
                R best = o.h.x - abs_(local.x);
                V3<R> rel = rb.vel - box.vel;
                if (box.mass <= 0)
                {
                    hitBranch(RB_INSIDE_STATIC);
                    axis = 0;
                    if (o.h.y < o.h[axis])
                        axis = 1;
                    if (o.h.z < o.h[axis])
                        axis = 2;
                    sign = dot(o.ax[axis], V3<R>{0, 0, 1}) >= 0 ? R(1) : R(-1);
                }
                else if (lengthSq(rel) > K<R>::satEps)
                {
                    hitBranch(RB_INSIDE_MOTION);
                    V3<R> against = -rel;
                    R bestDot = R(-FLT_MAX);
                    for (int k = 0; k < 3; ++k)
                        for (int si = 0; si < 2; ++si)
                        {
                            R s = si == 0 ? R(-1) : R(1);
                            R d = dot(o.ax[k] * s, against);
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
                    hitBranch(RB_INSIDE_NEAREST);
                    R dy = o.h.y - abs_(local.y);
                    if (dy < best)
                    {
                        axis = 1;
                        sign = local.y >= 0 ? R(1) : R(-1);
                        best = dy;
                    }
                    R dz = o.h.z - abs_(local.z);
                    if (dz < best)
                    {
                        axis = 2;
                        sign = local.z >= 0 ? R(1) : R(-1);
                    }
                }
                V3<R> surf = local;
                surf[axis] = o.h[axis] * sign;
                closest = obbWorld(o, surf);
                toRound = o.ax[axis] * sign;
                d2 = 1;
            }
        }
        else if (d2 <= r * r)
            hitBranch(RB_ROUND_BOX_NEAR);
        else if (d2 > r * r)
        {
            if (rb.shape != SHAPE_SPHERE || !sweptSphere(o, rb, box, dt, sn, sp))
                return 0;
            hitBranch(RB_OUTSIDE_SWEPT);
            toRound = sn;
            closest = sp;
            d2 = 1;
        }
    }
    V3<R> n = d2 > K<R>::satEps ? toRound / sqrt_(d2) : V3<R>{0, 0, 1};
    V3<R> nAB = roundIsA ? -n : n;
    V3<R> xRound = centre - n * r;
    basis = orthonormal(-nAB);
    ContactBuilder<R> cb(bA, bB, out);
    cb.add(roundIsA ? xRound : closest, roundIsA ? closest : xRound, rb.shape == SHAPE_CAPSULE ? 0x04000000 : 0x02000000);
    return cb.count;
}

// Rim point of a thin static box top against a cylinder; collide.cpp:978-1007.
template <class R>
inline void thinTopCylinderContact(ContactBuilder<R> &cb, const Obb<R> &o, const Body<R> &cyl, bool cylIsA, int axis, R sign,
                                   V3<R> xCyl, int32_t key)
{
    V3<R> local = obbLocal(o, xCyl);
    if (local[axis] * sign - o.h[axis] > K<R>::margin)
        return;
    if (!thinTopFoot(o, axis, local, cyl.shapeRadius))
        return;
    local[axis] = (o.h[axis] + K<R>::margin) * sign;
    clampOthers(o, axis, local);
    V3<R> xBox = obbWorld(o, local);
    cb.add(cylIsA ? xCyl : xBox, cylIsA ? xBox : xCyl, key);
}

// Unit vector perpendicular to n (orthonormal's first tangent, normalised on its own).
template <class R> inline V3<R> perpUnit(V3<R> n)
{
    return abs_(n.x) > abs_(n.z) ? normalize(V3<R>{-n.y, n.x, 0}) : normalize(V3<R>{0, -n.z, n.y});
}

// collide.cpp:1009-1053: a standing cylinder's cap rim (4 points) or a lying one's two base
// rims against a thin static top; -1 when that does not apply.
template <class R>
inline int thinTopCylinder(const Obb<R> &o, const Body<R> &cyl, const Body<R> &box, bool cylIsA, Contact<R> *out, M3<R> &basis)
{
    int axis;
    R sign;
    V3<R> normal;
    if (!thinTop(o, box, axis, sign, normal))
        return -1;
    V3<R> nAB = cylIsA ? -normal : normal;
    basis = orthonormal(-nAB);
    const Body<R> &bA = cylIsA ? cyl : box;
    const Body<R> &bB = cylIsA ? box : cyl;
    ContactBuilder<R> cb(bA, bB, out);
    V3<R> z = localZ(cyl);
    R zd = dot(z, normal);
    R r = cyl.shapeRadius;
    if (abs_(zd) > R(0.65f))
    {
        R zs = zd >= 0 ? R(1) : R(-1);
        V3<R> cap = cyl.pos - z * (cyl.halfLength * zs);
        V3<R> u = perpUnit(z);
        V3<R> v = normalize(cross(z, u));
        V3<R> rim[4] = {u * r, -u * r, v * r, -v * r};
        for (int i = 0; i < 4; ++i)
            thinTopCylinderContact(cb, o, cyl, cylIsA, axis, sign, cap + rim[i], 0x05000000 | i);
        if (cb.count)
            hitBranch(RB_CYL_TOP_RIM);
    }
    else
    {
        V3<R> radial = normal - z * zd;
        if (lengthSq(radial) < K<R>::satEps)
            radial = normal;
        radial = normalize(radial);
        V3<R> e0 = cyl.pos - z * cyl.halfLength;
        V3<R> e1 = cyl.pos + z * cyl.halfLength;
        thinTopCylinderContact(cb, o, cyl, cylIsA, axis, sign, e0 - radial * r, 0x05000100);
        thinTopCylinderContact(cb, o, cyl, cylIsA, axis, sign, e1 - radial * r, 0x05000101);
        if (cb.count)
            hitBranch(RB_CYL_TOP_SIDE);
    }
    return cb.count > 0 ? cb.count : -1;
}

// Two cylinders; collide.cpp:1080-1152.
template <class R>
inline int collideCylinders(const Body<R> &bA, const Body<R> &bB, Contact<R> *out, M3<R> &basis)
{
    V3<R> zA = localZ(bA);
    V3<R> zB = localZ(bB);
    R par = dot(zA, zB);
    V3<R> delta = bB.pos - bA.pos;
    if (abs_(par) > R(0.95f))
    {
        R axial = dot(delta, zA);
        V3<R> radial = delta - zA * axial;
        R rl2 = lengthSq(radial);
        R rs = bA.shapeRadius + bB.shapeRadius;
        if (rl2 > rs * rs)
            return 0;
        if (bA.halfLength + bB.halfLength - abs_(axial) < -K<R>::margin)
            return 0;
        ContactBuilder<R> cb(bA, bB, out);
        if (abs_(axial) > (bA.halfLength + bB.halfLength) * R(0.5f))
        {
            // end caps facing each other: four rim points of the smaller cap
            R side = axial >= 0 ? R(1) : R(-1);
            V3<R> n = zA * side;
            V3<R> capA = bA.pos + n * bA.halfLength;
            V3<R> capB = bB.pos - n * bB.halfLength;
            hitBranch(RB_CYL_CAPS);
            R rr = fmin_(bA.shapeRadius, bB.shapeRadius) * R(0.85f);
            basis = orthonormal(-n);
            V3<R> u = perpUnit(n);
            V3<R> v = normalize(cross(n, u));
            V3<R> rim[4] = {u * rr, -u * rr, v * rr, -v * rr};
            for (int i = 0; i < 4; ++i)
                cb.add(capA + rim[i], capB + rim[i], 0x06000000 | i);
            return cb.count;
        }
        // side by side: the two ends of the axial overlap
        hitBranch(RB_CYL_SIDE);
        V3<R> n = rl2 > K<R>::satEps ? radial / sqrt_(rl2) : orthonormal(zA)[1];
        basis = orthonormal(-n);
        R ends[2] = {fmax_(-bA.halfLength, axial - bB.halfLength), fmin_(bA.halfLength, axial + bB.halfLength)};
        for (int i = 0; i < 2; ++i)
        {
            R tA = ends[i];
            R tB = tA - axial;
            V3<R> xA = bA.pos + zA * tA + n * bA.shapeRadius;
            V3<R> xB = bB.pos + zA * tB - n * bB.shapeRadius;
            cb.add(xA, xB, 0x06000100 | i);
        }
        return cb.count;
    }
    V3<R> a0, a1, b0, b1, cA, cB;
    cylinderCore(bA, a0, a1);
    cylinderCore(bB, b0, b1);
    closestSegPts(a0, a1, b0, b1, cA, cB);
    hitBranch(RB_CYL_CROSSED);
    return roundPair(bA, bB, cA, bA.shapeRadius, cB, bB.shapeRadius, 0x06000200, out, basis);
}

// Manifold::collide (collide.cpp:1154-1181): the pair dispatch. A cylinder against a box that is
// not a thin static top is collided as the box of its bounding size (Rigid::size).
template <class R>
inline int collidePair(const Body<R> &bA, const Body<R> &bB, R dt, Contact<R> *out, M3<R> &basis)
{
    bool cylA = bA.shape == SHAPE_CYLINDER, cylB = bB.shape == SHAPE_CYLINDER;
    bool rndA = sphereOrCapsule(bA), rndB = sphereOrCapsule(bB);
    if (cylA && cylB)
        return collideCylinders(bA, bB, out, basis);
    if ((cylA && rndB) || (rndA && cylB))
    {
        // collide.cpp:1065-1078
        const Body<R> &cyl = cylA ? bA : bB;
        const Body<R> &rb = cylA ? bB : bA;
        V3<R> c0, c1, r0, r1, cc, cr;
        cylinderCore(cyl, c0, c1);
        roundCore(rb, r0, r1);
        closestSegPts(c0, c1, r0, r1, cc, cr);
        hitBranch(RB_ROUND_ROUND);
        if (cylA)
            return roundPair(bA, bB, cc, cyl.shapeRadius, cr, rb.shapeRadius, 0x07000000, out, basis);
        return roundPair(bA, bB, cr, rb.shapeRadius, cc, cyl.shapeRadius, 0x07000000, out, basis);
    }
    if ((cylA && bB.shape == SHAPE_BOX) || (bA.shape == SHAPE_BOX && cylB))
    {
        const Body<R> &cyl = cylA ? bA : bB;
        const Body<R> &box = cylA ? bB : bA;
        Obb<R> o = makeObb(box);
        int top = thinTopCylinder(o, cyl, box, cylA, out, basis);
        if (top >= 0)
            return top;
        hitBranch(RB_CYL_BOX);
        return collideBoxes(bA, bB, out, basis);
    }
    if (rndA && rndB)
    {
        // collide.cpp:756-765
        V3<R> a0, a1, b0, b1, cA, cB;
        roundCore(bA, a0, a1);
        roundCore(bB, b0, b1);
        closestSegPts(a0, a1, b0, b1, cA, cB);
        hitBranch(RB_ROUND_ROUND);
        return roundPair(bA, bB, cA, bA.shapeRadius, cB, bB.shapeRadius, 0x03000000, out, basis);
    }
    if (rndA && bB.shape == SHAPE_BOX)
        return collideRoundBox(bA, bB, true, dt, out, basis);
    if (bA.shape == SHAPE_BOX && rndB)
        return collideRoundBox(bB, bA, false, dt, out, basis);
    return collideBoxes(bA, bB, out, basis);
}

// manifold.cpp:14-27: contact anchors of every round shape (cylinders included) are re-taken at
// each initialisation; cylinder cap contacts keep their rim direction.
template <class R> inline bool anyRound(const Body<R> &b)
{
    return b.shape == SHAPE_SPHERE || b.shape == SHAPE_CAPSULE || b.shape == SHAPE_CYLINDER;
}
inline bool cylinderCapKey(int32_t key)
{
    return (key & (int32_t)0xFFFFFF00) == 0x05000000 || (key & (int32_t)0xFFFFFF00) == 0x06000000;
}

// World-frame offset of a contact point from the body centre (manifold.cpp:29-90): a box rotates
// its stored local offset; a sphere's point is its radius along the manifold normal; a capsule's
// slides along its core to the stored height moved by the normal; a cylinder's is the rim point
// in the support direction (or the stored cap direction for cap contacts).
template <class R>
inline V3<R> contactOffset(const Body<R> &b, V3<R> lo, const M3<R> &basis, bool isA, int32_t key)
{
    if (b.shape == SHAPE_SPHERE)
        return basis[0] * (isA ? -b.shapeRadius : b.shapeRadius);
    if (b.shape == SHAPE_CAPSULE)
    {
        V3<R> z = localZ(b);
        V3<R> n = basis[0];
        V3<R> nl = rotate(conjugate(b.rot), n);
        R cz = isA ? lo.z + nl.z * b.shapeRadius : lo.z - nl.z * b.shapeRadius;
        cz = clampv(cz, -b.halfLength, b.halfLength);
        V3<R> along = z * cz;
        return along + n * (isA ? -b.shapeRadius : b.shapeRadius);
    }
    if (b.shape == SHAPE_CYLINDER)
    {
        V3<R> dirW = basis[0] * (isA ? R(-1) : R(1));
        V3<R> dirL = rotate(conjugate(b.rot), dirW);
        R storedRad = sqrt_(lo.x * lo.x + lo.y * lo.y);
        if (cylinderCapKey(key))
        {
            V3<R> l{0, 0, 0};
            if (storedRad > R(1.0e-6f))
            {
                l.x = lo.x / storedRad * b.shapeRadius;
                l.y = lo.y / storedRad * b.shapeRadius;
            }
            R capSign = abs_(lo.z) > R(1.0e-6f) ? (lo.z >= 0 ? R(1) : R(-1)) : (dirL.z >= 0 ? R(1) : R(-1));
            l.z = b.halfLength * capSign;
            return rotate(b.rot, l);
        }
        R dirRad = sqrt_(dirL.x * dirL.x + dirL.y * dirL.y);
        R z = abs_(dirL.z) > R(0.15f) ? (dirL.z >= 0 ? b.halfLength : -b.halfLength) : clampv(lo.z, -b.halfLength, b.halfLength);
        V3<R> l{0, 0, z};
        if (dirRad > R(1.0e-6f))
        {
            l.x = dirL.x / dirRad * b.shapeRadius;
            l.y = dirL.y / dirRad * b.shapeRadius;
        }
        else if (storedRad > R(1.0e-6f))
        {
            l.x = lo.x / storedRad * b.shapeRadius;
            l.y = lo.y / storedRad * b.shapeRadius;
        }
        return rotate(b.rot, l);
    }
    return rotate(b.rot, lo);
}

// ---- the world and its step ----
template <class R>
struct World
{
    Params<R> p;
    std::vector<Body<R>> bodies;        // body-list order
    std::vector<Joint<R>> joints;       // ascending serial
    std::vector<Spring<R>> springs;     // ascending serial
    std::vector<Manifold<R>> manifolds; // ascending serial
    std::vector<std::pair<int, int>> ignores; // IgnoreCollision pairs (force.cpp:75-79)
    int64_t nextSerial = 0;
    int orderMode = ORDER_COLOR;
    int broadphaseMode = BP_SAP;
    // External pair and contact sources (Solver::externalBroadphasePairs / externalManifoldContacts,
    // solver.h:361-365): with externalPairsOn the broadphase is broadphaseExternal; a non-empty
    // contact table feeds initManifold (Solver::setExternalManifoldContacts, solver.cpp:958-967).
    bool externalPairsOn = false;
    std::vector<std::pair<int, int>> externalPairs;
    std::vector<ExternalContact<R>> externalContacts;

    // introspection of the last step
    std::vector<int> colors;
    std::vector<Hit> hits;
    int sweepAxis = 0;
    int pairChecks = 0;
    int colorCount = 0;
    std::vector<std::vector<ForceRef>> bodyForces; // descending serial

    // IgnoreCollision forces add nothing to a body's system but still make it constrained
    // (solver.cpp:1161-1166, 1190-1191): body->forces counts them.
    std::vector<int> ignoreCount;
    bool hasForces(int i) const { return !bodyForces[i].empty() || (!ignoreCount.empty() && ignoreCount[i] > 0); }
    bool solved(int i) const { return bodies[i].mass > 0 && hasForces(i); }

    // Existing force between two bodies (Rigid::constrainedTo, rigid.cpp:105-116)
    bool constrained(int i, int j, const std::vector<std::vector<int>> &partners) const
    {
        const std::vector<int> &pi = partners[i];
        return std::find(pi.begin(), pi.end(), j) != pi.end();
    }

    void partnerLists(std::vector<std::vector<int>> &partners) const
    {
        partners.assign(bodies.size(), {});
        auto add = [&](int a, int b) {
            if (a >= 0)
                partners[a].push_back(b);
            if (b >= 0)
                partners[b].push_back(a);
        };
        for (const auto &m : manifolds)
            add(m.a, m.b);
        for (const auto &j : joints)
            add(j.a, j.b);
        for (const auto &s : springs)
            add(s.a, s.b);
        for (const auto &g : ignores)
            add(g.first, g.second);
    }

    // The broadphase of the world's mode, then the serial, ordered apply of new manifolds
    // (solver.cpp:220-248): new manifolds get consecutive serials in emission order.
    void broadphase()
    {
        hits.clear();
        pairChecks = 0;
        std::vector<std::pair<int, int>> created;
        if (bodies.size() >= 2)
        {
            std::vector<std::vector<int>> partners;
            partnerLists(partners);
            if (externalPairsOn)
                broadphaseExternal(partners, created);
            else if (broadphaseMode == BP_GRID)
                broadphaseGrid(partners, created);
            else if (broadphaseMode == BP_ALL_PAIRS)
                broadphaseAllPairs(partners, created);
            else
                broadphaseSap(partners, created);
        }
        for (auto &pr : created)
        {
            Manifold<R> m{};
            m.a = pr.first;
            m.b = pr.second;
            m.serial = nextSerial++;
            m.n = 0;
            manifolds.push_back(m);
        }
    }

    // A bounding-sphere hit (a, b): skipped if a force already joins the pair (processSphereHit /
    // classifyHit, solver.cpp:131-164, 202-216), else a new manifold (a, b) in emission order.
    void classify(int a, int b, const std::vector<std::vector<int>> &partners,
                  std::vector<std::pair<int, int>> &created)
    {
        bool con = partners[a].size() <= partners[b].size() ? constrained(a, b, partners) : constrained(b, a, partners);
        hits.push_back(Hit{a, b, con ? 0 : 1});
        if (!con)
            created.push_back({a, b});
    }

    // generatePair (solver.cpp:166-174): one bounding-sphere test.
    void spherePair(int a, int b, const std::vector<std::vector<int>> &partners,
                    std::vector<std::pair<int, int>> &created)
    {
        ++pairChecks;
        V3<R> dp = bodies[a].pos - bodies[b].pos;
        R r = bodies[a].radius + bodies[b].radius;
        if (dot(dp, dp) <= r * r)
            classify(a, b, partners, created);
    }

    // broadphaseExternalPairs (solver.cpp:579-626): the external pairs in their order, each with the
    // bounding-sphere test of generatePair (invalid ids, a body paired with itself: skipped), then
    // every external contact entry whose bodies are valid and in ascending dense-id order, counted
    // as a check and a hit with no sphere test. A new manifold joins its bodies at once
    // (Manifold's constructor), so a later entry for the same two bodies is a constrained hit.
    void broadphaseExternal(std::vector<std::vector<int>> &partners, std::vector<std::pair<int, int>> &created)
    {
        const int n = (int)bodies.size();
        auto valid = [&](int i) { return i >= 0 && i < n; };
        auto hit = [&](int a, int b) {
            const size_t before = created.size();
            classify(a, b, partners, created);
            if (created.size() != before)
            {
                partners[a].push_back(b);
                partners[b].push_back(a);
            }
        };
        for (const auto &pr : externalPairs)
        {
            const int a = pr.first, b = pr.second;
            if (!valid(a) || !valid(b) || a == b)
                continue;
            ++pairChecks;
            V3<R> dp = bodies[a].pos - bodies[b].pos;
            R r = bodies[a].radius + bodies[b].radius;
            if (dot(dp, dp) <= r * r)
                hit(a, b);
        }
        for (const auto &e : externalContacts)
        {
            if (!valid(e.a) || !valid(e.b) || (uint32_t)bodies[e.a].denseId >= (uint32_t)bodies[e.b].denseId)
                continue;
            ++pairChecks;
            hit(e.a, e.b);
        }
    }

    // Solver::findExternalManifoldContact (solver.cpp:976-987): the table's last entry for (a, b),
    // else its last entry for (b, a); null when the table is empty or has neither.
    const ExternalContact<R> *findExternal(int a, int b) const
    {
        for (int pass = 0; pass < 2; ++pass)
        {
            const int x = pass == 0 ? a : b, y = pass == 0 ? b : a;
            for (int k = (int)externalContacts.size() - 1; k >= 0; --k)
                if (externalContacts[k].a == x && externalContacts[k].b == y)
                    return &externalContacts[k];
        }
        return nullptr;
    }

    // broadphaseAllPairs (solver.cpp:176-185): every pair in list order.
    void broadphaseAllPairs(const std::vector<std::vector<int>> &partners, std::vector<std::pair<int, int>> &created)
    {
        const int n = (int)bodies.size();
        for (int a = 0; a < n; ++a)
            for (int b = a + 1; b < n; ++b)
                spherePair(a, b, partners, created);
    }

    // broadphaseUniformGrid (solver.cpp:284-486). Cell size twice the 95th-percentile bounding
    // radius (at least 1e-3), grown until the grid has at most 2^22 cells; bodies no larger than
    // half a cell are binned by centre, sorted by (cell key, list index) and tested within their
    // own cell and the 13 forward neighbour cells; larger bodies carry a world AABB (boxes) or
    // their bounding cube (round shapes) plus a |v| dt slack and are tested against every binned
    // body, then against each other by the plain sphere test. Emission order: binned bodies in
    // sorted order, each with its own-cell, forward-cell and large-body hits, then the large pairs.
    void broadphaseGrid(const std::vector<std::vector<int>> &partners, std::vector<std::pair<int, int>> &created)
    {
        using S = decltype(plain_(R()));
        const int n = (int)bodies.size();
        std::vector<S> radii(n);
        for (int i = 0; i < n; ++i)
            radii[i] = plain_(bodies[i].radius);
        const size_t pct = std::min<size_t>((size_t)n - 1, (size_t)n * 95 / 100);
        std::nth_element(radii.begin(), radii.begin() + pct, radii.end());
        const S binnedRadius = fmax_(radii[pct], S(1e-3f));
        S cellSize = binnedRadius * S(2.0f);
        struct Entry
        {
            uint32_t key;
            int idx;
            V3<S> c;
            S r, slack;
        };
        struct Large
        {
            int idx;
            V3<S> lo, hi;
            S slack;
        };
        std::vector<Entry> entries;
        std::vector<Large> large;
        V3<S> bmin{S(FLT_MAX), S(FLT_MAX), S(FLT_MAX)}, bmax{S(-FLT_MAX), S(-FLT_MAX), S(-FLT_MAX)};
        for (int i = 0; i < n; ++i)
        {
            const Body<R> &b = bodies[i];
            V3<S> v{plain_(b.vel.x), plain_(b.vel.y), plain_(b.vel.z)};
            const S slack = sqrt_(dot(v, v)) * S(plain_(p.dt));
            const V3<S> c{plain_(b.pos.x), plain_(b.pos.y), plain_(b.pos.z)};
            const S r = plain_(b.radius);
            if (r > binnedRadius)
            {
                V3<S> h{r, r, r};
                if (b.shape == SHAPE_BOX)
                {
                    const V3<S> e = V3<S>{plain_(b.size.x), plain_(b.size.y), plain_(b.size.z)} * S(0.5f);
                    const Q<S> q{plain_(b.rot.x), plain_(b.rot.y), plain_(b.rot.z), plain_(b.rot.w)};
                    const V3<S> cx = rotate(q, V3<S>{1, 0, 0}), cy = rotate(q, V3<S>{0, 1, 0}),
                                cz = rotate(q, V3<S>{0, 0, 1});
                    h = {std::fabs(cx.x) * e.x + std::fabs(cy.x) * e.y + std::fabs(cz.x) * e.z,
                         std::fabs(cx.y) * e.x + std::fabs(cy.y) * e.y + std::fabs(cz.y) * e.z,
                         std::fabs(cx.z) * e.x + std::fabs(cy.z) * e.y + std::fabs(cz.z) * e.z};
                }
                large.push_back(Large{i, c - h, c + h, slack});
                continue;
            }
            bmin = {fmin_(bmin.x, c.x), fmin_(bmin.y, c.y), fmin_(bmin.z, c.z)};
            bmax = {fmax_(bmax.x, c.x), fmax_(bmax.y, c.y), fmax_(bmax.z, c.z)};
            entries.push_back(Entry{0u, i, c, r, slack});
        }
        const size_t binned = entries.size();
        int nx = 1, ny = 1, nz = 1;
        std::vector<uint32_t> cellStart;
        if (binned > 0)
        {
            const uint64_t maxCells = 1ull << 22;
            const V3<S> extent = bmax - bmin;
            auto dimsFor = [&](S cs) -> uint64_t {
                nx = (int)(extent.x / cs) + 1;
                ny = (int)(extent.y / cs) + 1;
                nz = (int)(extent.z / cs) + 1;
                return (uint64_t)nx * (uint64_t)ny * (uint64_t)nz;
            };
            uint64_t numCells = dimsFor(cellSize);
            while (numCells > maxCells)
            {
                cellSize *= fmax_(S(std::cbrt((S)numCells / (S)maxCells)), S(1.1f));
                numCells = dimsFor(cellSize);
            }
            for (Entry &e : entries)
            {
                const int ix = (int)clampv((e.c.x - bmin.x) / cellSize, S(0.0f), (S)(nx - 1));
                const int iy = (int)clampv((e.c.y - bmin.y) / cellSize, S(0.0f), (S)(ny - 1));
                const int iz = (int)clampv((e.c.z - bmin.z) / cellSize, S(0.0f), (S)(nz - 1));
                e.key = (uint32_t)(ix + nx * (iy + ny * iz));
            }
            std::sort(entries.begin(), entries.end(), [](const Entry &a, const Entry &b) {
                return a.key == b.key ? a.idx < b.idx : a.key < b.key;
            });
            cellStart.assign((size_t)numCells, UINT32_MAX);
            for (size_t i = 0; i < binned; ++i)
                if (i == 0 || entries[i].key != entries[i - 1].key)
                    cellStart[entries[i].key] = (uint32_t)i;
        }
        static const int forward[13][3] = {{1, 0, 0},   {-1, 1, 0}, {0, 1, 0},  {1, 1, 0},  {-1, -1, 1},
                                           {0, -1, 1},  {1, -1, 1}, {-1, 0, 1}, {0, 0, 1},  {1, 0, 1},
                                           {-1, 1, 1},  {0, 1, 1},  {1, 1, 1}};
        const S reachMargin = S(4.0f * 0.01f);  // 4 * COLLISION_MARGIN (solver.h:23)
        for (size_t i = 0; i < binned; ++i)
        {
            const Entry &a = entries[i];
            auto test = [&](const Entry &b) {
                ++pairChecks;
                const V3<S> dp = a.c - b.c;
                const S r = a.r + b.r;
                if (!(dot(dp, dp) > r * r))
                    classify(a.idx, b.idx, partners, created);
            };
            for (size_t j = i + 1; j < binned && entries[j].key == a.key; ++j)
                test(entries[j]);
            const int ix = (int)(a.key % (uint32_t)nx);
            const uint32_t t = a.key / (uint32_t)nx;
            const int iy = (int)(t % (uint32_t)ny), iz = (int)(t / (uint32_t)ny);
            for (const int *d : forward)
            {
                const int cx = ix + d[0], cy = iy + d[1], cz = iz + d[2];
                if (cx < 0 || cx >= nx || cy < 0 || cy >= ny || cz < 0 || cz >= nz)
                    continue;
                const uint32_t nkey = (uint32_t)(cx + nx * (cy + ny * cz));
                const uint32_t s0 = cellStart[nkey];
                if (s0 == UINT32_MAX)
                    continue;
                for (size_t j = s0; j < binned && entries[j].key == nkey; ++j)
                    test(entries[j]);
            }
            for (const Large &l : large)
            {
                ++pairChecks;
                const V3<S> q{clampv(a.c.x, l.lo.x, l.hi.x), clampv(a.c.y, l.lo.y, l.hi.y),
                              clampv(a.c.z, l.lo.z, l.hi.z)};
                const V3<S> dp = a.c - q;
                const S reach = a.r + a.slack + l.slack + reachMargin;
                if (!(dot(dp, dp) > reach * reach))
                    classify(a.idx, l.idx, partners, created);
            }
        }
        for (size_t i = 0; i < large.size(); ++i)
            for (size_t j = i + 1; j < large.size(); ++j)
                spherePair(large[i].idx, large[j].idx, partners, created);
    }

    // Sweep and prune (solver.cpp:488-577).
    void broadphaseSap(const std::vector<std::vector<int>> &partners, std::vector<std::pair<int, int>> &created)
    {
        size_t n = bodies.size();
        V3<R> mean{0, 0, 0};
        for (const auto &b : bodies)
            mean += b.pos;
        mean = mean / (R)n;
        V3<R> var{0, 0, 0};
        for (const auto &b : bodies)
        {
            V3<R> d = b.pos - mean;
            var += V3<R>{d.x * d.x, d.y * d.y, d.z * d.z};
        }
        int axis = 0;
        if (var.y > var[axis])
            axis = 1;
        if (var.z > var[axis])
            axis = 2;
        sweepAxis = axis;

        struct Iv
        {
            R lo, hi;
            int idx;
        };
        std::vector<Iv> iv(n);
        for (size_t i = 0; i < n; ++i)
        {
            R c = bodies[i].pos[axis];
            iv[i] = Iv{c - bodies[i].radius, c + bodies[i].radius, (int)i};
        }
        std::sort(iv.begin(), iv.end(), [](const Iv &a, const Iv &b) {
            if (a.lo == b.lo)
                return a.idx < b.idx;
            return a.lo < b.lo;
        });

        for (size_t i = 0; i < n; ++i)
        {
            const Iv &a = iv[i];
            for (size_t j = i + 1; j < n; ++j)
            {
                const Iv &b = iv[j];
                if (b.lo > a.hi)
                    break;
                ++pairChecks;
                V3<R> dp = bodies[a.idx].pos - bodies[b.idx].pos;
                R r = bodies[a.idx].radius + bodies[b.idx].radius;
                if (dot(dp, dp) > r * r)
                    continue;
                classify(a.idx, b.idx, partners, created);
            }
        }
    }

    // Manifold::initialize (manifold.cpp:100-174). A sphere in the pair makes the manifold
    // frictionless; with any round shape (cylinders included) matched contacts always take the
    // new anchors and restart their friction rows (lambda, penalty, stick).
    bool initManifold(Manifold<R> &m)
    {
        const Body<R> &A = bodies[m.a];
        const Body<R> &B = bodies[m.b];
        m.friction = sqrt_(A.friction * B.friction);
        if (A.shape == SHAPE_SPHERE || B.shape == SHAPE_SPHERE)
            m.friction = R(0);
        Contact<R> nc[8];
        std::memset((void *)nc, 0, sizeof(nc));
        int nn;
        // an external contact entry for the pair replaces collide (manifold.cpp:110-121): its basis
        // and its first min(numContacts, 8) contacts, penalties, lambdas and stick flags included
        if (const ExternalContact<R> *e = findExternal(m.a, m.b))
        {
            m.basis = e->basis;
            nn = std::min(e->n, 8);
            for (int i = 0; i < nn; ++i)
                nc[i] = e->c[i];
        }
        else
            nn = collidePair(A, B, p.dt, nc, m.basis);
        bool refresh = anyRound(A) || anyRound(B);
        for (int i = 0; i < nn; ++i)
        {
            for (int j = 0; j < m.n; ++j)
            {
                if (nc[i].feature == m.c[j].feature)
                {
                    V3<R> rA = nc[i].rA, rB = nc[i].rB;
                    nc[i] = m.c[j];
                    if (refresh || !m.c[j].stick)
                    {
                        nc[i].rA = rA;
                        nc[i].rB = rB;
                    }
                    if (refresh)
                    {
                        nc[i].lambda[1] = R(0);
                        nc[i].lambda[2] = R(0);
                        nc[i].penalty[1] = K<R>::penaltyMin;
                        nc[i].penalty[2] = K<R>::penaltyMin;
                        nc[i].stick = false;
                    }
                    break;
                }
            }
        }
        m.n = nn;
        for (int i = 0; i < nn; ++i)
            m.c[i] = nc[i];
        for (int i = 0; i < m.n; ++i)
        {
            Contact<R> &c = m.c[i];
            V3<R> xA = A.pos + contactOffset(A, c.rA, m.basis, true, c.feature);
            V3<R> xB = B.pos + contactOffset(B, c.rB, m.basis, false, c.feature);
            c.C0 = m.basis * (xA - xB) + V3<R>{K<R>::margin, 0, 0};
            c.lambda = c.lambda * p.alpha * p.gamma;
            c.penalty = clampv(c.penalty * p.gamma, K<R>::penaltyMin, K<R>::penaltyMax);
        }
        return m.n > 0;
    }

    // Joint::initialize (joint.cpp:35-54)
    bool initJoint(Joint<R> &j)
    {
        const Body<R> &B = bodies[j.b];
        V3<R> pa = j.a >= 0 ? transform(bodies[j.a].pos, bodies[j.a].rot, j.rA) : j.rA;
        j.C0Lin = pa - transform(B.pos, B.rot, j.rB);
        Q<R> qa = j.a >= 0 ? bodies[j.a].rot : Q<R>{0, 0, 0, 1};
        j.C0Ang = (qa - B.rot) * j.torqueArm;
        j.lamLin = j.lamLin * p.alpha * p.gamma;
        j.lamAng = j.lamAng * p.alpha * p.gamma;
        j.penLin = clampv(j.penLin * p.gamma, K<R>::penaltyMin, K<R>::penaltyMax);
        j.penAng = clampv(j.penAng * p.gamma, K<R>::penaltyMin, K<R>::penaltyMax);
        j.penLin = minv(j.penLin, j.kLin);
        j.penAng = minv(j.penAng, j.kAng);
        return !j.broken;
    }

    void buildBodyForces()
    {
        bodyForces.assign(bodies.size(), {});
        auto add = [&](int body, int kind, int idx, int64_t serial) {
            if (body >= 0)
                bodyForces[body].push_back(ForceRef{kind, idx, serial});
        };
        for (int k = 0; k < (int)manifolds.size(); ++k)
        {
            add(manifolds[k].a, KIND_MANIFOLD, k, manifolds[k].serial);
            add(manifolds[k].b, KIND_MANIFOLD, k, manifolds[k].serial);
        }
        for (int k = 0; k < (int)joints.size(); ++k)
        {
            add(joints[k].a, KIND_JOINT, k, joints[k].serial);
            add(joints[k].b, KIND_JOINT, k, joints[k].serial);
        }
        for (int k = 0; k < (int)springs.size(); ++k)
        {
            add(springs[k].a, KIND_SPRING, k, springs[k].serial);
            add(springs[k].b, KIND_SPRING, k, springs[k].serial);
        }
        for (auto &l : bodyForces)
            std::sort(l.begin(), l.end(), [](const ForceRef &x, const ForceRef &y) { return x.serial > y.serial; });
    }

    // Greedy coloring over solved bodies in list order (avbd_gpu_solver.cpp:1812-1865).
    void color()
    {
        int n = (int)bodies.size();
        colors.assign(n, -1);
        int maxColor = -1;
        std::vector<char> used;
        for (int i = 0; i < n; ++i)
        {
            if (!solved(i))
                continue;
            used.assign((size_t)maxColor + 2, 0);
            for (const ForceRef &f : bodyForces[i])
            {
                int o = partnerOf(f, i);
                if (o >= 0)
                {
                    int c = colors[o];
                    if (c >= 0 && c < (int)used.size())
                        used[c] = 1;
                }
            }
            int c = 0;
            while (c < (int)used.size() && used[c])
                ++c;
            colors[i] = c;
            maxColor = std::max(maxColor, c);
        }
        colorCount = maxColor + 1;
    }

    int partnerOf(const ForceRef &f, int i) const
    {
        int a, b;
        if (f.kind == KIND_MANIFOLD)
        {
            a = manifolds[f.index].a;
            b = manifolds[f.index].b;
        }
        else if (f.kind == KIND_JOINT)
        {
            a = joints[f.index].a;
            b = joints[f.index].b;
        }
        else
        {
            a = springs[f.index].a;
            b = springs[f.index].b;
        }
        return a == i ? b : a;
    }

    // Solver::prepareStep (solver.cpp:1043-1170)
    void prepare()
    {
        broadphase();
        // force initialization; inactive forces removed in order
        std::vector<Manifold<R>> keptM;
        keptM.reserve(manifolds.size());
        for (auto &m : manifolds)
            if (initManifold(m))
                keptM.push_back(m);
        manifolds.swap(keptM);
        std::vector<Joint<R>> keptJ;
        for (auto &j : joints)
            if (initJoint(j))
                keptJ.push_back(j);
        joints.swap(keptJ);
        buildBodyForces();
        ignoreCount.assign(bodies.size(), 0);
        for (const auto &g : ignores)
        {
            ++ignoreCount[g.first];
            ++ignoreCount[g.second];
        }

        for (int i = 0; i < (int)bodies.size(); ++i)
        {
            Body<R> &b = bodies[i];
            b.inerLin = b.pos + b.vel * p.dt;
            if (b.mass > 0)
                b.inerLin += V3<R>{0, 0, p.gravity} * (p.dt * p.dt);
            b.inerAng = b.rot + b.angVel * p.dt;
            V3<R> accel = (b.vel - b.prevVel) / p.dt;
            R accelExt = accel.z * signv(p.gravity);
            R w = clampk_(accelExt / abs_(p.gravity), R(0), R(1));
            if (!isfinite_(w))
                w = 0;
            b.initLin = b.pos;
            b.initAng = b.rot;
            if (b.mass > 0)
            {
                b.pos = b.pos + b.vel * p.dt + V3<R>{0, 0, p.gravity} * (w * p.dt * p.dt);
                b.rot = b.rot + b.angVel * p.dt;
            }
            if (b.mass > 0 && !hasForces(i))
            {
                b.pos = b.inerLin;
                b.rot = b.inerAng;
            }
        }
        color();
    }

    // Manifold::updatePrimal contribution of one manifold (manifold.cpp:176-229)
    void stampManifold(const Manifold<R> &m, int body, M3<R> &lLin, M3<R> &lAng, M3<R> &lCross, V3<R> &rLin, V3<R> &rAng) const
    {
        const Body<R> &A = bodies[m.a];
        const Body<R> &B = bodies[m.b];
        V3<R> dqAL = A.pos - A.initLin, dqAA = A.rot - A.initAng;
        V3<R> dqBL = B.pos - B.initLin, dqBA = B.rot - B.initAng;
        for (int i = 0; i < m.n; ++i)
        {
            const Contact<R> &c = m.c[i];
            V3<R> rAW = contactOffset(A, c.rA, m.basis, true, c.feature);
            V3<R> rBW = contactOffset(B, c.rB, m.basis, false, c.feature);
            M3<R> jAL = m.basis, jBL = -m.basis;
            M3<R> jAA{cross(rAW, jAL[0]), cross(rAW, jAL[1]), cross(rAW, jAL[2])};
            M3<R> jBA{cross(rBW, jBL[0]), cross(rBW, jBL[1]), cross(rBW, jBL[2])};
            M3<R> Km = diag(c.penalty.x, c.penalty.y, c.penalty.z);
            V3<R> C = c.C0 * (R(1) - p.alpha) + jAL * dqAL + jBL * dqBL + jAA * dqAA + jBA * dqBA;
            V3<R> F = Km * C + c.lambda;
            F[0] = fmin_(F[0], R(0));
            R bounds = abs_(F[0]) * m.friction;
            R fs = sqrt_(F[1] * F[1] + F[2] * F[2]);
            if (fs > bounds && fs > 0)
            {
                F[1] *= bounds / fs;
                F[2] *= bounds / fs;
            }
            const M3<R> &jL = body == m.a ? jAL : jBL;
            const M3<R> &jA = body == m.a ? jAA : jBA;
            M3<R> jLT = transpose(jL), jAT = transpose(jA);
            M3<R> jATk = jAT * Km;
            lLin += jLT * Km * jL;
            lAng += jATk * jA;
            lCross += jATk * jL;
            rLin += jLT * F;
            rAng += jAT * F;
        }
    }

    // Manifold::updateDual (manifold.cpp:231-280)
    void dualManifold(Manifold<R> &m) const
    {
        const Body<R> &A = bodies[m.a];
        const Body<R> &B = bodies[m.b];
        V3<R> dqAL = A.pos - A.initLin, dqAA = A.rot - A.initAng;
        V3<R> dqBL = B.pos - B.initLin, dqBA = B.rot - B.initAng;
        for (int i = 0; i < m.n; ++i)
        {
            Contact<R> &c = m.c[i];
            V3<R> rAW = contactOffset(A, c.rA, m.basis, true, c.feature);
            V3<R> rBW = contactOffset(B, c.rB, m.basis, false, c.feature);
            M3<R> jAL = m.basis, jBL = -m.basis;
            M3<R> jAA{cross(rAW, jAL[0]), cross(rAW, jAL[1]), cross(rAW, jAL[2])};
            M3<R> jBA{cross(rBW, jBL[0]), cross(rBW, jBL[1]), cross(rBW, jBL[2])};
            M3<R> Km = diag(c.penalty.x, c.penalty.y, c.penalty.z);
            V3<R> C = c.C0 * (R(1) - p.alpha) + jAL * dqAL + jBL * dqBL + jAA * dqAA + jBA * dqBA;
            V3<R> F = Km * C + c.lambda;
            F[0] = fmin_(F[0], R(0));
            R bounds = abs_(F[0]) * m.friction;
            R fs = sqrt_(F[1] * F[1] + F[2] * F[2]);
            if (fs > bounds && fs > 0)
            {
                F[1] *= bounds / fs;
                F[2] *= bounds / fs;
            }
            c.lambda = F;
            if (F[0] < 0)
                c.penalty[0] = fmin_(c.penalty[0] + p.betaLin * absk_(C[0]), K<R>::penaltyMax);
            if (fs <= bounds)
            {
                c.penalty[1] = fmin_(c.penalty[1] + p.betaLin * absk_(C[1]), K<R>::penaltyMax);
                c.penalty[2] = fmin_(c.penalty[2] + p.betaLin * absk_(C[2]), K<R>::penaltyMax);
                c.stick = sqrt_(C[1] * C[1] + C[2] * C[2]) < K<R>::stickThresh;
            }
        }
    }

    // geometric stiffness helper (joint.cpp:14-23)
    static M3<R> gsBall(int k, V3<R> v)
    {
        M3<R> m = diag(-v[k], -v[k], -v[k]);
        m[0][k] += v[0];
        m[1][k] += v[1];
        m[2][k] += v[2];
        return m;
    }

    // Joint::updatePrimal (joint.cpp:56-123)
    void stampJoint(const Joint<R> &j, int body, M3<R> &lLin, M3<R> &lAng, M3<R> &lCross, V3<R> &rLin, V3<R> &rAng) const
    {
        const Body<R> &B = bodies[j.b];
        bool isA = body == j.a;
        if (lengthSq(j.penLin) > 0)
        {
            V3<R> rAW = j.a >= 0 ? rotate(bodies[j.a].rot, j.rA) : V3<R>{0, 0, 0};
            V3<R> rBW = rotate(B.rot, j.rB);
            M3<R> Km = diag(j.penLin.x, j.penLin.y, j.penLin.z);
            V3<R> C = (j.a >= 0 ? bodies[j.a].pos + rAW : j.rA) - (B.pos + rBW);
            if (isinf_(j.kLin))
                C -= j.C0Lin * p.alpha;
            V3<R> F = Km * C + j.lamLin;
            M3<R> jL = isA ? ident(R(1)) : ident(R(-1));
            M3<R> jA = isA ? skew(-rAW) : skew(rBW);
            M3<R> jLT = transpose(jL), jAT = transpose(jA);
            M3<R> jATk = jAT * Km;
            lLin += jLT * Km * jL;
            lAng += jATk * jA;
            lCross += jATk * jL;
            V3<R> r = isA ? rAW : -rBW;
            M3<R> H = gsBall(0, r) * F[0] + gsBall(1, r) * F[1] + gsBall(2, r) * F[2];
            lAng += diagonalize(H);
            rLin += jLT * F;
            rAng += jAT * F;
        }
        if (lengthSq(j.penAng) > 0)
        {
            M3<R> Km = diag(j.penAng.x, j.penAng.y, j.penAng.z);
            Q<R> qa = j.a >= 0 ? bodies[j.a].rot : Q<R>{0, 0, 0, 1};
            V3<R> C = (qa - B.rot) * j.torqueArm;
            if (isinf_(j.kAng))
                C -= j.C0Ang * p.alpha;
            V3<R> F = Km * C + j.lamAng;
            M3<R> jA = (isA ? ident(R(1)) : ident(R(-1))) * j.torqueArm;
            lAng += transpose(jA) * Km * jA;
            rAng += transpose(jA) * F;
        }
    }

    // Joint::updateDual (joint.cpp:125-182)
    void dualJoint(Joint<R> &j) const
    {
        const Body<R> &B = bodies[j.b];
        if (lengthSq(j.penLin) > 0)
        {
            M3<R> Km = diag(j.penLin.x, j.penLin.y, j.penLin.z);
            V3<R> pa = j.a >= 0 ? transform(bodies[j.a].pos, bodies[j.a].rot, j.rA) : j.rA;
            V3<R> C = pa - transform(B.pos, B.rot, j.rB);
            if (isinf_(j.kLin))
            {
                C -= j.C0Lin * p.alpha;
                j.lamLin = Km * C + j.lamLin;
            }
            j.penLin = minv(j.penLin + absvk(C) * p.betaLin, fmin_(j.kLin, K<R>::penaltyMax));
        }
        if (lengthSq(j.penAng) > 0)
        {
            M3<R> Km = diag(j.penAng.x, j.penAng.y, j.penAng.z);
            Q<R> qa = j.a >= 0 ? bodies[j.a].rot : Q<R>{0, 0, 0, 1};
            V3<R> C = (qa - B.rot) * j.torqueArm;
            if (isinf_(j.kAng))
            {
                C -= j.C0Ang * p.alpha;
                j.lamAng = Km * C + j.lamAng;
            }
            j.penAng = minv(j.penAng + absvk(C) * p.betaAng, fmin_(j.kAng, K<R>::penaltyMax));
        }
        if (lengthSq(j.lamAng) > j.fracture * j.fracture)
        {
            j.penLin = j.penAng = j.lamLin = j.lamAng = V3<R>{0, 0, 0};
            j.broken = true;
        }
    }

    // Spring::updatePrimal (spring.cpp:27-69)
    void stampSpring(const Spring<R> &s, int body, M3<R> &lLin, M3<R> &lAng, M3<R> &lCross, V3<R> &rLin, V3<R> &rAng) const
    {
        const Body<R> &A = bodies[s.a];
        const Body<R> &B = bodies[s.b];
        V3<R> d = transform(A.pos, A.rot, s.rA) - transform(B.pos, B.rot, s.rB);
        R dl = length(d);
        if (dl <= R(1.0e-6f))
            return;
        V3<R> n = d / dl;
        R f = s.k * (dl - s.rest);
        V3<R> jL, jA;
        if (body == s.a)
        {
            V3<R> rW = rotate(A.rot, s.rA);
            jL = n;
            jA = cross(rW, n);
        }
        else
        {
            V3<R> rW = rotate(B.rot, s.rB);
            jL = -n;
            jA = -cross(rW, n);
        }
        lLin += outer(jL, jL) * s.k;
        lAng += outer(jA, jA) * s.k;
        lCross += outer(jA, jL) * s.k;
        rLin += jL * f;
        rAng += jA * f;
    }

    // One body's primal update (solver.cpp:1184-1240)
    void primalBody(int i)
    {
        Body<R> &b = bodies[i];
        R dt2 = p.dt * p.dt;
        M3<R> ML = diag(b.mass, b.mass, b.mass);
        M3<R> MA = diag(b.moment.x, b.moment.y, b.moment.z);
        M3<R> lLin = ML / dt2, lAng = MA / dt2;
        M3<R> lCross{V3<R>{0, 0, 0}, V3<R>{0, 0, 0}, V3<R>{0, 0, 0}};
        V3<R> rLin = ML / dt2 * (b.pos - b.inerLin);
        V3<R> rAng = MA / dt2 * (b.rot - b.inerAng);
        for (const ForceRef &f : bodyForces[i])
        {
            if (f.kind == KIND_MANIFOLD)
                stampManifold(manifolds[f.index], i, lLin, lAng, lCross, rLin, rAng);
            else if (f.kind == KIND_JOINT)
                stampJoint(joints[f.index], i, lLin, lAng, lCross, rLin, rAng);
            else
                stampSpring(springs[f.index], i, lLin, lAng, lCross, rLin, rAng);
        }
        V3<R> dxL, dxA;
        solve6(lLin, lAng, lCross, -rLin, -rAng, dxL, dxA);
        b.pos = b.pos + dxL;
        b.rot = b.rot + dxA;
    }

    // Solver::iteratePrimalDualCpu (solver.cpp:1172-1273)
    void iterate()
    {
        int n = (int)bodies.size();
        std::vector<std::vector<int>> byColor;
        if (orderMode == ORDER_COLOR)
        {
            byColor.assign(colorCount, {});
            for (int i = 0; i < n; ++i)
                if (colors[i] >= 0)
                    byColor[colors[i]].push_back(i);
        }
        for (int it = 0; it < p.iterations; ++it)
        {
            if (orderMode == ORDER_COLOR)
            {
                for (auto &list : byColor)
                    for (int i : list)
                        primalBody(i);
            }
            else
            {
                for (int i = 0; i < n; ++i)
                    if (solved(i))
                        primalBody(i);
            }
            for (auto &m : manifolds)
                dualManifold(m);
            for (auto &j : joints)
                dualJoint(j);
        }
    }

    // Solver::finishStep (solver.cpp:1275-1298): BDF1 velocities, then the sphere rolling
    // friction pass over the manifolds in force-list order (newest first).
    void finish()
    {
        for (auto &b : bodies)
        {
            b.prevVel = b.vel;
            if (b.mass > 0)
            {
                b.vel = (b.pos - b.initLin) / p.dt;
                b.angVel = (b.rot - b.initAng) / p.dt;
            }
        }
        for (size_t k = manifolds.size(); k-- > 0;)
            rollingFriction(manifolds[k]);
    }

    // applySphereRollingFriction (solver.cpp:628-660): a dynamic sphere resting on a static body
    // loses a friction-scaled share of its contact-point slip and is pulled toward rolling,
    // once per contact point.
    void rollingFriction(const Manifold<R> &m)
    {
        int si = -1, oi = -1;
        if (bodies[m.a].shape == SHAPE_SPHERE)
        {
            si = m.a;
            oi = m.b;
        }
        else if (bodies[m.b].shape == SHAPE_SPHERE)
        {
            si = m.b;
            oi = m.a;
        }
        if (si < 0)
            return;
        Body<R> &s = bodies[si];
        const Body<R> &o = bodies[oi];
        if (s.mass <= 0 || o.mass > 0 || m.n <= 0)
            return;
        R mu = sqrt_(s.friction * o.friction);
        if (mu <= 0)
            return;
        R blend = clampv(mu * R(0.08f), R(0), R(0.18f));
        hitBranch(RB_ROLLING);
        for (int i = 0; i < m.n; ++i)
        {
            V3<R> n = si == m.a ? m.basis[0] : -m.basis[0];
            V3<R> rW = -n * s.shapeRadius;
            V3<R> cv = s.vel + cross(s.angVel, rW);
            V3<R> slip = cv - n * dot(cv, n);
            s.vel -= slip * blend;
            V3<R> roll = cross(n, s.vel) / s.shapeRadius;
            s.angVel += (roll - s.angVel) * blend;
        }
    }

    void step()
    {
        prepare();
        iterate();
        finish();
    }
};

} // namespace avbd_oracle
