// avbd_math.hpp -- FP32 vector / quaternion / 3x3 algebra for the HIP kernels.
//
// Every operation keeps the operand order of the reference math library
// (maths.h:117-277, 317-547) so that, compiled with -ffp-contract=off and
// HIP's default correctly rounded division and square root, a kernel produces
// the same FP32 bits as the reference CPU solver. Expressions that multiply by
// a structural zero or one of the reference's dense 3x3 products are written
// out with the zero/one terms removed only where that is exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define AVBD_HD __host__ __device__ __forceinline__

namespace avbd
{

struct V3
{
    float x, y, z;
};
struct Q4
{
    float x, y, z, w;
};
struct M3
{
    V3 r0, r1, r2;
};

AVBD_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
AVBD_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AVBD_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AVBD_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
AVBD_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
AVBD_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
AVBD_HD void operator+=(V3 &a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
AVBD_HD void operator-=(V3 &a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
AVBD_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
AVBD_HD float lensq(V3 a) { return dot(a, a); }
AVBD_HD float len(V3 a) { return sqrtf(lensq(a)); }
AVBD_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
AVBD_HD float fminr(float a, float b) { return a < b ? a : b; }
AVBD_HD float fmaxr(float a, float b) { return a > b ? a : b; }
AVBD_HD float clampr(float x, float a, float b) { return fmaxr(a, fminr(b, x)); }
AVBD_HD float signr(float x) { return x < 0.0f ? -1.0f : x > 0.0f ? 1.0f : 0.0f; }
AVBD_HD float comp(V3 v, int i) { return i == 0 ? v.x : i == 1 ? v.y : v.z; }

AVBD_HD M3 operator-(const M3 &m) { return M3{-m.r0, -m.r1, -m.r2}; }
AVBD_HD V3 mul(const M3 &m, V3 v) { return V3{dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)}; }
AVBD_HD void add(M3 &a, const M3 &b) { a.r0 += b.r0; a.r1 += b.r1; a.r2 += b.r2; }

// Quaternions (maths.h:255-277, 347-365, 431-446)
AVBD_HD Q4 qmul(Q4 a, Q4 b)
{
    return Q4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
              a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
AVBD_HD float qlensq(Q4 q) { return (q.x * q.x + q.y * q.y + q.z * q.z) + q.w * q.w; }
AVBD_HD Q4 qdiv(Q4 q, float s) { return Q4{q.x / s, q.y / s, q.z / s, q.w / s}; }
AVBD_HD Q4 qconj(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
// q - p = 2 vec(q * p^-1), p^-1 = conj(p) / |p|^2
AVBD_HD V3 qdiff(Q4 a, Q4 b)
{
    Q4 d = qmul(a, qdiv(qconj(b), qlensq(b)));
    return V3{d.x * 2.0f, d.y * 2.0f, d.z * 2.0f};
}
// q + w = normalize(q + ((w,0) q) * 0.5)
AVBD_HD Q4 qintegrate(Q4 a, V3 w)
{
    Q4 t = qmul(Q4{w.x, w.y, w.z, 0.0f}, a);
    Q4 s = Q4{a.x + t.x * 0.5f, a.y + t.y * 0.5f, a.z + t.z * 0.5f, a.w + t.w * 0.5f};
    return qdiv(s, sqrtf(qlensq(s)));
}
AVBD_HD V3 rotate(Q4 q, V3 v)
{
    V3 u{q.x, q.y, q.z};
    V3 t = cross(u, v) * 2.0f;
    return v + t * q.w + cross(u, t);
}
AVBD_HD V3 rotate_inv(Q4 q, V3 v) { return rotate(qconj(q), v); }

// orthonormal basis with the normal in row 0 (maths.h:468-474)
AVBD_HD M3 orthonormal(V3 n)
{
    V3 t1 = fabsf(n.x) > fabsf(n.z) ? V3{-n.y, n.x, 0.0f} : V3{0.0f, -n.z, n.y};
    t1 = t1 / len(t1);
    return M3{n, t1, cross(n, t1)};
}

// 6x6 LDL^T solve of [[Lin, Cross^T], [Cross, Ang]] x = b (maths.h:481-547).
// The lower triangles of the three blocks are passed element-wise.
struct Sys6
{
    float l00, l10, l11, l20, l21, l22;     // Lin lower triangle
    float a00, a10, a11, a20, a21, a22;     // Ang lower triangle
    float c00, c01, c02, c10, c11, c12, c20, c21, c22; // Cross (full)
    V3 bl, ba;
};

AVBD_HD void solve6(const Sys6 &s, V3 bLin, V3 bAng, V3 &xLin, V3 &xAng)
{
    float A11 = s.l00;
    float A21 = s.l10, A22 = s.l11;
    float A31 = s.l20, A32 = s.l21, A33 = s.l22;
    float A41 = s.c00, A42 = s.c01, A43 = s.c02, A44 = s.a00;
    float A51 = s.c10, A52 = s.c11, A53 = s.c12, A54 = s.a10, A55 = s.a11;
    float A61 = s.c20, A62 = s.c21, A63 = s.c22, A64 = s.a20, A65 = s.a21, A66 = s.a22;

    float D1 = A11;
    float L21 = A21 / A11, L31 = A31 / A11, L41 = A41 / A11, L51 = A51 / A11, L61 = A61 / A11;
    float D2 = A22 - L21 * L21 * D1;
    float L32 = (A32 - L21 * L31 * D1) / D2;
    float L42 = (A42 - L21 * L41 * D1) / D2;
    float L52 = (A52 - L21 * L51 * D1) / D2;
    float L62 = (A62 - L21 * L61 * D1) / D2;
    float D3 = A33 - (L31 * L31 * D1 + L32 * L32 * D2);
    float L43 = (A43 - L31 * L41 * D1 - L32 * L42 * D2) / D3;
    float L53 = (A53 - L31 * L51 * D1 - L32 * L52 * D2) / D3;
    float L63 = (A63 - L31 * L61 * D1 - L32 * L62 * D2) / D3;
    float D4 = A44 - (L41 * L41 * D1 + L42 * L42 * D2 + L43 * L43 * D3);
    float L54 = (A54 - L41 * L51 * D1 - L42 * L52 * D2 - L43 * L53 * D3) / D4;
    float L64 = (A64 - L41 * L61 * D1 - L42 * L62 * D2 - L43 * L63 * D3) / D4;
    float D5 = A55 - (L51 * L51 * D1 + L52 * L52 * D2 + L53 * L53 * D3 + L54 * L54 * D4);
    float L65 = (A65 - L51 * L61 * D1 - L52 * L62 * D2 - L53 * L63 * D3 - L54 * L64 * D4) / D5;
    float D6 = A66 - (L61 * L61 * D1 + L62 * L62 * D2 + L63 * L63 * D3 + L64 * L64 * D4 + L65 * L65 * D5);

    float y1 = bLin.x;
    float y2 = bLin.y - L21 * y1;
    float y3 = bLin.z - L31 * y1 - L32 * y2;
    float y4 = bAng.x - L41 * y1 - L42 * y2 - L43 * y3;
    float y5 = bAng.y - L51 * y1 - L52 * y2 - L53 * y3 - L54 * y4;
    float y6 = bAng.z - L61 * y1 - L62 * y2 - L63 * y3 - L64 * y4 - L65 * y5;

    float z1 = y1 / D1, z2 = y2 / D2, z3 = y3 / D3, z4 = y4 / D4, z5 = y5 / D5, z6 = y6 / D6;

    xAng.z = z6;
    xAng.y = z5 - L65 * xAng.z;
    xAng.x = z4 - L54 * xAng.y - L64 * xAng.z;
    xLin.z = z3 - L43 * xAng.x - L53 * xAng.y - L63 * xAng.z;
    xLin.y = z2 - L32 * xLin.z - L42 * xAng.x - L52 * xAng.y - L62 * xAng.z;
    xLin.x = z1 - L21 * xLin.y - L31 * xLin.z - L41 * xAng.x - L51 * xAng.y - L61 * xAng.z;
}

AVBD_HD V3 ld3(const float4 &f) { return V3{f.x, f.y, f.z}; }
AVBD_HD Q4 ldq(const float4 &f) { return Q4{f.x, f.y, f.z, f.w}; }
AVBD_HD float4 st3(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
AVBD_HD float4 stq(Q4 q) { return make_float4(q.x, q.y, q.z, q.w); }

} // namespace avbd
