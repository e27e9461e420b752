// avbd_adjoint.hip -- hand-written adjoint (reverse mode) of the resident AVBD step for worlds of
// boxes with contacts, joints and springs. Included by avbd_capi.hip after the handle definition.
//
// The forward step runs in FP32 (bit-exact with the reference, solver.cpp:1036-1041); the
// adjoint differentiates that FP32 trajectory, with every branch the step took (contact set,
// features, friction-cone and normal clamps, penalty saturation, stick flags) held fixed, in
// FP64 arithmetic. Pieces, in reverse order of the step:
//   finishStep (solver.cpp:1275-1288)            v = (x - x0)/dt, w = (q - q0)/dt
//   iterations (solver.cpp:1172-1273), reversed iteration by iteration:
//     dual (manifold.cpp:231-280)                 lambda = F(C), k += beta |C| (unsaturated rows)
//     joint dual (joint.cpp:125-182)              lambda += K C (infinite stiffness), k += beta |C|
//     primal colours in reverse (manifold.cpp:176-229, solver.cpp:1184-1240): the 6x6 solve's
//       adjoint is one more LDL^T solve with the same (symmetric) matrix, w = H^-1 dx_bar, then
//       r_bar = -w, H_bar = -w dx^T are pushed through every contact, joint (linear rows,
//       geometric-stiffness diagonal, angular rows; joint.cpp:56-123) and spring (spring.cpp:27-69)
//       stamp and the inertia terms
//   warm start (solver.cpp:1134-1169)             inertial targets, adaptive gravity weight
//   joint init (joint.cpp:35-54)                  C0 from the start poses, lambda*alpha*gamma, clamped k*gamma
//   force init (manifold.cpp:100-174)             C0 from the start poses, lambda*alpha*gamma,
//       clamp(k*gamma); anchors kept by the stick rule pass their adjoint to the previous step,
//       fresh anchors and the basis go through the collide pose Jacobian, evaluated by running
//       collide_boxes<DualF> once per pose component (14 per manifold) on the step's branch.
// Checkpoints: the state at every K-th step start is kept in HBM; a segment is replayed to
// recover each step's start state, and each step is recomputed once more with its iterations
// taped (poses and lambda/penalty after every iteration) right before it is reversed.
#pragma once

namespace avbd
{

// ---- FP64 vector / quaternion helpers ----
struct DV
{
    double x, y, z;
};
struct DQ
{
    double x, y, z, w;
};
__device__ __forceinline__ DV dv(V3 v) { return DV{v.x, v.y, v.z}; }
__device__ __forceinline__ DV dv4(float4 v) { return DV{v.x, v.y, v.z}; }
__device__ __forceinline__ DQ dq4(float4 q) { return DQ{q.x, q.y, q.z, q.w}; }
__device__ __forceinline__ DV operator+(DV a, DV b) { return DV{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ DV operator-(DV a, DV b) { return DV{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ DV operator-(DV a) { return DV{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ DV operator*(DV a, double s) { return DV{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ void operator+=(DV &a, DV b) { a.x += b.x; a.y += b.y; a.z += b.z; }
__device__ __forceinline__ void operator-=(DV &a, DV b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }
__device__ __forceinline__ double ddot(DV a, DV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ DV dcross(DV a, DV b) { return DV{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ double dcomp(DV v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
__device__ __forceinline__ DQ operator+(DQ a, DQ b) { return DQ{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ DQ operator*(DQ a, double s) { return DQ{a.x * s, a.y * s, a.z * s, a.w * s}; }
__device__ __forceinline__ void operator+=(DQ &a, DQ b) { a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w; }
__device__ __forceinline__ DQ dqmul(DQ a, DQ b)
{
    return DQ{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__device__ __forceinline__ DQ dqconj(DQ q) { return DQ{-q.x, -q.y, -q.z, q.w}; }
__device__ __forceinline__ double dqdot(DQ a, DQ b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
__device__ __forceinline__ DV drot(DQ q, DV v)
{
    DV u{q.x, q.y, q.z};
    DV t = dcross(u, v) * 2.0;
    return v + t * q.w + dcross(u, t);
}
// out = rotate(q, v): accumulate q_bar, v_bar from out_bar
__device__ __forceinline__ void drot_adj(DQ q, DV v, DV ob, DQ &qb, DV &vb)
{
    DV u{q.x, q.y, q.z};
    DV t = dcross(u, v) * 2.0;
    double wb = ddot(ob, t);
    DV tb = ob * q.w + dcross(ob, u);
    DV ub = dcross(t, ob) + dcross(v, tb) * 2.0;
    vb += ob + dcross(tb, u) * 2.0;
    qb += DQ{ub.x, ub.y, ub.z, wb};
}
// d = a - b = 2 vec(a b^-1), b^-1 = conj(b)/|b|^2
__device__ __forceinline__ DV dqdiff(DQ a, DQ b)
{
    double s = dqdot(b, b);
    DQ d = dqmul(a, dqconj(b) * (1.0 / s));
    return DV{2.0 * d.x, 2.0 * d.y, 2.0 * d.z};
}
__device__ __forceinline__ void dqdiff_adj(DQ a, DQ b, DV db, DQ &ab, DQ &bb)
{
    double s = dqdot(b, b);
    const double rs = 1.0 / s;  // one division; the rest multiply by it
    DQ inv = dqconj(b) * rs;
    DQ pb{2.0 * db.x, 2.0 * db.y, 2.0 * db.z, 0.0};
    ab += dqmul(pb, dqconj(inv));
    DQ ib = dqmul(dqconj(a), pb);
    // inv = c o b / s with c = (-1,-1,-1,1)
    double cib = -b.x * ib.x - b.y * ib.y - b.z * ib.z + b.w * ib.w;
    double k2 = 2.0 * cib * (rs * rs);
    bb += DQ{-ib.x * rs - k2 * b.x, -ib.y * rs - k2 * b.y, -ib.z * rs - k2 * b.z, ib.w * rs - k2 * b.w};
}
// q' = normalize(u), u = q + ((w,0) q) / 2: accumulate q_bar, w_bar from q'_bar
__device__ __forceinline__ void dqint_adj(DQ q, DV w, DQ qpb, DQ &qb, DV &wb)
{
    DQ t = dqmul(DQ{w.x, w.y, w.z, 0.0}, q);
    DQ u = q + t * 0.5;
    double n = sqrt(dqdot(u, u));
    DQ qp = u * (1.0 / n);
    double pr = dqdot(qp, qpb);
    DQ ub = DQ{qpb.x - qp.x * pr, qpb.y - qp.y * pr, qpb.z - qp.z * pr, qpb.w - qp.w * pr} * (1.0 / n);
    DQ tb = ub * 0.5;
    DQ wq = dqmul(tb, dqconj(q));
    wb += DV{wq.x, wq.y, wq.z};
    qb += ub + dqmul(DQ{-w.x, -w.y, -w.z, 0.0}, tb);
}

// FP64 adds into adjoint accumulators: the hardware float atomic (the buffers are ordinary
// hipMalloc device memory), no compare-and-swap loop.
__device__ __forceinline__ void aadd(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add3(double *p, DV v)
{
    aadd(p, v.x);
    aadd(p + 1, v.y);
    aadd(p + 2, v.z);
}
__device__ __forceinline__ void atomic_add4(double *p, DQ v)
{
    aadd(p, v.x);
    aadd(p + 1, v.y);
    aadd(p + 2, v.z);
    aadd(p + 3, v.w);
}
__device__ __forceinline__ DV ld_dv(const double *p) { return DV{p[0], p[1], p[2]}; }
__device__ __forceinline__ DQ ld_dq(const double *p) { return DQ{p[0], p[1], p[2], p[3]}; }
__device__ __forceinline__ void st_dv(double *p, DV v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
__device__ __forceinline__ void st_dq(double *p, DQ v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.w; }
__device__ __forceinline__ void add_dv(double *p, DV v) { p[0] += v.x; p[1] += v.y; p[2] += v.z; }

// Adjoint buffers (device, FP64). Bodies: pose / velocity / init pose / inertial target adjoints,
// parameter adjoints. Contacts by slot (manifold * 8 + contact): penalty, lambda, C0, anchors.
// Carry: contact adjoints handed to the previous step's final manifolds, by their slot.
struct AdjBuf
{
    double *xb, *qb, *vb, *wb, *pvb, *ilb, *iab, *nlb, *nab, *mb, *mub;
    double *kb, *lb, *c0b, *rAb, *rBb, *basisb, *mumb;
    double *ckb, *clb, *crAb, *crBb;
    // joints, kJointAdj per joint: penalty lin, penalty ang, lambda lin, lambda ang (the state's
    // adjoint at the current point of the sweep), its carry to the previous step's final state,
    // and the C0 lin / ang adjoints (kJointC0 per joint) the joint init's reversal takes
    double *jb, *cjb, *jc0b;
};
constexpr int kJointAdj = 12, kJointC0 = 6;
enum { JA_PL = 0, JA_PA = 3, JA_LL = 6, JA_LA = 9 };

// ---- one contact's force and its adjoint (manifold.cpp:176-229 / 231-280) ----
struct CIn
{
    DV pA, pB, iLA, iLB, rA, rB, C0, n[3], K, lam;
    DQ qA, qB, iAA, iAB;
    double mu, oma;
    bool clamp0, scaled;
    int code;                        // the record's shape code (0: box pair)
    float radA, hlA, radB, hlB;      // radii and half lengths of the two shapes
};
struct CFwd
{
    DV rAW, rBW, dqAL, dqBL, dqAA, dqBA;
    double C[3], F0[3], F[3];
    double fs, bounds;
};
struct COut
{
    DV pA, pB, iLA, iLB, rA, rB, C0, n[3], K, lam;
    DQ qA, qB, iAA, iAB;
    double mu;
};

// Adjoint of a contact point's world offset (manifold.cpp:29-90). Box sides rotate the stored
// anchor (drot_adj). A sphere's offset is -+radius n. The other round offsets (capsule: the core
// point along the axis, clamped; cylinder: rim or cap point) are functions of the rotation, the
// stored anchor and the manifold normal, reversed through their forward-mode Jacobian: one DualF
// evaluation per input component at the step's FP32 values, so the branches are the step's.
__device__ void offset_adj(int code, bool isA, float radius, float hl, DQ q, DV lo, DV n, DV ob, DQ &qb, DV &lob,
                           DV &nb)
{
    const int shape = code == 0 ? kShapeBox : (isA ? (code & 3) : ((code >> 2) & 3));
    if (shape == kShapeBox)
    {
        drot_adj(q, lo, ob, qb, lob);
        return;
    }
    if (shape == kShapeSphere)
    {
        nb += ob * (isA ? -(double)radius : (double)radius);
        return;
    }
    const bool cap = (code & kCodeCap) != 0;
    const float qv[4] = {(float)q.x, (float)q.y, (float)q.z, (float)q.w};
    const float lv[3] = {(float)lo.x, (float)lo.y, (float)lo.z}, nv[3] = {(float)n.x, (float)n.y, (float)n.z};
    double g[10];
    for (int c = 0; c < 10; ++c)
    {
        auto sd = [&](float v, int k) { return DualF(v, c == k ? 1.0 : 0.0); };
        const Q4T<DualF> qd{sd(qv[0], 0), sd(qv[1], 1), sd(qv[2], 2), sd(qv[3], 3)};
        const V3T<DualF> ld{sd(lv[0], 4), sd(lv[1], 5), sd(lv[2], 6)}, nd{sd(nv[0], 7), sd(nv[1], 8), sd(nv[2], 9)};
        const V3T<DualF> r = contact_offset_t<DualF>(shape, radius, hl, cap, qd, ld, nd, isA);
        g[c] = ob.x * r.x.d + ob.y * r.y.d + ob.z * r.z.d;
    }
    qb += DQ{g[0], g[1], g[2], g[3]};
    lob += DV{g[4], g[5], g[6]};
    nb += DV{g[7], g[8], g[9]};
}

// The contact's forward values at which the reversal linearises: exactly those the FP32 step
// computed (contact_eval, manifold.cpp:176-229), promoted to FP64. The inputs are FP32 values held
// in FP64, so converting back is exact. Recomputing them in FP64 instead would not be the step's
// linearisation point: pose differences such as x - x- are small against a position's rounding,
// and F = k C + lambda multiplies a one-ulp change of C by penalties of up to 1e10, so the force
// values that scale the friction cone and the friction derivative would differ by percents.
__device__ __forceinline__ V3 f3(DV v) { return V3{(float)v.x, (float)v.y, (float)v.z}; }
__device__ __forceinline__ Q4 f4(DQ q) { return Q4{(float)q.x, (float)q.y, (float)q.z, (float)q.w}; }
// Round: the world may hold round shapes (without them the shape paths are not compiled in).
template <bool Round> __device__ __forceinline__ void cfwd(const CIn &in, CFwd &f)
{
    const Q4 qA = f4(in.qA), qB = f4(in.qB);
    V3 rAW, rBW;
    if (!Round || in.code == 0)
    {
        rAW = rotate(qA, f3(in.rA));
        rBW = rotate(qB, f3(in.rB));
    }
    else
    {
        const bool cap = (in.code & kCodeCap) != 0;
        const V3 n0 = f3(in.n[0]);
        rAW = contact_offset(in.code & 3, in.radA, in.hlA, cap, qA, f3(in.rA), n0, true);
        rBW = contact_offset((in.code >> 2) & 3, in.radB, in.hlB, cap, qB, f3(in.rB), n0, false);
    }
    const V3 dqAL = f3(in.pA) - f3(in.iLA), dqBL = f3(in.pB) - f3(in.iLB);
    const V3 dqAA = qdiff(qA, f4(in.iAA)), dqBA = qdiff(qB, f4(in.iAB));
    const V3 b0 = f3(in.n[0]), b1 = f3(in.n[1]), b2 = f3(in.n[2]);
    const V3 pen = f3(in.K), lam = f3(in.lam);
    const ContactEval ev = contact_eval(f3(in.C0), pen, lam, (float)in.mu, (float)in.oma, b0, b1, b2, -b0, -b1, -b2,
                                        cross(rAW, b0), cross(rAW, b1), cross(rAW, b2), cross(rBW, -b0),
                                        cross(rBW, -b1), cross(rBW, -b2), dqAL, dqBL, dqAA, dqBA);
    f.rAW = dv(rAW);
    f.rBW = dv(rBW);
    f.dqAL = dv(dqAL);
    f.dqBL = dv(dqBL);
    f.dqAA = dv(dqAA);
    f.dqBA = dv(dqBA);
    f.C[0] = ev.C0;
    f.C[1] = ev.C1;
    f.C[2] = ev.C2;
    // unclamped, unscaled forces as the step formed them
    f.F0[0] = pen.x * ev.C0 + lam.x;
    f.F0[1] = pen.y * ev.C1 + lam.y;
    f.F0[2] = pen.z * ev.C2 + lam.z;
    f.F[0] = in.clamp0 ? 0.0 : f.F0[0];
    f.bounds = ev.bounds;
    f.fs = ev.fs;
    f.F[1] = in.scaled ? (double)ev.F1 : f.F0[1];
    f.F[2] = in.scaled ? (double)ev.F2 : f.F0[2];
}

// Given F_bar, an extra C_bar and pre-accumulated rAW_bar / rBW_bar / n_bar (Jacobian terms of a
// stamp), accumulate every input adjoint into o.
template <bool Round>
__device__ __forceinline__ void cadj(const CIn &in, const CFwd &f, const double Fb[3], const double Cx[3], DV rAWb,
                                     DV rBWb, COut &o)
{
    double F0b[3];
    F0b[0] = Fb[0];
    F0b[1] = Fb[1];
    F0b[2] = Fb[2];
    if (in.scaled)
    {
        double s = f.bounds / f.fs;
        double ux = f.F0[1] / f.fs, uy = f.F0[2] / f.fs;
        double bb = Fb[1] * ux + Fb[2] * uy;
        F0b[1] = s * (Fb[1] - ux * bb);
        F0b[2] = s * (Fb[2] - uy * bb);
        // bounds = |F_0| mu
        double sg = f.F[0] < 0.0 ? -1.0 : (f.F[0] > 0.0 ? 1.0 : 0.0);
        F0b[0] += bb * in.mu * sg;
        o.mu += bb * fabs(f.F[0]);
    }
    if (in.clamp0)
        F0b[0] = 0.0;
    double Cb[3];
    for (int k = 0; k < 3; ++k)
    {
        Cb[k] = dcomp(in.K, k) * F0b[k] + Cx[k];
        double kb = f.C[k] * F0b[k];
        if (k == 0) { o.K.x += kb; o.lam.x += F0b[0]; o.C0.x += in.oma * Cb[0]; }
        if (k == 1) { o.K.y += kb; o.lam.y += F0b[1]; o.C0.y += in.oma * Cb[1]; }
        if (k == 2) { o.K.z += kb; o.lam.z += F0b[2]; o.C0.z += in.oma * Cb[2]; }
    }
    DV dALb{0, 0, 0}, dAAb{0, 0, 0}, dBAb{0, 0, 0};
    for (int k = 0; k < 3; ++k)
    {
        DV n = in.n[k];
        dALb += n * Cb[k];
        o.n[k] += (f.dqAL - f.dqBL) * Cb[k] + dcross(f.dqAA, f.rAW) * Cb[k] - dcross(f.dqBA, f.rBW) * Cb[k];
        rAWb += dcross(n, f.dqAA) * Cb[k];
        rBWb -= dcross(n, f.dqBA) * Cb[k];
        dAAb += dcross(f.rAW, n) * Cb[k];
        dBAb -= dcross(f.rBW, n) * Cb[k];
    }
    o.pA += dALb;
    o.iLA -= dALb;
    o.pB -= dALb;
    o.iLB += dALb;
    dqdiff_adj(in.qA, in.iAA, dAAb, o.qA, o.iAA);
    dqdiff_adj(in.qB, in.iAB, dBAb, o.qB, o.iAB);
    if (!Round)
    {
        drot_adj(in.qA, in.rA, rAWb, o.qA, o.rA);
        drot_adj(in.qB, in.rB, rBWb, o.qB, o.rB);
        return;
    }
    offset_adj(in.code, true, in.radA, in.hlA, in.qA, in.rA, in.n[0], rAWb, o.qA, o.rA, o.n[0]);
    offset_adj(in.code, false, in.radB, in.hlB, in.qB, in.rB, in.n[0], rBWb, o.qB, o.rB, o.n[0]);
}

__device__ __forceinline__ void cout_zero(COut &o)
{
    DV z{0, 0, 0};
    DQ zq{0, 0, 0, 0};
    o.pA = o.pB = o.iLA = o.iLB = o.rA = o.rB = o.C0 = o.K = o.lam = z;
    o.n[0] = o.n[1] = o.n[2] = z;
    o.qA = o.qB = o.iAA = o.iAB = zq;
    o.mu = 0.0;
}

// Loads one contact as seen at a point of the step: poses given, penalty/lambda from the tape.
__device__ __forceinline__ void cin_load(CIn &in, const float4 *cr, float4 pen, float4 lam, float4 pA, float4 qA,
                                         float4 pB, float4 qB, float4 iLA, float4 iAA, float4 iLB, float4 iAB,
                                         float oma)
{
    in.pA = dv4(pA);
    in.qA = dq4(qA);
    in.pB = dv4(pB);
    in.qB = dq4(qB);
    in.iLA = dv4(iLA);
    in.iAA = dq4(iAA);
    in.iLB = dv4(iLB);
    in.iAB = dq4(iAB);
    in.rA = dv4(cr[CR_RA]);
    in.rB = dv4(cr[CR_RB]);
    in.C0 = dv4(cr[CR_C0]);
    in.n[0] = dv4(cr[CR_N]);
    in.n[1] = dv4(cr[CR_T1]);
    in.n[2] = dv4(cr[CR_T2]);
    in.K = dv4(pen);
    in.lam = dv4(lam);
    in.mu = cr[CR_RA].w;
    in.oma = (double)oma;
    in.code = __float_as_int(cr[CR_RB].w);
    in.radA = cr[CR_N].w;
    in.hlA = cr[CR_T2].w;
    in.radB = cr[CR_T1].w;
    in.hlB = cr[CR_C0].w;
}

__device__ __forceinline__ bool eval_clamp0(const ContactEval &ev, float4 pen, float4 lam)
{
    // F0 = min(k C0 + lambda, 0): clamped when the unclamped value is not negative
    float f = pen.x * ev.C0 + lam.x;
    return !(f < 0.0f);
}
__device__ __forceinline__ bool eval_scaled(const ContactEval &ev) { return ev.fs > ev.bounds && ev.fs > 0.0f; }

// ---- 6x6 symmetric solve in FP64 (the LDL^T of maths.h:481-547, unpivoted) ----
__device__ __forceinline__ void solve6d(double H[6][6], const double b[6], double x[6])
{
    double L[6][6], D[6], rD[6];  // rD: the six reciprocals, the only divisions
    for (int i = 0; i < 6; ++i)
    {
        for (int j = 0; j < i; ++j)
        {
            double s = H[i][j];
            for (int k = 0; k < j; ++k)
                s -= L[i][k] * L[j][k] * D[k];
            L[i][j] = s * rD[j];
        }
        double d = H[i][i];
        for (int k = 0; k < i; ++k)
            d -= L[i][k] * L[i][k] * D[k];
        D[i] = d;
        rD[i] = 1.0 / d;
    }
    double y[6];
    for (int i = 0; i < 6; ++i)
    {
        double s = b[i];
        for (int k = 0; k < i; ++k)
            s -= L[i][k] * y[k];
        y[i] = s;
    }
    for (int i = 5; i >= 0; --i)
    {
        double s = y[i] * rD[i];
        for (int k = i + 1; k < 6; ++k)
            s -= L[k][i] * x[k];
        x[i] = s;
    }
}

// ------------------------------------------------------------------------------------------
// reverse kernels
// ------------------------------------------------------------------------------------------

// finishStep (solver.cpp:1275-1288): prevVel' = v, and for dynamic bodies v' = (x - x0)/dt,
// w' = (q - q0)/dt.
__device__ __forceinline__ void adj_finish_body(const Bodies &B, const Params &P, int i, AdjBuf &A)
{
    DV vb = ld_dv(A.vb + 3 * i), wb = ld_dv(A.wb + 3 * i), pvb = ld_dv(A.pvb + 3 * i);
    if (B.massMoment[i].x > 0.0f)
    {
        double idt = 1.0 / (double)P.dt;
        add_dv(A.xb + 3 * i, vb * idt);
        add_dv(A.ilb + 3 * i, vb * (-idt));
        DQ qb = ld_dq(A.qb + 4 * i), iab = ld_dq(A.iab + 4 * i);
        dqdiff_adj(dq4(B.pose[4 * i + P_ROT]), dq4(B.pose[4 * i + P_IANG]), wb * idt, qb, iab);
        st_dq(A.qb + 4 * i, qb);
        st_dq(A.iab + 4 * i, iab);
        st_dv(A.vb + 3 * i, pvb);
        st_dv(A.wb + 3 * i, DV{0, 0, 0});
    }
    else
        st_dv(A.vb + 3 * i, vb + pvb);
    st_dv(A.pvb + 3 * i, DV{0, 0, 0});
}

// Start of a step's reversal: contact adjoints of the step's final contacts come from the carry
// (written by the reversal of the next step, only ever at slots of these contacts, and cleared here
// as they are taken); everything per step is cleared; then (finish) finishStep is reversed per
// body. Worlds with spheres reverse their rolling friction in between (k_adj_rolling, then
// k_adj_finish).
__global__ void k_adj_begin(Bodies B, Params P, const int4 *contactRec, int nc, int nm, int nb, bool finish, AdjBuf A)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nc)
    {
        int4 r = contactRec[t];
        int sl = r.x * kMaxContacts + r.y;
        for (int k = 0; k < 3; ++k)
        {
            A.kb[3 * sl + k] = A.ckb[3 * sl + k];
            A.lb[3 * sl + k] = A.clb[3 * sl + k];
            A.rAb[3 * sl + k] = A.crAb[3 * sl + k];
            A.rBb[3 * sl + k] = A.crBb[3 * sl + k];
            A.c0b[3 * sl + k] = 0.0;
            A.ckb[3 * sl + k] = 0.0;
            A.clb[3 * sl + k] = 0.0;
            A.crAb[3 * sl + k] = 0.0;
            A.crBb[3 * sl + k] = 0.0;
        }
    }
    if (t < nm)
    {
        for (int k = 0; k < 9; ++k)
            A.basisb[9 * t + k] = 0.0;
        A.mumb[t] = 0.0;
    }
    if (t < nb)
    {
        for (int k = 0; k < 3; ++k)
        {
            A.ilb[3 * t + k] = 0.0;
            A.nlb[3 * t + k] = 0.0;
        }
        for (int k = 0; k < 4; ++k)
        {
            A.iab[4 * t + k] = 0.0;
            A.nab[4 * t + k] = 0.0;
        }
        if (finish)
            adj_finish_body(B, P, t, A);
    }
}

__global__ void k_adj_finish(Bodies B, Params P, int nb, AdjBuf A)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nb)
        adj_finish_body(B, P, t, A);
}

// applySphereRollingFriction (solver.cpp:628-660) reversed, one thread per solved body. A dynamic
// sphere's velocities after finishStep (recomputed from its pose records as k_finish forms them)
// pass, once per contact point of each manifold with a static partner where it is the pair's
// chosen sphere, through v -= blend slip, w += blend (roll - w), in its stamps' order (the
// force-list order the forward visits). The end-of-step velocity adjoints go back through those
// passes (each pass's input recomputed from the start, so nothing is stored), the normal's
// adjoint to the manifold's basis row 0, and blend = clamp(0.08 sqrt(mu_s mu_o), 0, 0.18) hands
// its adjoint to both friction coefficients.
__global__ void k_adj_rolling(Bodies B, Manifolds M, const int4 *stamps, const int4 *tasks, int numTasks, Params P,
                              AdjBuf A)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= numTasks)
        return;
    const int4 task = tasks[k];
    const int i = task.x;
    const float4 hs = B.shape[i];
    if ((int)hs.z != kShapeSphere || !(B.massMoment[i].x > 0.0f))
        return;
    const float r = hs.x, fi = B.friction[i];
    const V3 v0 = (ld3(B.pose[4 * i + P_POS]) - ld3(B.pose[4 * i + P_ILIN])) / P.dt;
    const V3 w0 = qdiff(ldq(B.pose[4 * i + P_ROT]), ldq(B.pose[4 * i + P_IANG])) / P.dt;
    // visit(limit, f): the passes in forward order, f(pass, m, isA, o, mu, blend, nrm) for pass < limit;
    // returns the pass count
    auto visit = [&](int limit, auto &&f) {
        int pass = 0, last = -1;
        for (int t = 0; t < task.z; ++t)
        {
            const int4 st = stamps[task.y + t];
            if (adjKind(st.x) != kAdjManifold)
                continue;
            const int m = st.w / kMaxContacts;
            if (m == last)
                continue;  // the manifold's other contact points: its passes are counted below
            last = m;
            const int2 ab = M.ab[m];
            const bool isA = ab.x == i;
            if (!isA && (int)B.shape[ab.x].z == kShapeSphere)
                continue;
            const int o = isA ? ab.y : ab.x, nc = M.count[m];
            if (B.massMoment[o].x > 0.0f || nc <= 0)
                continue;
            const float mu = sqrtf(fi * B.friction[o]);
            if (mu <= 0.0f)
                continue;
            const float blend = clampr(mu * 0.08f, 0.0f, 0.18f);
            const V3 n0 = ld3(M.cr[(size_t)m * kMaxContacts * kCR + CR_N]);
            const V3 nrm = isA ? n0 : -n0;
            for (int c = 0; c < nc; ++c, ++pass)
                if (pass < limit)
                    f(pass, m, isA, o, mu, blend, nrm);
        }
        return pass;
    };
    const int passes = visit(0, [](int, int, bool, int, float, float, V3) {});
    if (passes == 0)
        return;
    DV vb = ld_dv(A.vb + 3 * i), wb = ld_dv(A.wb + 3 * i);
    for (int p = passes - 1; p >= 0; --p)
    {
        // this pass's input velocities, and its constants
        V3 v = v0, w = w0, nrm{0, 0, 0};
        float blend = 0.0f, mu = 0.0f;
        int m = 0, o = 0;
        bool isA = true;
        visit(p + 1, [&](int q, int mq, bool aq, int oq, float muq, float bq, V3 nq) {
            if (q < p)
            {
                const V3 rW = -nq * r;
                const V3 cv = v + cross(w, rW);
                const V3 slip = cv - nq * dot(cv, nq);
                v = v - slip * bq;
                const V3 roll = cross(nq, v) / r;
                w = w + (roll - w) * bq;
            }
            else
            {
                m = mq, isA = aq, o = oq, mu = muq, blend = bq, nrm = nq;
            }
        });
        const DV n = dv(nrm), vd = dv(v), wd = dv(w);
        const double bl = blend, rr = r;
        const DV rW = n * (-rr);
        const DV cv = vd + dcross(wd, rW);
        const DV slip = cv - n * ddot(cv, n);
        const DV v1 = vd - slip * bl;
        const DV roll = dcross(n, v1) * (1.0 / rr);
        double blb = ddot(wb, roll - wd);
        const DV rollb = wb * bl;
        DV wprev = wb * (1.0 - bl);
        DV v1b = vb + dcross(rollb, n) * (1.0 / rr);
        DV nb = dcross(v1, rollb) * (1.0 / rr);
        DV vprev = v1b;
        const DV slipb = v1b * (-bl);
        blb -= ddot(v1b, slip);
        const double ns = ddot(n, slipb), cn = ddot(cv, n);
        const DV cvb = slipb - n * ns;
        nb += slipb * (-cn) - cv * ns;
        vprev += cvb;
        wprev += dcross(rW, cvb);
        const DV rWb = dcross(cvb, wd);
        nb += rWb * (-rr);
        vb = vprev;
        wb = wprev;
        atomic_add3(A.basisb + 9 * m, isA ? nb : nb * -1.0);
        const float x = mu * 0.08f;
        if (x >= 0.0f && x <= 0.18f)
        {
            const double mub = blb * 0.08, fo = B.friction[o];
            aadd(A.mub + i, mub * fo / (2.0 * (double)mu));
            aadd(A.mub + o, mub * (double)fi / (2.0 * (double)mu));
        }
    }
    st_dv(A.vb + 3 * i, vb);
    st_dv(A.wb + 3 * i, wb);
}


// Pushes a contact's input adjoints out: bodies and init poses (atomic), the contact's own slot
// (exclusive) and its manifold (atomic unless `exclusive`).
__device__ __forceinline__ void cout_store(const COut &o, int a, int b, int sl, int m, bool massA, bool massB,
                                           bool exclusive, AdjBuf &A)
{
    if (massA)
    {
        atomic_add3(A.xb + 3 * a, o.pA);
        atomic_add4(A.qb + 4 * a, o.qA);
        atomic_add3(A.ilb + 3 * a, o.iLA);
        atomic_add4(A.iab + 4 * a, o.iAA);
    }
    if (massB)
    {
        atomic_add3(A.xb + 3 * b, o.pB);
        atomic_add4(A.qb + 4 * b, o.qB);
        atomic_add3(A.ilb + 3 * b, o.iLB);
        atomic_add4(A.iab + 4 * b, o.iAB);
    }
    add_dv(A.c0b + 3 * sl, o.C0);
    add_dv(A.rAb + 3 * sl, o.rA);
    add_dv(A.rBb + 3 * sl, o.rB);
    if (exclusive)
    {
        for (int k = 0; k < 3; ++k)
            add_dv(A.basisb + 9 * m + 3 * k, o.n[k]);
        A.mumb[m] += o.mu;
    }
    else
    {
        for (int k = 0; k < 3; ++k)
            atomic_add3(A.basisb + 9 * m + 3 * k, o.n[k]);
        aadd(A.mumb + m, o.mu);
    }
}

// Dual update of one contact, reversed (manifold.cpp:231-280). Poses after this iteration's
// primal (tapeCur), penalty/lambda before the dual (tapeLamPrev), branch flags by contact slot.
// 64-thread blocks (the launch's shape): the bound lets the compiler keep the FP64 chain in registers.
template <bool Round>
__global__ __launch_bounds__(64) void k_adj_dual(Bodies B, Manifolds M, const int4 *contactRec, int nc, const float4 *tapeCur,
                           const float4 *tapeLamPrev, const unsigned char *flags, Params P, AdjBuf A)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc)
        return;
    int4 r = contactRec[t];
    int m = r.x, sl = r.x * kMaxContacts + r.y, a = r.z, b = r.w;
    const float4 *cr = M.cr + (size_t)sl * kCR;
    float4 pen4 = tapeLamPrev[2 * sl], lam4 = tapeLamPrev[2 * sl + 1];
    float4 pA = tapeCur[2 * a], qA = tapeCur[2 * a + 1], pB = tapeCur[2 * b], qB = tapeCur[2 * b + 1];
    float4 iA = B.pose[4 * a + P_ILIN], jA = B.pose[4 * a + P_IANG], iB = B.pose[4 * b + P_ILIN],
           jB = B.pose[4 * b + P_IANG];
    float oma = 1.0f - P.alpha;
    const int fl = flags[sl];
    CIn in;
    cin_load(in, cr, pen4, lam4, pA, qA, pB, qB, iA, jA, iB, jB, oma);
    in.clamp0 = (fl & kFlagClamp0) != 0;
    in.scaled = (fl & kFlagScaled) != 0;
    CFwd f;
    cfwd<Round>(in, f);
    // adjoints of the updated lambda and penalty
    DV lbn = ld_dv(A.lb + 3 * sl), kbn = ld_dv(A.kb + 3 * sl);
    double Fb[3] = {lbn.x, lbn.y, lbn.z};
    double Cx[3] = {0.0, 0.0, 0.0};
    double kold[3] = {kbn.x, kbn.y, kbn.z};
    const double beta = P.betaLin;
    // unrolled, so the three rows stay in registers (a rolled loop indexes them from scratch memory)
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const bool upd = (fl & (k == 0 ? kFlagUpd0 : kFlagUpd12)) != 0;
        const bool sat = (fl & (kFlagSat0 << k)) != 0;
        const double kb = k == 0 ? kbn.x : (k == 1 ? kbn.y : kbn.z);
        const double sg = f.C[k] < 0.0 ? -1.0 : (f.C[k] > 0.0 ? 1.0 : 0.0);
        kold[k] = upd && sat ? 0.0 : kold[k];
        Cx[k] = upd && !sat ? kb * beta * sg : 0.0;
    }
    COut o;
    cout_zero(o);
    cadj<Round>(in, f, Fb, Cx, DV{0, 0, 0}, DV{0, 0, 0}, o);
    // the dual overwrote lambda and penalty: their old adjoints are what cadj produced, plus the
    // penalty carried through the update
    st_dv(A.lb + 3 * sl, o.lam);
    st_dv(A.kb + 3 * sl, DV{kold[0], kold[1], kold[2]} + o.K);
    o.lam = o.K = DV{0, 0, 0};
    cout_store(o, a, b, sl, m, B.massMoment[a].x > 0.0f, B.massMoment[b].x > 0.0f, false, A);
}

// Primal update of the bodies of colour `col`, reversed (solver.cpp:1184-1240,
// manifold.cpp:176-229); kAdjLanes lanes per body, two waves per block. tapePrev: poses before
// this iteration, tapeCur: after it (lower colours already updated); tapeLamPrev: penalty/lambda
// the iteration's primal read. The solver wave's lane 0 of each group solves H w = dx_bar in FP64
// with the step's taped system and dx, while the stamp wave loads its first round of stamps and
// re-evaluates their forces (the two do not depend on each other, so their latencies overlap);
// after the block barrier the stamp wave reverses every stamp: own-pose adjoints are summed over
// the lanes, partner ones go out atomically. A whole wave per body: the reversal is bound by the
// latency of its slowest waves (a colour of a small world is a few hundred bodies, far fewer waves
// than SIMDs), so one round of stamps for bodies of up to 64 stamps and one body per block took the
// projectile workload's backward from 289 to 242 ms (16 lanes per body: 4 bodies per block).
constexpr int kAdjLanes = 64;
constexpr int kAdjBodies = 64 / kAdjLanes;  // bodies per block
constexpr int kAdjThreads = 128;            // a solver wave and a stamp wave

struct AdjBodyShared
{
    double G[27];      // G_LL (9), G_AA (9), G_AL (9)
    double rLb[3], rAb[3];
    double xb[3], qb[4];  // the body's pose adjoints through its update and inertia terms
};

// One spring stamp of body i reversed (Spring::updatePrimal, spring.cpp:27-69): with rW the own
// anchor, s = +1 for body A and -1 for body B, the record is H_LL += k gl gl^T, H_AA += k ga ga^T,
// H_AL += k ga gl^T, r_L += f gl, r_A += f ga for gl = s n, ga = s (rW x n), n = d / |d|,
// f = k (|d| - rest), d = (rotate(qA, rA) + xA) - (rotate(qB, rB) + xB). G: the symmetrised
// adjoint of the body's 6x6 matrix (G_LL, G_AA, G_AL), rLb / rAb: of its right-hand side. The
// forward values are the FP32 step's own (as spring_record computes them); the pose adjoints of
// body A and B come out in xAb, qAb, xBb, qBb.
__device__ __forceinline__ void spring_adj(const Spring &sp, bool ownA, float4 pA4, float4 qA4, float4 pB4,
                                           float4 qB4, const double *G, DV rLb, DV rAb, DV &xAb, DQ &qAb, DV &xBb,
                                           DQ &qBb)
{
    const Q4 qA = ldq(qA4), qB = ldq(qB4);
    const V3 rAw = rotate(qA, ld3(sp.rA)), rBw = rotate(qB, ld3(sp.rB));
    const V3 d = (rAw + ld3(pA4)) - (rBw + ld3(pB4));
    const float dl = len(d);
    if (dl <= 1.0e-6f)
        return;
    const V3 n = d / dl;
    const float k = sp.rB.w;
    const float f = k * (dl - sp.rA.w);
    const V3 rW = ownA ? rAw : rBw;
    const double sg = ownA ? 1.0 : -1.0;
    const DV nd = dv(n), rWd = dv(rW);
    const DV gl = nd * sg, ga = dv(ownA ? cross(rW, n) : -cross(rW, n));
    const double *GLL = G, *GAA = G + 9, *GAL = G + 18;
    auto mv = [](const double *M, DV v) {
        return DV{M[0] * v.x + M[1] * v.y + M[2] * v.z, M[3] * v.x + M[4] * v.y + M[5] * v.z,
                  M[6] * v.x + M[7] * v.y + M[8] * v.z};
    };
    auto mtv = [](const double *M, DV v) {
        return DV{M[0] * v.x + M[3] * v.y + M[6] * v.z, M[1] * v.x + M[4] * v.y + M[7] * v.z,
                  M[2] * v.x + M[5] * v.y + M[8] * v.z};
    };
    const double kd = k, fd = f;
    const double fb = ddot(gl, rLb) + ddot(ga, rAb);
    const DV glb = rLb * fd + (mv(GLL, gl) + mtv(GAL, ga)) * kd;
    const DV gab = rAb * fd + (mv(GAA, ga) + mv(GAL, gl)) * kd;
    // gl = s n, ga = s (rW x n)
    DV nb = glb * sg + dcross(gab, rWd) * sg;
    const DV rWb = dcross(nd, gab) * sg;
    // f = k (|d| - rest), n = d / |d|
    const double dlb = kd * fb;
    const double dld = dl;
    // n = d / |d| and |d| = sqrt(d . d) differentiated at the FP32 d and |d| (as the step's values)
    const DV dd = dv(d);
    const DV db = nb * (1.0 / dld) - dd * (ddot(dd, nb) / (dld * dld * dld)) + dd * (dlb / dld);
    DV rAwb = db, rBwb = -db;
    if (ownA)
        rAwb += rWb;
    else
        rBwb += rWb;
    xAb += db;
    xBb -= db;
    DV unusedA{0, 0, 0}, unusedB{0, 0, 0};
    drot_adj(dq4(qA4), dv4(sp.rA), rAwb, qAb, unusedA);
    drot_adj(dq4(qB4), dv4(sp.rB), rBwb, qBb, unusedB);
}

// ---- joints (joint.cpp:35-182) ----
// A joint's poses as the reversal sees them: body A (or the world anchor for a = -1) and body B.
struct JointPoses
{
    float4 pA, qA, pB, qB;
};

// Linear part of a joint at poses `jp` (FP32, as joint_record / dual_joint compute it):
// rAW = rotate(qA, rA) (0 for a world anchor), rBW = rotate(qB, rB), C = (xA + rAW) - (xB + rBW),
// stabilised by -C0 alpha for an infinite stiffness.
__device__ __forceinline__ void joint_lin(const Joint &j, const JointPoses &jp, float alpha, V3 &rAW, V3 &rBW, V3 &C)
{
    const Q4 qB = ldq(jp.qB);
    rAW = j.a >= 0 ? rotate(ldq(jp.qA), ld3(j.rA)) : V3{0.0f, 0.0f, 0.0f};
    rBW = rotate(qB, ld3(j.rB));
    C = (j.a >= 0 ? ld3(jp.pA) + rAW : ld3(j.rA)) - (ld3(jp.pB) + rBW);
    if (isinf(j.rA.w))
        C -= ld3(j.c0Lin) * alpha;
}
// Angular part: C = (qA - qB) arm, stabilised likewise.
__device__ __forceinline__ V3 joint_ang(const Joint &j, const JointPoses &jp, float alpha)
{
    const Q4 qa = j.a >= 0 ? ldq(jp.qA) : Q4{0.0f, 0.0f, 0.0f, 1.0f};
    V3 C = qdiff(qa, ldq(jp.qB)) * j.c0Lin.w;
    if (isinf(j.rB.w))
        C -= ld3(j.c0Ang) * alpha;
    return C;
}

// Pose and C0 adjoints of a joint's constraint values: Clb of the linear C, Cab of the angular C.
__device__ __forceinline__ void joint_c_adj(const Joint &j, const JointPoses &jp, float alpha, DV Clb, DV Cab,
                                            DV rAWb, DV rBWb, DV &xAb, DQ &qAb, DV &xBb, DQ &qBb, DV &c0lb, DV &c0ab)
{
    if (j.a >= 0)
    {
        xAb += Clb;
        rAWb += Clb;
    }
    xBb -= Clb;
    rBWb -= Clb;
    if (isinf(j.rA.w))
        c0lb -= Clb * (double)alpha;
    if (isinf(j.rB.w))
        c0ab -= Cab * (double)alpha;
    DV unused{0, 0, 0};
    if (j.a >= 0)
        drot_adj(dq4(jp.qA), dv4(j.rA), rAWb, qAb, unused);
    drot_adj(dq4(jp.qB), dv4(j.rB), rBWb, qBb, unused);
    const DQ qa = j.a >= 0 ? dq4(jp.qA) : DQ{0, 0, 0, 1};
    DQ qab{0, 0, 0, 0};
    dqdiff_adj(qa, dq4(jp.qB), Cab * (double)j.c0Lin.w, qab, qBb);
    if (j.a >= 0)
        qAb += qab;
}

// One joint stamp of body i reversed (Joint::updatePrimal, joint.cpp:56-123), part 0 (the linear
// rows: jL = +-I, jA = skew(-rAW) or skew(rBW), K = diag(penalty), F = K C + lambda), part 1 (the
// ball joint's geometric-stiffness diagonal: diag |column b of sum_k F_k G_k(rr)|) or part 2 (the
// angular rows, jA = +-arm I). G / rLb / rAb: the body system's adjoints; the joint state's
// adjoints go to jsb (kJointAdj), the C0 adjoints to jc0 (kJointC0), the poses' to xAb ... qBb.
__device__ __forceinline__ void joint_stamp_adj(const Joint &j, int part, bool isA, const JointPoses &jp, float alpha,
                                                const double *G, DV rLb, DV rAb, double *jsb, double *jc0,
                                                DV &xAb, DQ &qAb, DV &xBb, DQ &qBb)
{
    const double *GLL = G, *GAA = G + 9, *GAL = G + 18;
    const double sg = isA ? 1.0 : -1.0;
    DV Clb{0, 0, 0}, Cab{0, 0, 0}, rAWb{0, 0, 0}, rBWb{0, 0, 0}, c0lb{0, 0, 0}, c0ab{0, 0, 0};
    if (part < 2)
    {
        const V3 penL = ld3(j.penLin);
        if (!(lensq(penL) > 0.0f))
            return;
        V3 rAW, rBW, C;
        joint_lin(j, jp, alpha, rAW, rBW, C);
        const V3 F = V3{penL.x * C.x, penL.y * C.y, penL.z * C.z} + ld3(j.lamLin);
        const double K[3] = {penL.x, penL.y, penL.z}, Cd[3] = {C.x, C.y, C.z}, Fd[3] = {F.x, F.y, F.z};
        double Fb[3] = {0, 0, 0}, Kb[3] = {0, 0, 0};
        if (part == 0)
        {
            const DV rW = dv(isA ? rAW : rBW);
            DV rWb{0, 0, 0};
            for (int k = 0; k < 3; ++k)
            {
                const DV n{k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
                const DV av = dcross(rW, n) * sg;
                const DV GLLn{GLL[k], GLL[3 + k], GLL[6 + k]};
                const DV GAAa{GAA[0] * av.x + GAA[1] * av.y + GAA[2] * av.z, GAA[3] * av.x + GAA[4] * av.y + GAA[5] * av.z,
                              GAA[6] * av.x + GAA[7] * av.y + GAA[8] * av.z};
                const DV GALn{GAL[k], GAL[3 + k], GAL[6 + k]};
                Kb[k] += 0.5 * ddot(n, GLLn) + 0.5 * ddot(av, GAAa) + sg * ddot(av, GALn);
                Fb[k] += sg * dcomp(rLb, k) + ddot(av, rAb);
                const DV ab = rAb * Fd[k] + GAAa * K[k] + GALn * (K[k] * sg);
                rWb += dcross(n, ab) * sg;
            }
            if (isA)
                rAWb += rWb;
            else
                rBWb += rWb;
        }
        else
        {
            // column b of H = sum_k F_k G_k(rr): c_b[a] = -(F . rr) [a == b] + rr_a F_b, as FP32
            const V3 rr = isA ? rAW : -rBW;
#define AVBD_GS_H(a, b) ((gs_entry<0, a, b>(rr) * F.x + gs_entry<1, a, b>(rr) * F.y) + gs_entry<2, a, b>(rr) * F.z)
            const V3 c0{AVBD_GS_H(0, 0), AVBD_GS_H(1, 0), AVBD_GS_H(2, 0)};
            const V3 c1{AVBD_GS_H(0, 1), AVBD_GS_H(1, 1), AVBD_GS_H(2, 1)};
            const V3 c2{AVBD_GS_H(0, 2), AVBD_GS_H(1, 2), AVBD_GS_H(2, 2)};
#undef AVBD_GS_H
            const V3 cols[3] = {c0, c1, c2};
            DV cb[3];
            for (int b = 0; b < 3; ++b)
            {
                const float lb = len(cols[b]);
                cb[b] = lb > 0.0f ? dv(cols[b]) * (0.5 * GAA[4 * b] / (double)lb) : DV{0, 0, 0};
            }
            const DV rrd = dv(rr);
            const double trace = cb[0].x + cb[1].y + cb[2].z;
            DV rrb = DV{Fd[0], Fd[1], Fd[2]} * (-trace);
            for (int b = 0; b < 3; ++b)
                rrb += cb[b] * Fd[b];
            for (int k = 0; k < 3; ++k)
                Fb[k] += -dcomp(rrd, k) * trace + ddot(cb[k], rrd);
            if (isA)
                rAWb += rrb;
            else
                rBWb -= rrb;
        }
        for (int k = 0; k < 3; ++k)
        {
            Kb[k] += Fb[k] * Cd[k];
            jsb[JA_PL + k] += Kb[k];
            jsb[JA_LL + k] += Fb[k];
        }
        Clb = DV{K[0] * Fb[0], K[1] * Fb[1], K[2] * Fb[2]};
    }
    else
    {
        const V3 penA = ld3(j.penAng);
        if (!(lensq(penA) > 0.0f))
            return;
        const V3 C = joint_ang(j, jp, alpha);
        const double arm = j.c0Lin.w, K[3] = {penA.x, penA.y, penA.z}, Cd[3] = {C.x, C.y, C.z};
        for (int k = 0; k < 3; ++k)
        {
            const double Fb = sg * arm * dcomp(rAb, k);
            jsb[JA_PA + k] += 0.5 * GAA[4 * k] * arm * arm + Fb * Cd[k];
            jsb[JA_LA + k] += Fb;
            if (k == 0) Cab.x = K[0] * Fb;
            if (k == 1) Cab.y = K[1] * Fb;
            if (k == 2) Cab.z = K[2] * Fb;
        }
    }
    joint_c_adj(j, jp, alpha, Clb, Cab, rAWb, rBWb, xAb, qAb, xBb, qBb, c0lb, c0ab);
    for (int k = 0; k < 3; ++k)
    {
        jc0[k] += dcomp(c0lb, k);
        jc0[3 + k] += dcomp(c0ab, k);
    }
}

// Joint dual of one iteration reversed (Joint::updateDual, joint.cpp:125-182): before / after its
// update (tape slots it and it + 1), poses after the iteration's primal. With an infinite
// stiffness lambda' = K C + lambda; penalty' = min(penalty + beta |C|, cap) where not capped; a
// joint that fractured in this update has all-zero outputs and passes nothing back.
__global__ void k_adj_joint_dual(const Joint *before, const Joint *after, int nj, const float4 *pose, Bodies B,
                                 Params P, AdjBuf A)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nj)
        return;
    const Joint j = before[q];
    if (j.dead)
        return;
    double *jb = A.jb + (size_t)kJointAdj * q;
    if (after[q].broken && !j.broken)
    {
        for (int k = 0; k < kJointAdj; ++k)
            jb[k] = 0.0;
        return;
    }
    JointPoses jp;
    jp.pA = j.a >= 0 ? pose[2 * j.a] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    jp.qA = j.a >= 0 ? pose[2 * j.a + 1] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    jp.pB = pose[2 * j.b];
    jp.qB = pose[2 * j.b + 1];
    DV Clb{0, 0, 0}, Cab{0, 0, 0};
    // linear rows, then angular rows: penalty and lambda adjoints before the update
    for (int part = 0; part < 2; ++part)
    {
        const V3 pen = part == 0 ? ld3(j.penLin) : ld3(j.penAng);
        if (!(lensq(pen) > 0.0f))
            continue;
        V3 C;
        if (part == 0)
        {
            V3 rAW, rBW;
            joint_lin(j, jp, P.alpha, rAW, rBW, C);
        }
        else
            C = joint_ang(j, jp, P.alpha);
        const float stiff = part == 0 ? j.rA.w : j.rB.w, beta = part == 0 ? P.betaLin : P.betaAng;
        const int op = part == 0 ? JA_PL : JA_PA, ol = part == 0 ? JA_LL : JA_LA;
        const float cap = fminr(stiff, kPenaltyMax);
        const float Cf[3] = {C.x, C.y, C.z}, pf[3] = {pen.x, pen.y, pen.z};
        double Cb[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k)
        {
            const double lb = jb[ol + k], pb = jb[op + k];
            double penb = 0.0;
            if (isinf(stiff))
            {
                penb += lb * (double)Cf[k];
                Cb[k] += (double)pf[k] * lb;
            }
            const float t = pf[k] + fabsf(Cf[k]) * beta;
            if (t < cap)
            {
                penb += pb;
                Cb[k] += (double)beta * (Cf[k] < 0.0f ? -1.0 : (Cf[k] > 0.0f ? 1.0 : 0.0)) * pb;
            }
            jb[op + k] = penb;
            jb[ol + k] = lb;
        }
        if (part == 0)
            Clb = DV{Cb[0], Cb[1], Cb[2]};
        else
            Cab = DV{Cb[0], Cb[1], Cb[2]};
    }
    DV xAb{0, 0, 0}, xBb{0, 0, 0}, rAWb{0, 0, 0}, rBWb{0, 0, 0}, c0lb{0, 0, 0}, c0ab{0, 0, 0};
    DQ qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
    joint_c_adj(j, jp, P.alpha, Clb, Cab, rAWb, rBWb, xAb, qAb, xBb, qBb, c0lb, c0ab);
    if (j.a >= 0 && B.massMoment[j.a].x > 0.0f)
    {
        atomic_add3(A.xb + 3 * j.a, xAb);
        atomic_add4(A.qb + 4 * j.a, qAb);
    }
    if (B.massMoment[j.b].x > 0.0f)
    {
        atomic_add3(A.xb + 3 * j.b, xBb);
        atomic_add4(A.qb + 4 * j.b, qBb);
    }
    atomic_add3(A.jc0b + kJointC0 * q, c0lb);
    atomic_add3(A.jc0b + kJointC0 * q + 3, c0ab);
}

// Start of a step's reversal for the joints: the state adjoint of the step's final joints is the
// carry the next step's init reversal left; the carry and the C0 adjoints are cleared.
__global__ void k_adj_joint_begin(int nj, AdjBuf A)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nj)
        return;
    for (int k = 0; k < kJointAdj; ++k)
    {
        A.jb[(size_t)kJointAdj * q + k] = A.cjb[(size_t)kJointAdj * q + k];
        A.cjb[(size_t)kJointAdj * q + k] = 0.0;
    }
    for (int k = 0; k < kJointC0; ++k)
        A.jc0b[(size_t)kJointC0 * q + k] = 0.0;
}

// Joint init reversed (Joint::initialize, joint.cpp:35-54): C0 lin = (xA + rotate(qA, rA)) -
// (xB + rotate(qB, rB)) and C0 ang = (qA - qB) arm at the step's start poses; lambda = lambda_prev
// alpha gamma; penalty = min(clamp(penalty_prev gamma, min, max), stiffness). The previous step's
// final state receives its adjoint as the carry; joints removed or fractured before this step pass
// nothing.
__global__ void k_adj_joint_init(const Joint *start, int nj, Bodies B, Params P, AdjBuf A)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nj)
        return;
    const Joint j = start[q];
    if (j.dead || j.broken)
        return;
    const double *jb = A.jb + (size_t)kJointAdj * q, *jc = A.jc0b + (size_t)kJointC0 * q;
    double *cj = A.cjb + (size_t)kJointAdj * q;
    const double ag = (double)P.alpha * (double)P.gamma;
    for (int k = 0; k < 3; ++k)
    {
        cj[JA_LL + k] = jb[JA_LL + k] * ag;
        cj[JA_LA + k] = jb[JA_LA + k] * ag;
    }
    const float pl[3] = {j.penLin.x, j.penLin.y, j.penLin.z}, pa[3] = {j.penAng.x, j.penAng.y, j.penAng.z};
    for (int k = 0; k < 3; ++k)
        for (int part = 0; part < 2; ++part)
        {
            const float v = (part == 0 ? pl[k] : pa[k]) * P.gamma, stiff = part == 0 ? j.rA.w : j.rB.w;
            const float v1 = clampr(v, kPenaltyMin, kPenaltyMax);
            const bool inside = !(kPenaltyMax < v) && !(kPenaltyMin > v) && v1 < stiff;
            cj[(part == 0 ? JA_PL : JA_PA) + k] = inside ? jb[(part == 0 ? JA_PL : JA_PA) + k] * (double)P.gamma : 0.0;
        }
    // C0 from the start poses (the x- rows of the pose records)
    JointPoses jp;
    jp.pA = j.a >= 0 ? B.pose[4 * j.a + P_ILIN] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    jp.qA = j.a >= 0 ? B.pose[4 * j.a + P_IANG] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    jp.pB = B.pose[4 * j.b + P_ILIN];
    jp.qB = B.pose[4 * j.b + P_IANG];
    DV Clb{jc[0], jc[1], jc[2]}, Cab{jc[3], jc[4], jc[5]};
    DV xAb{0, 0, 0}, xBb{0, 0, 0}, rAWb{0, 0, 0}, rBWb{0, 0, 0}, c0lb{0, 0, 0}, c0ab{0, 0, 0};
    DQ qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
    // C0 has no stabilisation term: a copy of the joint with finite stiffness flags
    Joint jf = j;
    jf.rA.w = 0.0f;
    jf.rB.w = 0.0f;
    joint_c_adj(jf, jp, P.alpha, Clb, Cab, rAWb, rBWb, xAb, qAb, xBb, qBb, c0lb, c0ab);
    if (j.a >= 0 && B.massMoment[j.a].x > 0.0f)
    {
        atomic_add3(A.xb + 3 * j.a, xAb);
        atomic_add4(A.qb + 4 * j.a, qAb);
    }
    if (B.massMoment[j.b].x > 0.0f)
    {
        atomic_add3(A.xb + 3 * j.b, xBb);
        atomic_add4(A.qb + 4 * j.b, qBb);
    }
}

__device__ __forceinline__ void wave_sum_group(double &v)
{
#pragma unroll
    for (int off = 1; off < kAdjLanes; off <<= 1)
        v += __shfl_xor(v, off, kAdjLanes);
}

template <bool Round>
__global__ __launch_bounds__(kAdjThreads) void k_adj_primal(Bodies B, Manifolds M, const Spring *springs,
                                                            const Joint *jointsTape, const int4 *stamps,
                                                            const int4 *tasks, int begin, int end, const int *color,
                                                            int col, const float4 *tapePrev, const float4 *tapeCur,
                                                            const float4 *tapeLamPrev, const float4 *sysTape,
                                                            const unsigned char *stampFlags, Params P, AdjBuf A)
{
    constexpr int L = kAdjLanes, G = kAdjBodies;
    __shared__ AdjBodyShared sh[G];
    const bool solver = threadIdx.x < 64;
    const int lane = (threadIdx.x & 63) % L, g = (threadIdx.x & 63) / L;
    const int k = begin + blockIdx.x * G + g;
    const bool valid = k < end;
    const int4 task = valid ? tasks[k] : make_int4(0, 0, 0, 0);
    const int i = task.x, s0 = task.y, ns = valid ? task.z : 0;
    auto seen = [&](int j) -> const float4 * {
        int cj = color[j];
        return (j != i && cj >= 0 && cj < col) ? tapeCur : tapePrev;
    };
    const float oma = 1.0f - P.alpha;
    AdjBodyShared &S = sh[g];
    const float4 mm = B.massMoment[i];
    const float4 rot4 = tapePrev[2 * i + 1];
    const V3 pos = ld3(tapePrev[2 * i]);
    // ---- solver wave, lane 0 of each group: the step's dx, the FP64 solve of H w = dx_bar, inertia ----
    if (solver && valid && lane == 0)
    {
        DV xb{0, 0, 0};
        DQ qb{0, 0, 0, 0};
        // the step's accumulated system and dx, as its primal kernel taped them
        float a[36];
        const float4 *src = sysTape + 9 * (size_t)i;
#pragma unroll
        for (int q = 0; q < 9; ++q)
        {
            float4 v = src[q];
            a[4 * q] = v.x;
            a[4 * q + 1] = v.y;
            a[4 * q + 2] = v.z;
            a[4 * q + 3] = v.w;
        }
        V3 dxL{a[27], a[28], a[29]}, dxA{a[30], a[31], a[32]};
        // x' = x + dxL, q' = q (+) dxA
        DV xbo = ld_dv(A.xb + 3 * i);
        DQ qbo = ld_dq(A.qb + 4 * i);
        xb = xbo;
        DV dxAb{0, 0, 0};
        DQ q = dq4(rot4);
        dqint_adj(q, dv(dxA), qbo, qb, dxAb);
        (void)oma;
        double H[6][6];
        const int lo[3][3] = {{R_L00, R_L10, R_L20}, {R_L10, R_L11, R_L21}, {R_L20, R_L21, R_L22}};
        const int an[3][3] = {{R_A00, R_A10, R_A20}, {R_A10, R_A11, R_A21}, {R_A20, R_A21, R_A22}};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
            {
                H[r][c] = a[lo[r][c]];
                H[3 + r][3 + c] = a[an[r][c]];
                H[3 + r][c] = a[R_C + 3 * r + c];
                H[c][3 + r] = a[R_C + 3 * r + c];
            }
        double g6[6] = {xbo.x, xbo.y, xbo.z, dxAb.x, dxAb.y, dxAb.z}, w6[6];
        solve6d(H, g6, w6);
        const double dx6[6] = {dxL.x, dxL.y, dxL.z, dxA.x, dxA.y, dxA.z};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
            {
                S.G[3 * r + c] = -(w6[r] * dx6[c] + w6[c] * dx6[r]);
                S.G[9 + 3 * r + c] = -(w6[3 + r] * dx6[3 + c] + w6[3 + c] * dx6[3 + r]);
                S.G[18 + 3 * r + c] = -(w6[3 + r] * dx6[c] + w6[c] * dx6[3 + r]);
            }
        DV rLb{-w6[0], -w6[1], -w6[2]}, rAb{-w6[3], -w6[4], -w6[5]};
        S.rLb[0] = rLb.x;
        S.rLb[1] = rLb.y;
        S.rLb[2] = rLb.z;
        S.rAb[0] = rAb.x;
        S.rAb[1] = rAb.y;
        S.rAb[2] = rAb.z;
        // inertia: H_LL += m/dt^2 I, H_AA += diag(I)/dt^2, r_L += m/dt^2 (x - xt), r_A += I/dt^2 o (q - qt)
        double idt2 = 1.0 / ((double)P.dt * (double)P.dt);
        // the inertial residuals with the values the primal used: x - xt and q (-) qt rounded in
        // FP32 (for a resting body x - xt is ~1e-4 m against a position ulp of ~5e-7 m, so the
        // exact difference of the FP32 poses would differ from the primal's by ~1e-2 relative in
        // the mass derivative, which these residuals multiply)
        DV dld = dv(pos - ld3(B.inerLin[i]));
        DQ inA = dq4(B.inerAng[i]);
        DV dad = dv(qdiff(ldq(rot4), ldq(B.inerAng[i])));
        double mbar = idt2 * (-(w6[0] * dx6[0] + w6[1] * dx6[1] + w6[2] * dx6[2]) + ddot(rLb, dld));
        double Ib[3];
        for (int c = 0; c < 3; ++c)
            Ib[c] = idt2 * (-(w6[3 + c] * dx6[3 + c]) + dcomp(rAb, c) * dcomp(dad, c));
        double mlD = (double)mm.x * idt2;
        xb += rLb * mlD;
        add_dv(A.nlb + 3 * i, rLb * (-mlD));
        DV dab{dcomp(rAb, 0) * (double)mm.y * idt2, dcomp(rAb, 1) * (double)mm.z * idt2,
               dcomp(rAb, 2) * (double)mm.w * idt2};
        DQ nab{0, 0, 0, 0};
        dqdiff_adj(q, inA, dab, qb, nab);
        DQ old = ld_dq(A.nab + 4 * i);
        st_dq(A.nab + 4 * i, old + nab);
        if (mm.x > 0.0f)
            mbar += (Ib[0] * (double)mm.y + Ib[1] * (double)mm.z + Ib[2] * (double)mm.w) / (double)mm.x;
        A.mb[i] += mbar;
        S.xb[0] = xb.x;
        S.xb[1] = xb.y;
        S.xb[2] = xb.z;
        S.qb[0] = qb.x;
        S.qb[1] = qb.y;
        S.qb[2] = qb.z;
        S.qb[3] = qb.w;
    }
    // ---- stamp wave: its stamps' forward values (the first round before the barrier) ----
    int4 st = make_int4(0, 0, 0, 0);
    int fl = 0;
    CIn in;
    CFwd f;
    auto load_stamp = [&](int t) {
        st = stamps[s0 + t];
        if (adjKind(st.x) != kAdjManifold)
            return;  // springs and joints are evaluated where they are reversed
        fl = stampFlags[s0 + t];
        const bool isA = (st.y >> 16) != 0;
        const int partner = st.z, sl = st.w;
        const int a = isA ? i : partner, b = isA ? partner : i;
        const float4 *cr = M.cr + (size_t)sl * kCR;
        const float4 *sa = a == i ? tapePrev : seen(a), *sb = b == i ? tapePrev : seen(b);
        float4 pen4 = tapeLamPrev[2 * sl], lam4 = tapeLamPrev[2 * sl + 1];
        float4 pA = sa[2 * a], qA = sa[2 * a + 1], pB = sb[2 * b], qB = sb[2 * b + 1];
        float4 iA = B.pose[4 * a + P_ILIN], jA = B.pose[4 * a + P_IANG], iB = B.pose[4 * b + P_ILIN],
               jB = B.pose[4 * b + P_IANG];
        cin_load(in, cr, pen4, lam4, pA, qA, pB, qB, iA, jA, iB, jB, oma);
        in.clamp0 = (fl & kFlagClamp0) != 0;
        in.scaled = (fl & kFlagScaled) != 0;
        cfwd<Round>(in, f);
    };
    if (!solver && valid && lane < ns)
        load_stamp(lane);
    __syncthreads();
    if (solver)
        return;
    // ---- stamp wave: every lane reverses its own stamps ----
    DV rLb{S.rLb[0], S.rLb[1], S.rLb[2]}, rAb{S.rAb[0], S.rAb[1], S.rAb[2]};
    DV oxb{0, 0, 0};
    DQ oqb{0, 0, 0, 0};
    DV oil{0, 0, 0};
    DQ oia{0, 0, 0, 0};
    for (int t = lane; valid && t < ns; t += L)
    {
        if (t != lane)
            load_stamp(t);
        if (adjKind(st.x) == kAdjJoint)
        {
            // the joint's state as this iteration's primal read it (tape slot of the iteration)
            const Joint &jt = jointsTape[adjIndex(st.x)];
            const bool ownA = jt.a == i;
            const int j = ownA ? jt.b : jt.a;
            JointPoses jp;
            const float4 *sj = j >= 0 ? seen(j) : tapePrev;
            const float4 pI = tapePrev[2 * i], qI = tapePrev[2 * i + 1];
            const float4 pJ = j >= 0 ? sj[2 * j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float4 qJ = j >= 0 ? sj[2 * j + 1] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            jp.pA = ownA ? pI : pJ;
            jp.qA = ownA ? qI : qJ;
            jp.pB = ownA ? pJ : pI;
            jp.qB = ownA ? qJ : qI;
            double jsb[kJointAdj], jc0[kJointC0];
            for (int q = 0; q < kJointAdj; ++q)
                jsb[q] = 0.0;
            for (int q = 0; q < kJointC0; ++q)
                jc0[q] = 0.0;
            DV xAb{0, 0, 0}, xBb{0, 0, 0};
            DQ qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
            joint_stamp_adj(jt, st.y & 0xFFFF, ownA, jp, P.alpha, S.G, rLb, rAb, jsb, jc0, xAb, qAb, xBb, qBb);
            oxb += ownA ? xAb : xBb;
            oqb += ownA ? qAb : qBb;
            if (j >= 0 && B.massMoment[j].x > 0.0f)
            {
                atomic_add3(A.xb + 3 * j, ownA ? xBb : xAb);
                atomic_add4(A.qb + 4 * j, ownA ? qBb : qAb);
            }
            // a joint's parts sit on several lanes: its adjoints are added atomically
            double *jb = A.jb + (size_t)kJointAdj * adjIndex(st.x), *jc = A.jc0b + (size_t)kJointC0 * adjIndex(st.x);
            for (int q = 0; q < kJointAdj; ++q)
                if (jsb[q] != 0.0)
                    aadd(jb + q, jsb[q]);
            for (int q = 0; q < kJointC0; ++q)
                if (jc0[q] != 0.0)
                    aadd(jc + q, jc0[q]);
            continue;
        }
        if (adjKind(st.x) == kAdjSpring)
        {
            const Spring &sp = springs[adjIndex(st.x)];
            const bool ownA = sp.a == i;
            const int j = ownA ? sp.b : sp.a;
            const float4 *sj = seen(j);
            const float4 pI = tapePrev[2 * i], qI = tapePrev[2 * i + 1], pJ = sj[2 * j], qJ = sj[2 * j + 1];
            DV xAb{0, 0, 0}, xBb{0, 0, 0};
            DQ qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
            if (ownA)
                spring_adj(sp, true, pI, qI, pJ, qJ, S.G, rLb, rAb, xAb, qAb, xBb, qBb);
            else
                spring_adj(sp, false, pJ, qJ, pI, qI, S.G, rLb, rAb, xAb, qAb, xBb, qBb);
            oxb += ownA ? xAb : xBb;
            oqb += ownA ? qAb : qBb;
            if (B.massMoment[j].x > 0.0f)
            {
                atomic_add3(A.xb + 3 * j, ownA ? xBb : xAb);
                atomic_add4(A.qb + 4 * j, ownA ? qBb : qAb);
            }
            continue;
        }
        const bool isA = (st.y >> 16) != 0;
        const int partner = st.z, sl = st.w, m = adjIndex(st.x);
        const int a = isA ? i : partner, b = isA ? partner : i;
        const double sgn = isA ? 1.0 : -1.0;
        DV rW = isA ? f.rAW : f.rBW;
        COut o;
        cout_zero(o);
        double Fb[3];
        DV rWb{0, 0, 0};
        for (int kk = 0; kk < 3; ++kk)
        {
            DV n = in.n[kk];
            DV av = dcross(rW, n) * sgn;
            double Kk = dcomp(in.K, kk);
            const double *GLL = S.G, *GAA = S.G + 9, *GAL = S.G + 18;
            DV GLLn{GLL[0] * n.x + GLL[1] * n.y + GLL[2] * n.z, GLL[3] * n.x + GLL[4] * n.y + GLL[5] * n.z,
                    GLL[6] * n.x + GLL[7] * n.y + GLL[8] * n.z};
            DV GAAa{GAA[0] * av.x + GAA[1] * av.y + GAA[2] * av.z, GAA[3] * av.x + GAA[4] * av.y + GAA[5] * av.z,
                    GAA[6] * av.x + GAA[7] * av.y + GAA[8] * av.z};
            DV GALn{GAL[0] * n.x + GAL[1] * n.y + GAL[2] * n.z, GAL[3] * n.x + GAL[4] * n.y + GAL[5] * n.z,
                    GAL[6] * n.x + GAL[7] * n.y + GAL[8] * n.z};
            DV GALta{GAL[0] * av.x + GAL[3] * av.y + GAL[6] * av.z, GAL[1] * av.x + GAL[4] * av.y + GAL[7] * av.z,
                     GAL[2] * av.x + GAL[5] * av.y + GAL[8] * av.z};
            // K adjoint: n^T Hb_LL n + a^T Hb_AA a + s a^T G_AL n (each Hb quadratic form is half its G form)
            double kb = 0.5 * ddot(n, GLLn) + 0.5 * ddot(av, GAAa) + sgn * ddot(av, GALn);
            if (kk == 0) o.K.x += kb;
            if (kk == 1) o.K.y += kb;
            if (kk == 2) o.K.z += kb;
            Fb[kk] = sgn * ddot(n, rLb) + ddot(av, rAb);
            DV nb = rLb * (f.F[kk] * sgn) + GLLn * Kk + GALta * (Kk * sgn);
            DV ab = rAb * f.F[kk] + GAAa * Kk + GALn * (Kk * sgn);
            // a = s (rW x n)
            rWb += dcross(n, ab) * sgn;
            nb += dcross(ab, rW) * sgn;
            o.n[kk] += nb;
        }
        double Cx[3] = {0.0, 0.0, 0.0};
        DV rAWb = isA ? rWb : DV{0, 0, 0}, rBWb = isA ? DV{0, 0, 0} : rWb;
        cadj<Round>(in, f, Fb, Cx, rAWb, rBWb, o);
        if (isA)
        {
            oxb += o.pA;
            oqb += o.qA;
            o.pA = DV{0, 0, 0};
            o.qA = DQ{0, 0, 0, 0};
        }
        else
        {
            oxb += o.pB;
            oqb += o.qB;
            o.pB = DV{0, 0, 0};
            o.qB = DQ{0, 0, 0, 0};
        }
        add_dv(A.kb + 3 * sl, o.K);
        add_dv(A.lb + 3 * sl, o.lam);
        o.K = o.lam = DV{0, 0, 0};
        bool massA = B.massMoment[a].x > 0.0f, massB = B.massMoment[b].x > 0.0f;
        // own init-pose adjoints stay in registers
        if (isA)
        {
            oil += o.iLA;
            oia += o.iAA;
            o.iLA = DV{0, 0, 0};
            o.iAA = DQ{0, 0, 0, 0};
        }
        else
        {
            oil += o.iLB;
            oia += o.iAB;
            o.iLB = DV{0, 0, 0};
            o.iAB = DQ{0, 0, 0, 0};
        }
        // contacts of one manifold can sit on different lanes of this body: manifold sums are atomic
        cout_store(o, a, b, sl, m, massA && a != i, massB && b != i, false, A);
    }
    wave_sum_group(oxb.x);
    wave_sum_group(oxb.y);
    wave_sum_group(oxb.z);
    wave_sum_group(oqb.x);
    wave_sum_group(oqb.y);
    wave_sum_group(oqb.z);
    wave_sum_group(oqb.w);
    wave_sum_group(oil.x);
    wave_sum_group(oil.y);
    wave_sum_group(oil.z);
    wave_sum_group(oia.x);
    wave_sum_group(oia.y);
    wave_sum_group(oia.z);
    wave_sum_group(oia.w);
    if (valid && lane == 0)
    {
        st_dv(A.xb + 3 * i, DV{S.xb[0], S.xb[1], S.xb[2]} + oxb);
        st_dq(A.qb + 4 * i, DQ{S.qb[0], S.qb[1], S.qb[2], S.qb[3]} + oqb);
        // no other thread of this launch touches body i (neighbours have other colours)
        add_dv(A.ilb + 3 * i, oil);
        DQ ia = ld_dq(A.iab + 4 * i);
        st_dq(A.iab + 4 * i, ia + oia);
    }
}

// Warm start (solver.cpp:1134-1169) reversed: the pose adjoints now refer to the step's start.
__global__ void k_adj_warmstart(Bodies B, Params P, const int *deg, const int *ignoreDeg, int n, AdjBuf A)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    float mass = B.massMoment[i].x;
    DV xwb = ld_dv(A.xb + 3 * i);        // adjoint of the warm-started pose
    DQ qwb = ld_dq(A.qb + 4 * i);
    DV nlb = ld_dv(A.nlb + 3 * i), ilb = ld_dv(A.ilb + 3 * i);
    DQ nab = ld_dq(A.nab + 4 * i), iab = ld_dq(A.iab + 4 * i);
    DV vb = ld_dv(A.vb + 3 * i), wb = ld_dv(A.wb + 3 * i), pvb = ld_dv(A.pvb + 3 * i);
    double dt = P.dt;
    DQ q = dq4(B.pose[4 * i + P_IANG]);  // the start pose (init = start)
    DV w = dv4(B.angVel[i]);
    DV xb = ilb, qbz{0, 0, 0};
    DQ qb = iab;
    // solved as the forward's warm start decided: forces or IgnoreCollision pairs make a body constrained
    bool solved = mass > 0.0f && (deg[i] > 0 || (ignoreDeg && ignoreDeg[i] > 0));
    if (mass > 0.0f && !solved)
    {
        // x' = inertial target, q' = inertial rotation target
        nlb += xwb;
        nab += qwb;
        xwb = DV{0, 0, 0};
        qwb = DQ{0, 0, 0, 0};
    }
    DQ qintb = nab;  // q (+) w dt feeds both the inertial target and (dynamic) the warm-started rotation
    if (mass > 0.0f)
    {
        // x' = x + v dt + (0,0,g) aw dt^2
        xb += xwb;
        vb += xwb * dt;
        qintb += qwb;
        float4 v4 = B.vel[i], pv4 = B.prevVel[i];
        float accelExt = ((v4.z - pv4.z) / P.dt) * signr(P.gravity);
        float awr = accelExt / fabsf(P.gravity);
        if (awr > 0.0f && awr < 1.0f && isfinite(awr))
        {
            double dawdvz = (double)signr(P.gravity) / ((double)P.dt * fabs((double)P.gravity));
            double awb = xwb.z * (double)P.gravity * dt * dt;
            vb.z += awb * dawdvz;
            pvb.z -= awb * dawdvz;
        }
    }
    else
    {
        xb += xwb;
        qb += qwb;
    }
    // inertial target x + v dt (+ g dt^2)
    xb += nlb;
    vb += nlb * dt;
    DV wdtb{0, 0, 0};
    dqint_adj(q, w * dt, qintb, qb, wdtb);
    wb += wdtb * dt;
    (void)qbz;
    st_dv(A.xb + 3 * i, xb);
    st_dq(A.qb + 4 * i, qb);
    st_dv(A.vb + 3 * i, vb);
    st_dv(A.wb + 3 * i, wb);
    st_dv(A.pvb + 3 * i, pvb);
}

// Force init of one contact reversed (manifold.cpp:100-174): C0 = basis (xA - xB) + margin,
// lambda = lambda_old alpha gamma, k = clamp(k_old gamma); anchors kept by the stick rule hand
// their adjoint to the previous step's contact. Pairs with a round shape refresh their anchors
// (manifold.cpp:14-27): a matched contact keeps only its normal lambda and penalty (the
// tangential rows restart at 0 and the minimum penalty), and its offsets are the shapes'.
template <bool Round>
__global__ void k_adj_init(Bodies B, Manifolds M, const int4 *contactRec, int nc, const int *srcSlot,
                           const int *matchOld, const int *stickUsed, const float4 *oldCr, int numOld, Params P,
                           AdjBuf A)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc)
        return;
    int4 r = contactRec[t];
    int m = r.x, i = r.y, sl = m * kMaxContacts + i, a = r.z, b = r.w;
    const float4 *cr = M.cr + (size_t)sl * kCR;
    DQ qA = dq4(B.pose[4 * a + P_IANG]), qB = dq4(B.pose[4 * b + P_IANG]);
    DV rA = dv4(cr[CR_RA]), rB = dv4(cr[CR_RB]);
    DV pA = dv4(B.pose[4 * a + P_ILIN]), pB = dv4(B.pose[4 * b + P_ILIN]);
    const int code = __float_as_int(cr[CR_RB].w);
    DV d;
    if (!Round || code == 0)
        d = (pA + drot(qA, rA)) - (pB + drot(qB, rB));
    else
    {
        V3 oA, oB;
        contact_anchors(cr[CR_RA], cr[CR_RB], cr[CR_C0], cr[CR_N], cr[CR_T1], cr[CR_T2], ldq(B.pose[4 * a + P_IANG]),
                        ldq(B.pose[4 * b + P_IANG]), oA, oB);
        d = (pA + dv(oA)) - (pB + dv(oB));
    }
    DV c0b = ld_dv(A.c0b + 3 * sl);
    DV db{0, 0, 0};
    DV n[3] = {dv4(cr[CR_N]), dv4(cr[CR_T1]), dv4(cr[CR_T2])};
    for (int k = 0; k < 3; ++k)
    {
        double ck = dcomp(c0b, k);
        db += n[k] * ck;
        atomic_add3(A.basisb + 9 * m + 3 * k, d * ck);
    }
    DQ qAb{0, 0, 0, 0}, qBb{0, 0, 0, 0};
    DV rAb = ld_dv(A.rAb + 3 * sl), rBb = ld_dv(A.rBb + 3 * sl), n0b{0, 0, 0};
    const DV n0 = dv4(cr[CR_N]);
    if (!Round || code == 0)
    {
        drot_adj(qA, rA, db, qAb, rAb);
        drot_adj(qB, rB, -db, qBb, rBb);
    }
    else
    {
        offset_adj(code, true, cr[CR_N].w, cr[CR_T2].w, qA, rA, n0, db, qAb, rAb, n0b);
        offset_adj(code, false, cr[CR_T1].w, cr[CR_C0].w, qB, rB, n0, -db, qBb, rBb, n0b);
        atomic_add3(A.basisb + 9 * m, n0b);
    }
    bool massA = B.massMoment[a].x > 0.0f, massB = B.massMoment[b].x > 0.0f;
    if (massA)
    {
        atomic_add3(A.xb + 3 * a, db);
        atomic_add4(A.qb + 4 * a, qAb);
    }
    if (massB)
    {
        atomic_add3(A.xb + 3 * b, -db);
        atomic_add4(A.qb + 4 * b, qBb);
    }
    int src = srcSlot[m];
    int j = matchOld[src * kMaxContacts + i];
    bool kept = stickUsed[src * kMaxContacts + i] != 0;
    const bool refresh = Round && (any_round((int)B.shape[a].z) || any_round((int)B.shape[b].z));
    if (j >= 0 && src < numOld)
    {
        int os = src * kMaxContacts + j;
        double ag = (double)P.alpha * (double)P.gamma;
        DV lb = ld_dv(A.lb + 3 * sl), kb = ld_dv(A.kb + 3 * sl);
        float4 kold = oldCr[(size_t)os * kCR + CR_PEN];
        const float ko[3] = {kold.x, refresh ? kPenaltyMin : kold.y, refresh ? kPenaltyMin : kold.z};
        double kbo[3] = {kb.x, kb.y, kb.z};
        for (int k = 0; k < 3; ++k)
        {
            float kg = ko[k] * P.gamma;
            if (!(kg > kPenaltyMin && kg < kPenaltyMax) || (refresh && k > 0))
                kbo[k] = 0.0;
            else
                kbo[k] *= (double)P.gamma;
        }
        if (refresh)
            lb = DV{lb.x, 0.0, 0.0};
        st_dv(A.clb + 3 * os, lb * ag);
        st_dv(A.ckb + 3 * os, DV{kbo[0], kbo[1], kbo[2]});
        if (kept)
        {
            st_dv(A.crAb + 3 * os, rAb);
            st_dv(A.crBb + 3 * os, rBb);
            rAb = rBb = DV{0, 0, 0};
        }
    }
    // fresh anchors: their adjoint goes through the collide Jacobian
    st_dv(A.rAb + 3 * sl, kept && j >= 0 && src < numOld ? DV{0, 0, 0} : rAb);
    st_dv(A.rBb + 3 * sl, kept && j >= 0 && src < numOld ? DV{0, 0, 0} : rBb);
}

// Collide pose Jacobian (collide.cpp:435-1181), one thread per (manifold, component): the tangent
// of the basis and fresh anchors along that component, dotted with their adjoints. Components 0-6
// and 7-13 are A's and B's pose (x, q) at the step start; 14-16 and 17-19 their velocities, which
// a sphere's swept test and the push-out face rule read (pairs of boxes skip them). The thread of
// component 0 also reverses the manifold friction sqrt(mu_A mu_B) (manifold.cpp:104; 0 for pairs
// with a sphere).
template <bool Round> constexpr int collide_comps() { return Round ? 20 : 14; }
template <bool Round>
__global__ __launch_bounds__(64) void k_adj_collide(Bodies B, Manifolds M, int nm, float dt, AdjBuf A)
{
    constexpr int kComps = collide_comps<Round>();
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nm * kComps)
        return;
    int m = t / kComps, comp_ = t % kComps;
    int2 ab = M.ab[m];
    int a = ab.x, b = ab.y;
    const float4 hA = B.shape[a], hB = B.shape[b];
    const int shA = (int)hA.z, shB = (int)hB.z;
    const bool boxes = !Round || (shA == kShapeBox && shB == kShapeBox);
    if (comp_ >= 14 && boxes)
        return;
    if (comp_ == 0 && (!Round || (shA != kShapeSphere && shB != kShapeSphere)))
    {
        const double fa = B.friction[a], fb = B.friction[b];
        const double mub = A.mumb[m], sq = sqrt(fa * fb);
        if (sq > 0.0)
        {
            aadd(A.mub + a, mub * fb / (2.0 * sq));
            aadd(A.mub + b, mub * fa / (2.0 * sq));
        }
    }
    // c: 0-2 position, 3-6 rotation, 7-9 velocity of the seeded body
    const bool onA = comp_ < 7 || (comp_ >= 14 && comp_ < 17);
    if (!(B.massMoment[onA ? a : b].x > 0.0f))
        return;
    const int c = comp_ < 14 ? (onA ? comp_ : comp_ - 7) : 7 + (comp_ - 14) % 3;
    float4 pA = B.pose[4 * a + P_ILIN], qA = B.pose[4 * a + P_IANG], pB = B.pose[4 * b + P_ILIN],
           qB = B.pose[4 * b + P_IANG];
    float4 sA = B.sizeRad[a], sB = B.sizeRad[b];
    auto seedv = [&](float4 v, bool on, int base) {
        return V3T<DualF>{DualF(v.x, on && c == base + 0 ? 1.0 : 0.0), DualF(v.y, on && c == base + 1 ? 1.0 : 0.0),
                          DualF(v.z, on && c == base + 2 ? 1.0 : 0.0)};
    };
    auto seedq = [&](float4 q, bool on) {
        return Q4T<DualF>{DualF(q.x, on && c == 3 ? 1.0 : 0.0), DualF(q.y, on && c == 4 ? 1.0 : 0.0),
                          DualF(q.z, on && c == 5 ? 1.0 : 0.0), DualF(q.w, on && c == 6 ? 1.0 : 0.0)};
    };
    V3T<DualF> PA = seedv(pA, onA, 0), PB = seedv(pB, !onA, 0);
    Q4T<DualF> QA = seedq(qA, onA), QB = seedq(qB, !onA);
    V3T<DualF> SA{DualF(sA.x), DualF(sA.y), DualF(sA.z)}, SB{DualF(sB.x), DualF(sB.y), DualF(sB.z)};
    NewContactsT<DualF> nc;
    V3T<DualF> r0, r1, r2;
    int nn;
    if (boxes)
        nn = collide_boxes<DualF>(PA, QA, SA, PB, QB, SB, nc, r0, r1, r2);
    else if (Round)
    {
        const ShapeBodyT<DualF> SBA{PA, SA, seedv(B.vel[a], onA, 7), QA, B.massMoment[a].x, hA.x, hA.y, shA};
        const ShapeBodyT<DualF> SBB{PB, SB, seedv(B.vel[b], !onA, 7), QB, B.massMoment[b].x, hB.x, hB.y, shB};
        nn = collide_pair<DualF>(SBA, SBB, dt, nc, r0, r1, r2);
    }
    int n = M.count[m];
    if (nn != n)
        return;  // a branch the step did not take cannot happen: values are the step's FP32 values
    const double *bb = A.basisb + 9 * m;
    double g = bb[0] * r0.x.d + bb[1] * r0.y.d + bb[2] * r0.z.d + bb[3] * r1.x.d + bb[4] * r1.y.d + bb[5] * r1.z.d +
               bb[6] * r2.x.d + bb[7] * r2.y.d + bb[8] * r2.z.d;
    for (int i = 0; i < n; ++i)
    {
        const double *ra = A.rAb + 3 * (m * kMaxContacts + i), *rb = A.rBb + 3 * (m * kMaxContacts + i);
        g += ra[0] * nc.rA[i].x.d + ra[1] * nc.rA[i].y.d + ra[2] * nc.rA[i].z.d;
        g += rb[0] * nc.rB[i].x.d + rb[1] * nc.rB[i].y.d + rb[2] * nc.rB[i].z.d;
    }
    int body = onA ? a : b;
    if (c < 3)
        aadd(A.xb + 3 * body + c, g);
    else if (c < 7)
        aadd(A.qb + 4 * body + (c - 3), g);
    else
        aadd(A.vb + 3 * body + (c - 7), g);
}

} // namespace avbd

// ------------------------------------------------------------------------------------------
// host side: checkpoints, taped recompute, the reverse sweep
// ------------------------------------------------------------------------------------------
namespace
{

using namespace avbd;

// A step-start state in HBM: body pose records and velocities, the persistent manifolds and the
// counters (the next serial among them).
struct Snapshot
{
    float4 *pose = nullptr, *vel = nullptr, *angVel = nullptr, *prevVel = nullptr, *cr = nullptr;
    int2 *ab = nullptr;
    long long *serial = nullptr;
    int *count = nullptr, *feature = nullptr;
    Counters *ctr = nullptr;
    Joint *joints = nullptr;  // the joint records (penalties, lambdas, C0) at the step start
    int numManifolds = 0;
};

} // namespace

// One step of a checkpoint segment, as its reversal reads it: the tape (poses and penalty/lambda
// after every iteration, each body's system and dx, branch flags), the narrowphase/compaction
// recording, and copies of the handle's arrays as the step left them before finishStep (bodies,
// manifolds, the flat contact list, stamps, colour-sorted tasks, colours, degrees) with the
// per-step counters and colour offsets the host read. Filled while the segment is replayed, so
// each step is computed once in the backward pass.
struct StepStore
{
    float4 *tapePose = nullptr, *tapeContact = nullptr, *tapeSys = nullptr;
    unsigned char *tapeStampFlags = nullptr, *tapeDualFlags = nullptr;
    Joint *tapeJoint = nullptr;  // joint records after the init (slot 0) and after each iteration
    int *recMatch = nullptr, *recStick = nullptr, *recSrc = nullptr;
    float4 *pose = nullptr, *vel = nullptr, *angVel = nullptr, *prevVel = nullptr, *inerLin = nullptr, *inerAng = nullptr;
    int2 *ab = nullptr;
    long long *serial = nullptr;
    int *count = nullptr, *feature = nullptr;
    float4 *cr = nullptr;
    int4 *contactRec = nullptr, *stamps = nullptr, *tasks = nullptr;
    int *color = nullptr, *deg = nullptr;
    Counters ctr{};
    std::vector<int> offsets;
};

struct AdjointState
{
    int steps = 0, every = 1;
    bool tapeAll = false;  // the forward taped every step (checkpoint_every = 0)
    bool recorded = false;
    unsigned epoch = 0;  // the handle's upload count when the forward was recorded
    std::vector<Snapshot> ck, seg;
    Snapshot end;
    AdjBuf A{};
    int slotCap = 0;
    int tapeIters = 0;            // iterations the step stores' tapes hold
    std::vector<StepStore> store; // one per step of a checkpoint segment
};

namespace
{

int snap_alloc(avbd_handle *h, Snapshot &s)
{
    int N = h->maxBodies, M = h->maxManifolds, r;
    size_t C = (size_t)M * kMaxContacts;
    if ((r = h->alloc(s.pose, 4 * (size_t)N)) || (r = h->alloc(s.vel, N)) || (r = h->alloc(s.angVel, N)) ||
        (r = h->alloc(s.prevVel, N)) || (r = h->alloc(s.ab, M)) || (r = h->alloc(s.serial, M)) ||
        (r = h->alloc(s.count, M)) || (r = h->alloc(s.feature, C)) || (r = h->alloc(s.cr, C * kCR)) ||
        (r = h->alloc(s.ctr, 1)) || (h->maxJoints > 0 && (r = h->alloc(s.joints, h->maxJoints))))
        return r;
    return AVBD_OK;
}

// Device-to-device copies gathered into one k_copy_many launch (sizes are whole 4-byte words).
struct CopyBatch
{
    CopyList L{};
    CopyBatch() { L.start[0] = 0; }
    void add(void *dst, const void *src, size_t bytes)
    {
        if (!bytes)
            return;
        L.src[L.n] = (const unsigned *)src;
        L.dst[L.n] = (unsigned *)dst;
        L.start[L.n + 1] = L.start[L.n] + (long long)(bytes / 4);
        ++L.n;
    }
    int launch(hipStream_t s, const char *what)
    {
        const long long total = L.start[L.n];
        if (total > 0)
            k_copy_many<<<(int)std::min<long long>((total + 255) / 256, 4096), 256, 0, s>>>(L);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return fail(AVBD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
        return AVBD_OK;
    }
};

int snap_copy(avbd_handle *h, bool save, Snapshot &s)
{
    const size_t N = h->nb;
    if (save)
        s.numManifolds = h->lastNumManifolds;
    const size_t M = s.numManifolds, C = M * kMaxContacts;
    CopyBatch b;
    auto cp = [&](void *snap, void *live, size_t bytes) { save ? b.add(snap, live, bytes) : b.add(live, snap, bytes); };
    cp(s.pose, h->B.pose, 4 * N * sizeof(float4));
    cp(s.vel, h->B.vel, N * sizeof(float4));
    cp(s.angVel, h->B.angVel, N * sizeof(float4));
    cp(s.prevVel, h->B.prevVel, N * sizeof(float4));
    cp(s.ab, h->cur.ab, M * sizeof(int2));
    cp(s.serial, h->cur.serial, M * sizeof(long long));
    cp(s.count, h->cur.count, M * sizeof(int));
    cp(s.feature, h->cur.feature, C * sizeof(int));
    cp(s.cr, h->cur.cr, C * kCR * sizeof(float4));
    cp(s.ctr, h->ctr, sizeof(Counters));
    cp(s.joints, h->joints, (size_t)h->nj * sizeof(Joint));
    int r = b.launch(h->stream, "adjoint snapshot copy");
    if (r)
        return r;
    if (!save)
        h->lastNumManifolds = s.numManifolds;
    return AVBD_OK;
}

// every = 0: a store and a start snapshot for every step (the forward tapes them all)
int adj_alloc(avbd_handle *h, int steps, int every)
{
    if (!h->adj)
        h->adj = new AdjointState();
    AdjointState &S = *h->adj;
    int r;
    const bool all = every == 0;
    if (all)
        every = steps;
    int nck = all ? 1 : steps / every + 1;
    while ((int)S.ck.size() < nck)
    {
        S.ck.emplace_back();
        if ((r = snap_alloc(h, S.ck.back())))
            return r;
    }
    while ((int)S.seg.size() < every)
    {
        S.seg.emplace_back();
        if ((r = snap_alloc(h, S.seg.back())))
            return r;
    }
    if (!S.end.pose && (r = snap_alloc(h, S.end)))
        return r;
    const int N = h->maxBodies, M = h->maxManifolds, I = std::max(h->P.iterations, 1);
    S.slotCap = M * kMaxContacts;
    const size_t C = (size_t)S.slotCap;
    // step stores, sized for the current iteration count (more iterations later: a larger set)
    if (I > S.tapeIters)
    {
        S.store.clear();
        S.tapeIters = I;
    }
    while ((int)S.store.size() < every)
    {
        S.store.emplace_back();
        StepStore &t = S.store.back();
        t.offsets.assign(257, 0);
        if ((r = h->alloc(t.tapePose, (size_t)(I + 1) * 2 * N)) || (r = h->alloc(t.tapeSys, (size_t)I * 9 * N)) ||
            (r = h->alloc(t.tapeStampFlags, (size_t)I * h->maxStamps)) ||
            (r = h->alloc(t.tapeDualFlags, (size_t)I * C)) || (r = h->alloc(t.tapeContact, (size_t)(I + 1) * 2 * C)) ||
            (r = h->alloc(t.recMatch, C)) || (r = h->alloc(t.recStick, C)) || (r = h->alloc(t.recSrc, M)) ||
            (r = h->alloc(t.pose, 4 * (size_t)N)) || (r = h->alloc(t.vel, N)) || (r = h->alloc(t.angVel, N)) ||
            (r = h->alloc(t.prevVel, N)) || (r = h->alloc(t.inerLin, N)) || (r = h->alloc(t.inerAng, N)) ||
            (r = h->alloc(t.ab, M)) || (r = h->alloc(t.serial, M)) || (r = h->alloc(t.count, M)) ||
            (r = h->alloc(t.feature, C)) || (r = h->alloc(t.cr, C * kCR)) || (r = h->alloc(t.contactRec, C)) ||
            (r = h->alloc(t.stamps, h->maxStamps)) || (r = h->alloc(t.tasks, N)) || (r = h->alloc(t.color, N)) ||
            (r = h->alloc(t.deg, N)) ||
            (h->maxJoints > 0 && (r = h->alloc(t.tapeJoint, (size_t)(I + 1) * h->maxJoints))))
        {
            S.store.pop_back();
            return r;
        }
    }
    if (!S.A.xb)
    {
        AdjBuf &A = S.A;
        if ((r = h->alloc(A.xb, 3 * (size_t)N)) ||
            (r = h->alloc(A.qb, 4 * (size_t)N)) || (r = h->alloc(A.vb, 3 * (size_t)N)) ||
            (r = h->alloc(A.wb, 3 * (size_t)N)) || (r = h->alloc(A.pvb, 3 * (size_t)N)) ||
            (r = h->alloc(A.ilb, 3 * (size_t)N)) || (r = h->alloc(A.iab, 4 * (size_t)N)) ||
            (r = h->alloc(A.nlb, 3 * (size_t)N)) || (r = h->alloc(A.nab, 4 * (size_t)N)) ||
            (r = h->alloc(A.mb, N)) || (r = h->alloc(A.mub, N)) || (r = h->alloc(A.kb, 3 * C)) ||
            (r = h->alloc(A.lb, 3 * C)) || (r = h->alloc(A.c0b, 3 * C)) || (r = h->alloc(A.rAb, 3 * C)) ||
            (r = h->alloc(A.rBb, 3 * C)) || (r = h->alloc(A.basisb, 9 * (size_t)M)) || (r = h->alloc(A.mumb, M)) ||
            (r = h->alloc(A.ckb, 3 * C)) || (r = h->alloc(A.clb, 3 * C)) || (r = h->alloc(A.crAb, 3 * C)) ||
            (r = h->alloc(A.crBb, 3 * C)) ||
            (h->maxJoints > 0 && ((r = h->alloc(A.jb, (size_t)kJointAdj * h->maxJoints)) ||
                                  (r = h->alloc(A.cjb, (size_t)kJointAdj * h->maxJoints)) ||
                                  (r = h->alloc(A.jc0b, (size_t)kJointC0 * h->maxJoints)))))
        {
            A.xb = nullptr;
            return r;
        }
    }
    return AVBD_OK;
}

// One step computed with its iterations taped into the store t, which then receives copies of the
// arrays its reversal reads as the step left them before finishStep (finishStep overwrites the
// start velocities the reversal needs); the caller runs finishStep afterwards to go on.
int taped_step(avbd_handle *h, StepStore &t)
{
    h->tapePose = t.tapePose;
    h->tapeContact = t.tapeContact;
    h->tapeSys = t.tapeSys;
    h->tapeStampFlags = t.tapeStampFlags;
    h->tapeDualFlags = t.tapeDualFlags;
    h->tapeJoint = t.tapeJoint;
    h->recMatch = t.recMatch;
    h->recStick = t.recStick;
    h->recSrc = t.recSrc;
    int r = enqueue_prepare(h);
    if (!r)
        r = enqueue_iterate(h);
    h->tapePose = h->tapeContact = h->tapeSys = nullptr;
    h->tapeStampFlags = h->tapeDualFlags = nullptr;
    h->tapeJoint = nullptr;
    h->recMatch = h->recStick = h->recSrc = nullptr;
    if (r)
        return r;
    // enqueue_iterate read the counters and colour offsets back before its launches
    t.ctr = *h->hostCtr;
    std::memcpy(t.offsets.data(), h->hostOffsets, sizeof(int) * 257);
    const size_t N = h->nb, M = t.ctr.numManifolds, C = M * kMaxContacts;
    CopyBatch b;
    b.add(t.pose, h->B.pose, 4 * N * sizeof(float4));
    b.add(t.vel, h->B.vel, N * sizeof(float4));
    b.add(t.angVel, h->B.angVel, N * sizeof(float4));
    b.add(t.prevVel, h->B.prevVel, N * sizeof(float4));
    b.add(t.inerLin, h->B.inerLin, N * sizeof(float4));
    b.add(t.inerAng, h->B.inerAng, N * sizeof(float4));
    b.add(t.ab, h->cur.ab, M * sizeof(int2));
    b.add(t.serial, h->cur.serial, M * sizeof(long long));
    b.add(t.count, h->cur.count, M * sizeof(int));
    b.add(t.feature, h->cur.feature, C * sizeof(int));
    b.add(t.cr, h->cur.cr, C * kCR * sizeof(float4));
    b.add(t.contactRec, h->S.contactRec, (size_t)t.ctr.contactCount * sizeof(int4));
    b.add(t.stamps, h->S.stamps, (size_t)std::min(t.ctr.numStamps, h->maxStamps) * sizeof(int4));
    b.add(t.tasks, h->S.tasks, N * sizeof(int4));
    b.add(t.color, h->S.color, N * sizeof(int));
    b.add(t.deg, h->S.deg, N * sizeof(int));
    return b.launch(h->stream, "adjoint step store");
}

// The reverse sweep of one step from its store t; start: the step's start state (the previous
// step's final manifolds, which the force init's warm start read).
int reverse_step(avbd_handle *h, const StepStore &t, const Snapshot &start)
{
    hipStream_t s = h->stream;
    AdjointState &S = *h->adj;
    AdjBuf &A = S.A;
    const Counters &c = t.ctr;
    const int nb = h->nb, nc = c.contactCount, nm = c.numManifolds, I = h->P.iterations;
    const int colors = std::min(c.colorCount, kMaxColors);
    const size_t ps = 2 * (size_t)h->maxBodies, cs = 2 * (size_t)S.slotCap;
    const float4 *tp = t.tapePose, *tc = t.tapeContact;
    // the handle's views with the step's arrays in place of the live ones
    Bodies B = h->B;
    B.pose = t.pose;
    B.vel = t.vel;
    B.angVel = t.angVel;
    B.prevVel = t.prevVel;
    B.inerLin = t.inerLin;
    B.inerAng = t.inerAng;
    Manifolds M{t.ab, t.serial, t.count, t.feature, t.cr};
    int n0 = std::max(std::max(nb, nc), std::max(nm, 1));
    const bool rolling = h->numSpheres > 0, round = h->numRound > 0;
    k_adj_begin<<<ceil_div(n0, 256), 256, 0, s>>>(B, h->P, t.contactRec, nc, nm, nb, !rolling, A);
    if (rolling)
    {
        const int numTasks = colors > 0 ? t.offsets[colors] : 0;
        if (numTasks > 0)
            k_adj_rolling<<<ceil_div(numTasks, 64), 64, 0, s>>>(B, M, t.stamps, t.tasks, numTasks, h->P, A);
        k_adj_finish<<<ceil_div(nb, 256), 256, 0, s>>>(B, h->P, nb, A);
    }
    const int nj = h->nj;
    if (nj > 0)
        k_adj_joint_begin<<<ceil_div(nj, 256), 256, 0, s>>>(nj, A);
    for (int it = I - 1; it >= 0; --it)
    {
        const float4 *prev = tp + it * ps, *cur = tp + (it + 1) * ps, *lam = tc + it * cs;
        const float4 *sys = t.tapeSys + (size_t)it * 9 * h->maxBodies;
        const unsigned char *sfl = t.tapeStampFlags + (size_t)it * h->maxStamps;
        const unsigned char *dfl = t.tapeDualFlags + (size_t)it * S.slotCap;
        if (nc > 0)
        {
            if (round)
                k_adj_dual<true><<<ceil_div(nc, 64), 64, 0, s>>>(B, M, t.contactRec, nc, cur, lam, dfl, h->P, A);
            else
                k_adj_dual<false><<<ceil_div(nc, 64), 64, 0, s>>>(B, M, t.contactRec, nc, cur, lam, dfl, h->P, A);
        }
        const Joint *jt = nj > 0 ? t.tapeJoint + (size_t)it * h->maxJoints : nullptr;
        if (nj > 0)
            k_adj_joint_dual<<<ceil_div(nj, 64), 64, 0, s>>>(jt, jt + h->maxJoints, nj, cur, B, h->P, A);
        for (int col = colors - 1; col >= 0; --col)
        {
            int b = t.offsets[col], e = t.offsets[col + 1];
            if (e > b)
            {
                if (round)
                    k_adj_primal<true><<<ceil_div(e - b, kAdjBodies), kAdjThreads, 0, s>>>(
                        B, M, h->springs, jt, t.stamps, t.tasks, b, e, t.color, col, prev, cur, lam, sys, sfl, h->P, A);
                else
                    k_adj_primal<false><<<ceil_div(e - b, kAdjBodies), kAdjThreads, 0, s>>>(
                        B, M, h->springs, jt, t.stamps, t.tasks, b, e, t.color, col, prev, cur, lam, sys, sfl, h->P, A);
            }
        }
    }
    k_adj_warmstart<<<ceil_div(nb, 256), 256, 0, s>>>(B, h->P, t.deg, h->numIgnore > 0 ? h->ignoreDeg : nullptr,
                                                      nb, A);
    if (nj > 0)
        k_adj_joint_init<<<ceil_div(nj, 256), 256, 0, s>>>(start.joints, nj, B, h->P, A);
    if (nc > 0)
    {
        if (round)
            k_adj_init<true><<<ceil_div(nc, 256), 256, 0, s>>>(B, M, t.contactRec, nc, t.recSrc, t.recMatch, t.recStick,
                                                     start.cr, start.numManifolds, h->P, A);
        else
            k_adj_init<false><<<ceil_div(nc, 256), 256, 0, s>>>(B, M, t.contactRec, nc, t.recSrc, t.recMatch, t.recStick,
                                                     start.cr, start.numManifolds, h->P, A);
    }
    if (nm > 0)
    {
        if (round)
            k_adj_collide<true><<<ceil_div(nm * collide_comps<true>(), 64), 64, 0, s>>>(B, M, nm, h->P.dt, A);
        else
            k_adj_collide<false><<<ceil_div(nm * collide_comps<false>(), 64), 64, 0, s>>>(B, M, nm, h->P.dt, A);
    }
    HIP_TRY(hipGetLastError());
    return AVBD_OK;
}

int check_adjoint_scene(avbd_handle *h)
{
    if (h->extPairsOn || !h->extContacts.empty())
        return fail(AVBD_ERR_UNSUPPORTED, "adjoint: worlds without external pairs or contacts");
    if (h->P.iterations > 64)
        return fail(AVBD_ERR_UNSUPPORTED, "adjoint: at most 64 iterations per step");
    return AVBD_OK;
}

// Per body {dynamic, moment / mass} of the resident state (a dynamic body's moments scale with
// its mass when avbd_set_body_inputs changes the mass).
__global__ void k_mass_ratio(Bodies B, int n, double *ratio)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float4 mm = B.massMoment[i];
    const bool dyn = mm.x > 0.0f;
    const double m = dyn ? (double)mm.x : 1.0;
    ratio[4 * i + 0] = dyn ? 1.0 : 0.0;
    ratio[4 * i + 1] = dyn ? (double)mm.y / m : 0.0;
    ratio[4 * i + 2] = dyn ? (double)mm.z / m : 0.0;
    ratio[4 * i + 3] = dyn ? (double)mm.w / m : 0.0;
}

// FP64 inputs (staged: vel [3n], angular velocity [3n], mass [n], friction [n]) rounded into the
// resident bodies; `which` bit k set: input k given. A body uploaded static keeps mass 0.
__global__ void k_set_inputs(Bodies B, int n, const double *in, const double *ratio, int which)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *v = in, *w = in + 3 * (size_t)n, *m = in + 6 * (size_t)n, *f = in + 7 * (size_t)n;
    if (which & 1)
    {
        float4 x = B.vel[i];
        B.vel[i] = make_float4((float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2], x.w);
    }
    if (which & 2)
    {
        float4 x = B.angVel[i];
        B.angVel[i] = make_float4((float)w[3 * i], (float)w[3 * i + 1], (float)w[3 * i + 2], x.w);
    }
    if (which & 4)
    {
        const double mi = ratio[4 * i] > 0.0 ? m[i] : 0.0;
        B.massMoment[i] = make_float4((float)mi, (float)(ratio[4 * i + 1] * mi), (float)(ratio[4 * i + 2] * mi),
                                      (float)(ratio[4 * i + 3] * mi));
    }
    if (which & 8)
        B.friction[i] = (float)f[i];
}

// Resident pose, rotation, velocity and angular velocity as FP64: out = pos [3n], rot [4n],
// vel [3n], angular velocity [3n].
__global__ void k_get_state_f64(Bodies B, int n, double *out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float4 p = B.pose[4 * i + P_POS], q = B.pose[4 * i + P_ROT], v = B.vel[i], w = B.angVel[i];
    double *pos = out, *rot = out + 3 * (size_t)n, *vel = out + 7 * (size_t)n, *ang = out + 10 * (size_t)n;
    pos[3 * i] = p.x, pos[3 * i + 1] = p.y, pos[3 * i + 2] = p.z;
    rot[4 * i] = q.x, rot[4 * i + 1] = q.y, rot[4 * i + 2] = q.z, rot[4 * i + 3] = q.w;
    vel[3 * i] = v.x, vel[3 * i + 1] = v.y, vel[3 * i + 2] = v.z;
    ang[3 * i] = w.x, ang[3 * i + 1] = w.y, ang[3 * i + 2] = w.z;
}

int io_alloc(avbd_handle *h)
{
    int r;
    if (!h->io64 && ((r = h->alloc(h->io64, 13 * (size_t)std::max(h->maxBodies, 1))) ||
                     (r = h->alloc(h->massRatio, 4 * (size_t)std::max(h->maxBodies, 1)))))
        return r;
    return AVBD_OK;
}

} // namespace

extern "C" {

int avbd_set_body_inputs(avbd_handle *h, const double *vel, const double *ang_vel, const double *mass,
                         const double *friction)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_set_body_inputs: null handle");
    int r;
    if ((r = pick_device(h)) || (r = io_alloc(h)))
        return r;
    const int n = h->nb;
    if (n == 0)
        return AVBD_OK;
    hipStream_t s = h->stream;
    if (!h->ratioValid)
    {
        k_mass_ratio<<<ceil_div(n, 256), 256, 0, s>>>(h->B, n, h->massRatio);
        HIP_TRY(hipGetLastError());
        h->ratioValid = true;
    }
    const double *src[4] = {vel, ang_vel, mass, friction};
    const size_t off[4] = {0, 3 * (size_t)n, 6 * (size_t)n, 7 * (size_t)n}, len[4] = {3, 3, 1, 1};
    int which = 0;
    for (int k = 0; k < 4; ++k)
        if (src[k])
        {
            HIP_TRY(hipMemcpyAsync(h->io64 + off[k], src[k], len[k] * n * sizeof(double), hipMemcpyDefault, s));
            which |= 1 << k;
        }
    if (which)
    {
        k_set_inputs<<<ceil_div(n, 256), 256, 0, s>>>(h->B, n, h->io64, h->massRatio, which);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(s));
    return AVBD_OK;
}

int avbd_get_body_state_f64(avbd_handle *h, double *pos, double *rot, double *vel, double *ang_vel)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_get_body_state_f64: null handle");
    int r;
    if ((r = pick_device(h)) || (r = io_alloc(h)))
        return r;
    const int n = h->nb;
    if (n == 0)
        return AVBD_OK;
    hipStream_t s = h->stream;
    k_get_state_f64<<<ceil_div(n, 256), 256, 0, s>>>(h->B, n, h->io64);
    HIP_TRY(hipGetLastError());
    double *dst[4] = {pos, rot, vel, ang_vel};
    const size_t off[4] = {0, 3 * (size_t)n, 7 * (size_t)n, 10 * (size_t)n}, len[4] = {3, 4, 3, 3};
    for (int k = 0; k < 4; ++k)
        if (dst[k])
            HIP_TRY(hipMemcpyAsync(dst[k], h->io64 + off[k], len[k] * n * sizeof(double), hipMemcpyDefault, s));
    HIP_TRY(hipStreamSynchronize(s));
    return AVBD_OK;
}

int avbd_adjoint_rewind(avbd_handle *h)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_adjoint_rewind: null handle");
    if (!h->adj || !h->adj->recorded || h->adj->epoch != h->uploadEpoch)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_adjoint_rewind: no recorded forward of the uploaded state");
    int r;
    if ((r = pick_device(h)) || (r = snap_copy(h, false, h->adj->ck[0])) || (r = avbd_synchronize(h)))
        return r;
    return AVBD_OK;
}

int avbd_adjoint_forward(avbd_handle *h, int32_t num_steps, int32_t checkpoint_every)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_adjoint_forward: null handle");
    if (num_steps < 1 || checkpoint_every < 0)
        return fail(AVBD_ERR_INVALID_ARGUMENT,
                    "avbd_adjoint_forward: steps must be >= 1 and the checkpoint interval >= 0");
    int r;
    if ((r = pick_device(h)) || (r = check_adjoint_scene(h)) || (r = adj_alloc(h, num_steps, checkpoint_every)))
        return r;
    AdjointState &S = *h->adj;
    S.recorded = false;
    S.epoch = h->uploadEpoch;
    S.steps = num_steps;
    S.tapeAll = checkpoint_every == 0;
    S.every = S.tapeAll ? num_steps : checkpoint_every;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if ((r = snap_copy(h, true, S.ck[0])))
        return r;
    for (int k = 0; k < num_steps; ++k)
    {
        if (S.tapeAll)
        {
            // every step taped as it is taken: its start state and store, then finishStep
            if ((r = snap_copy(h, true, S.seg[k])) || (r = taped_step(h, S.store[k])) || (r = enqueue_finish(h)))
                return r;
            continue;
        }
        if ((r = enqueue_prepare(h)) || (r = enqueue_iterate(h)) || (r = enqueue_finish(h)))
            return r;
        if ((k + 1) % checkpoint_every == 0 && (k + 1) < num_steps && (r = snap_copy(h, true, S.ck[(k + 1) / checkpoint_every])))
            return r;
    }
    if ((r = snap_copy(h, true, S.end)) || (r = avbd_synchronize(h)))
        return r;
    S.recorded = true;
    return AVBD_OK;
}

int avbd_adjoint_backward(avbd_handle *h, const double *d_pos, const double *d_rot, const double *d_vel,
                          const double *d_ang_vel, double *g_pos, double *g_rot, double *g_vel, double *g_ang_vel,
                          double *g_mass, double *g_friction)
{
    if (!h)
        return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_adjoint_backward: null handle");
    if (!h->adj || !h->adj->recorded || h->adj->epoch != h->uploadEpoch)
        return fail(AVBD_ERR_INVALID_ARGUMENT,
                    "avbd_adjoint_backward: no recorded forward (avbd_adjoint_forward) of the uploaded state");
    int r;
    if ((r = pick_device(h)) || (r = check_adjoint_scene(h)))
        return r;
    AdjointState &S = *h->adj;
    AdjBuf &A = S.A;
    hipStream_t s = h->stream;
    const size_t N = h->nb;
    auto zero = [&](double *p, size_t n) { return hipMemsetAsync(p, 0, n * sizeof(double), s); };
    auto up = [&](double *dst, const double *src, size_t n) {
        return src ? hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDefault, s) : zero(dst, n);
    };
    const size_t C3 = 3 * (size_t)S.slotCap;
    hipError_t e;
    if ((e = up(A.xb, d_pos, 3 * N)) || (e = up(A.qb, d_rot, 4 * N)) || (e = up(A.vb, d_vel, 3 * N)) ||
        (e = up(A.wb, d_ang_vel, 3 * N)) || (e = zero(A.pvb, 3 * N)) || (e = zero(A.mb, N)) || (e = zero(A.mub, N)) ||
        (e = zero(A.ckb, C3)) || (e = zero(A.clb, C3)) || (e = zero(A.crAb, C3)) || (e = zero(A.crBb, C3)) ||
        (h->nj > 0 && (e = zero(A.cjb, (size_t)kJointAdj * h->nj))))
        return fail(AVBD_ERR_HIP, std::string("adjoint init: ") + hipGetErrorString(e));
    const int K = S.every, steps = S.steps;
    const int nseg = (steps + K - 1) / K;
    if (S.tapeAll)
    {
        if (h->P.iterations > S.tapeIters)
            return fail(AVBD_ERR_INVALID_ARGUMENT, "avbd_adjoint_backward: iterations changed since the taped forward");
        for (int st = steps - 1; st >= 0; --st)
            if ((r = reverse_step(h, S.store[st], S.seg[st])))
                return r;
    }
    else if (h->P.iterations > S.tapeIters && (r = adj_alloc(h, steps, K)))
        return r;
    for (int sg = S.tapeAll ? -1 : nseg - 1; sg >= 0; --sg)
    {
        const int s0 = sg * K, s1 = std::min(s0 + K, steps);
        // replay the segment from its checkpoint once, taping every step into its store and
        // keeping its start state (the previous step's manifolds) for the force init's reversal
        if ((r = snap_copy(h, false, S.ck[sg])))
            return r;
        for (int st = s0; st < s1; ++st)
        {
            if ((r = snap_copy(h, true, S.seg[st - s0])) || (r = taped_step(h, S.store[st - s0])) ||
                (st + 1 < s1 && (r = enqueue_finish(h))))
                return r;
        }
        for (int st = s1 - 1; st >= s0; --st)
            if ((r = reverse_step(h, S.store[st - s0], S.seg[st - s0])))
                return r;
    }
    // leave the handle at the forward's final state
    if ((r = snap_copy(h, false, S.end)))
        return r;
    auto down = [&](double *dst, const double *src, size_t n) {
        return dst ? hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDefault, s) : hipSuccess;
    };
    if ((e = down(g_pos, A.xb, 3 * N)) || (e = down(g_rot, A.qb, 4 * N)) || (e = down(g_vel, A.vb, 3 * N)) ||
        (e = down(g_ang_vel, A.wb, 3 * N)) || (e = down(g_mass, A.mb, N)) || (e = down(g_friction, A.mub, N)) ||
        (e = hipStreamSynchronize(s)))
        return fail(AVBD_ERR_HIP, std::string("adjoint results: ") + hipGetErrorString(e));
    return avbd_synchronize(h);
}

} // extern "C"
