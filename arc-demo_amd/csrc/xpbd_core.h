/*
 * xpbd_core.h -- FP32 XPBD primitives shared by the device kernels and the
 * host-side builder of libarcsim.
 *
 * Every routine evaluates its arithmetic in the order of the reference engine
 * (csrc/arc_phys.c, csrc/arc_math.h), one IEEE-754 binary32 rounding per
 * operation: the library is compiled with -ffp-contract=off, division and
 * square root are correctly rounded, and atan2 / sin / cos are evaluated in
 * binary64 and rounded once, which is what the reference computes when its
 * `double` is float.  The result tracks the FP32 build of the reference op for
 * op; only the last-bit rounding of binary64 atan2/sin/cos may differ between
 * the host libm and the device's.
 *
 * Bodies come in three kinds, chosen at compile time per call site:
 *   KISO  isotropic inertia (torso, coxas): world inverse inertia is the
 *         scalar invI.x, so products with it are three multiplies;
 *   KANI  anisotropic (femurs, tibias): full symmetric 3x3;
 *   KGEN  either, decided at run time from the body constants (the generic
 *         single-world path).
 * The reference multiplies the zero entries of an isotropic body's diagonal
 * matrix too; dropping those `0 * v` terms changes no finite non-zero result
 * (at most the sign of an exact zero).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

#define XD __host__ __device__ __forceinline__
/* Branch hints for conditions that hold in a small fraction of calls (degenerate
 * lengths, joint limits, touch-downs, lookups outside the LDS tile).  They only
 * move the rarely run code out of the hot instruction stream, so that the common
 * path runs without taken branches; the arithmetic is unchanged. */
#define ARC_LIKELY(c) __builtin_expect(!!(c), 1)
#define ARC_UNLIKELY(c) __builtin_expect(!!(c), 0)

namespace xpbd {

enum { KGEN = 0, KISO = 1, KANI = 2 };

struct V3 { float x, y, z; };
struct Q4 { float w, x, y, z; };
struct M3 { float e[3][3]; };

XD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
XD V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
XD V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
XD V3 scl(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
XD V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
XD V3 had(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
XD V3 mad(V3 a, V3 b, float s) { return v3(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }
XD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
XD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
XD float len2(V3 a) { return dot(a, a); }
XD float len(V3 a) { return sqrtf(dot(a, a)); }
XD V3 zero3() { return v3(0.0f, 0.0f, 0.0f); }

/* 1 / x correctly rounded.  On the device, for 2^-126 <= |x| < 2^126, one
 * Newton step on the hardware reciprocal: r + r (1 - x r) with fused
 * multiply-adds equals the IEEE quotient 1.0f / x for every such float (both
 * signs, all 2^32 bit patterns checked exhaustively on gfx950), in 3
 * instructions instead of the 10 of the general division; every other x
 * (zero, denormal, huge, infinite, NaN) takes the division. */
XD float recipf(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    if (ARC_LIKELY(fabsf(x) >= 0x1p-126f && fabsf(x) < 0x1p126f)) {
        const float r = __builtin_amdgcn_rcpf(x);
        return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
    }
#endif
    return 1.0f / x;
}

/* a / d for a divisor fixed per simulation (the column size).  On the device,
 * with y = RN(1 / d): q0 = a y, r = d q0 - a (one fused multiply-add),
 * q = q0 - r y, three instructions instead of the ten of the general
 * division.  `fast` is set only for a divisor for which this sequence has been
 * checked, on the device, to give the IEEE quotient a / d bit for bit for
 * every binary32 numerator that is zero or has 2^-96 <= |a| < 2^100 (all 2^32
 * patterns, selfcheck.hip); the caller guarantees its numerators lie there.
 * Unchecked divisors take the division. */
struct Divisor {
    float d, y;
    int fast;
};
XD bool qdiv_domain(float a) { return a == 0.0f || (fabsf(a) >= 0x1p-96f && fabsf(a) < 0x1p100f); }
/* the three-instruction sequence itself (host builds evaluate it for tests) */
XD float qdiv_fma(float a, float d, float y)
{
    const float q0 = a * y;
    const float r = __builtin_fmaf(d, q0, -a);
    return __builtin_fmaf(-r, y, q0);
}
XD float qdiv(float a, const Divisor &D)
{
#ifdef __HIP_DEVICE_COMPILE__
    if (D.fast)
        return qdiv_fma(a, D.d, D.y);
#endif
    return a / D.d;
}

/* binary64 evaluation rounded to binary32 */
XD float f_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
XD float f_sin(float x) { return (float)sin((double)x); }
XD float f_cos(float x) { return (float)cos((double)x); }

XD V3 unit(V3 a)
{
    float l = len(a);
    return ARC_LIKELY(l > 1e-15f) ? scl(a, recipf(l)) : zero3();
}
XD V3 perp(V3 n)
{
    V3 t = fabsf(n.x) < 0.57f ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 1.0f, 0.0f);
    return unit(cross(n, t));
}

XD Q4 q4(float w, float x, float y, float z) { Q4 r; r.w = w; r.x = x; r.y = y; r.z = z; return r; }
XD Q4 qid() { return q4(1.0f, 0.0f, 0.0f, 0.0f); }
XD Q4 conj(Q4 a) { return q4(a.w, -a.x, -a.y, -a.z); }
XD Q4 qmul(Q4 a, Q4 b)
{
    return q4(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
              a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
              a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w);
}
XD Q4 qunit(Q4 a)
{
    float l = sqrtf(a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z);
    if (ARC_UNLIKELY(l < 1e-15f))
        return qid();
    l = recipf(l);
    return q4(a.w * l, a.x * l, a.y * l, a.z * l);
}
XD Q4 qaxis(V3 axis, float angle)
{
    float s = f_sin(0.5f * angle);
    return q4(f_cos(0.5f * angle), axis.x * s, axis.y * s, axis.z * s);
}
XD V3 qrot(Q4 q, V3 v)
{
    V3 u = v3(q.x, q.y, q.z);
    V3 t = scl(cross(u, v), 2.0f);
    return add(mad(v, t, q.w), cross(u, t));
}
XD V3 qirot(Q4 q, V3 v) { return qrot(conj(q), v); }

/* small rotation without renormalisation, arc_phys.c:54-61 */
XD Q4 qturn(Q4 q, V3 w)
{
    float hx = 0.5f * w.x, hy = 0.5f * w.y, hz = 0.5f * w.z;
    return q4(q.w - hx * q.x - hy * q.y - hz * q.z,
              q.x + hx * q.w + hy * q.z - hz * q.y,
              q.y - hx * q.z + hy * q.w + hz * q.x,
              q.z + hx * q.y - hy * q.x + hz * q.w);
}

XD void m_from_q(M3 &r, Q4 q)
{
    /* 2 / s = 2 (1 / s) exactly: a power-of-two scale commutes with rounding */
    float k = 2.0f * recipf(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    float x2 = q.x * k, y2 = q.y * k, z2 = q.z * k;
    float xx = q.x * x2, yy = q.y * y2, zz = q.z * z2;
    float xy = q.x * y2, xz = q.x * z2, yz = q.y * z2;
    float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    r.e[0][0] = 1.0f - (yy + zz); r.e[0][1] = xy - wz; r.e[0][2] = xz + wy;
    r.e[1][0] = xy + wz; r.e[1][1] = 1.0f - (xx + zz); r.e[1][2] = yz - wx;
    r.e[2][0] = xz - wy; r.e[2][1] = yz + wx; r.e[2][2] = 1.0f - (xx + yy);
}
XD V3 mmul(const M3 &a, V3 v)
{
    return v3(a.e[0][0] * v.x + a.e[0][1] * v.y + a.e[0][2] * v.z,
              a.e[1][0] * v.x + a.e[1][1] * v.y + a.e[1][2] * v.z,
              a.e[2][0] * v.x + a.e[2][1] * v.y + a.e[2][2] * v.z);
}
XD V3 mtmul(const M3 &a, V3 v)
{
    return v3(a.e[0][0] * v.x + a.e[1][0] * v.y + a.e[2][0] * v.z,
              a.e[0][1] * v.x + a.e[1][1] * v.y + a.e[2][1] * v.z,
              a.e[0][2] * v.x + a.e[1][2] * v.y + a.e[2][2] * v.z);
}
XD float clampf_(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
XD float minf_(float a, float b) { return a < b ? a : b; }
XD float maxf_(float a, float b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* bodies                                                              */

/* per-body constants (identical in every world) */
struct BodyK {
    float mass, invm;
    V3 I, invI;
    int iso_I;     /* inertia isotropic: no gyroscopic term (arc_phys.c:667) */
    int iso_invI;  /* inverse inertia isotropic: scalar world inverse inertia (arc_phys.c:70) */
};

/* per-body dynamic state of one world */
struct Body {
    V3 x;
    Q4 q;
    V3 v, w;
    M3 R;   /* rotation of q / |q| */
    M3 Iw;  /* world inverse inertia (KANI / KGEN only) */
};

/* a body's pose at the start of the current substep (x_prev, q_prev of the
 * reference), written by integrate and read by the velocity update and by
 * friction touch-downs; kept outside Body so callers can store it apart */
struct Pose0 {
    V3 x0;
    Q4 q0;
};

template <int KIND> XD bool iso_inv(const BodyK &k) { return KIND == KISO ? true : KIND == KANI ? false : k.iso_invI != 0; }
template <int KIND> XD bool iso_in(const BodyK &k) { return KIND == KISO ? true : KIND == KANI ? false : k.iso_I != 0; }

/* world inverse inertia times v */
template <int KIND> XD V3 iw_mul(const Body &b, const BodyK &k, V3 v)
{
    if (KIND == KISO)
        return scl(v, k.invI.x);
    return mmul(b.Iw, v);
}

/* refresh R and Iw from q, arc_phys.c:64-82 */
template <int KIND> XD void bsync(Body &b, const BodyK &k)
{
    m_from_q(b.R, b.q);
    if (KIND == KISO)
        return;
    V3 d = k.invI;
    if (iso_inv<KIND>(k)) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                b.Iw.e[i][j] = 0.0f;
        b.Iw.e[0][0] = b.Iw.e[1][1] = b.Iw.e[2][2] = d.x;
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        V3 rd = v3(b.R.e[i][0] * d.x, b.R.e[i][1] * d.y, b.R.e[i][2] * d.z);
#pragma unroll
        for (int j = i; j < 3; ++j)
            b.Iw.e[i][j] = b.Iw.e[j][i] = rd.x * b.R.e[j][0] + rd.y * b.R.e[j][1] + rd.z * b.R.e[j][2];
    }
}

XD V3 brot(const Body &b, V3 v) { return mmul(b.R, v); }

template <int KIND> XD float wpos(const Body &b, const BodyK &k, V3 r, V3 n)
{
    V3 rn = cross(r, n);
    return k.invm + dot(rn, iw_mul<KIND>(b, k, rn));
}
template <int KIND> XD float wrot(const Body &b, const BodyK &k, V3 n) { return dot(n, iw_mul<KIND>(b, k, n)); }

template <int KIND> XD void push_pos(Body &b, const BodyK &k, V3 r, V3 p)
{
    b.x = mad(b.x, p, k.invm);
    b.q = qturn(b.q, iw_mul<KIND>(b, k, cross(r, p)));
    bsync<KIND>(b, k);
}
template <int KIND> XD void push_rot(Body &b, const BodyK &k, V3 p)
{
    b.q = qturn(b.q, iw_mul<KIND>(b, k, p));
    bsync<KIND>(b, k);
}
template <int KIND> XD void push_vel(Body &b, const BodyK &k, V3 r, V3 p)
{
    b.v = mad(b.v, p, k.invm);
    b.w = add(b.w, iw_mul<KIND>(b, k, cross(r, p)));
}

/* positional constraint against the static world with zero compliance
 * (solve_pos(b, r, NULL, r, d, 0, h), arc_phys.c:226-242): the static side
 * contributes exact zeros to the denominator */
template <int KIND> XD float pos_static(Body &b, const BodyK &k, V3 r, V3 d)
{
    float c = len(d);
    if (ARC_UNLIKELY(c < 1e-12f))
        return 0.0f;
    V3 n = scl(d, recipf(c));
    float den = wpos<KIND>(b, k, r, n);
    if (ARC_UNLIKELY(den <= 0.0f))
        return 0.0f;
    float dl = c / den;
    push_pos<KIND>(b, k, r, scl(n, dl));
    return dl;
}

/* symmetric 3x3 solve by adjugate, arc_phys.c:265-280.  The singularity
 * threshold 1e-300 is zero in binary32, so the test never fires. */
struct S3 { float a, b, c, e, f, g; };
XD bool sym3(const S3 &K, V3 d, V3 &x)
{
    float a = K.a, b = K.b, c = K.c, e = K.e, f = K.f, g = K.g;
    float A = e * g - f * f, B = c * f - b * g, C = b * f - c * e;
    float det = a * A + b * B + c * C;
    if (fabsf(det) < 0.0f)
        return false;
    float inv = recipf(det);
    float D = a * g - c * c;
    float E = b * c - a * f;
    float F = a * e - b * b;
    x.x = (A * d.x + B * d.y + C * d.z) * inv;
    x.y = (B * d.x + D * d.y + E * d.z) * inv;
    x.z = (C * d.x + E * d.y + F * d.z) * inv;
    return true;
}

/* K += point-mass matrix of the body at offset r, arc_phys.c:284-300 */
template <int KIND> XD void point_mass(S3 &K, const Body &b, const BodyK &k, V3 r)
{
    float x = r.x, y = r.y, z = r.z;
    if (KIND == KISO) {
        float d = k.invI.x; /* off-diagonal entries of the inverse inertia are zero */
        K.a += k.invm + d * z * z + d * y * y;
        K.e += k.invm + d * x * x + d * z * z;
        K.g += k.invm + d * y * y + d * x * x;
        K.b += -(d * x) * y;
        K.c += -(d * x) * z;
        K.f += -(d * y) * z;
        return;
    }
    const M3 &a = b.Iw;
    K.a += k.invm + a.e[1][1] * z * z - 2.0f * a.e[1][2] * y * z + a.e[2][2] * y * y;
    K.e += k.invm + a.e[2][2] * x * x - 2.0f * a.e[0][2] * x * z + a.e[0][0] * z * z;
    K.g += k.invm + a.e[0][0] * y * y - 2.0f * a.e[0][1] * x * y + a.e[1][1] * x * x;
    float k01 = a.e[1][2] * x * z + a.e[0][2] * y * z - a.e[0][1] * z * z - a.e[2][2] * x * y;
    float k02 = a.e[0][1] * y * z + a.e[1][2] * x * y - a.e[1][1] * x * z - a.e[0][2] * y * y;
    float k12 = a.e[0][2] * x * y + a.e[0][1] * x * z - a.e[0][0] * y * z - a.e[1][2] * x * x;
    K.b += k01;
    K.c += k02;
    K.f += k12;
}

template <int KIND> XD void iw_accum(S3 &K, const Body &b, const BodyK &k)
{
    if (KIND == KISO) {
        K.a += k.invI.x;
        K.e += k.invI.x;
        K.g += k.invI.x;
        return;
    }
    K.a += b.Iw.e[0][0]; K.b += b.Iw.e[0][1]; K.c += b.Iw.e[0][2];
    K.e += b.Iw.e[1][1]; K.f += b.Iw.e[1][2]; K.g += b.Iw.e[2][2];
}

/* ------------------------------------------------------------------ */
/* joints (hinge with limits and a torque-limited PD drive)             */

struct JointK {
    V3 r1, r2, a1, a2, b1, b2;
    Q4 qrel0;
    float lo, hi;
    float lim[4];     /* cos lo, sin lo, cos hi, sin hi (binary64 libm, rounded) */
    float kp, kd, tmax;
    float soft;       /* (1 / kp) / (h * h) */
    float cap_pos;    /* tmax * h * h */
    float tmax_h;     /* tmax * h */
    int wide;         /* hi - lo > pi */
    int has_limits;   /* lo <= hi */
};


/* rigid 3x3 angular block between P (or the static world) and C,
 * arc_phys.c:328-346 */
template <int PK, int CK>
XD void rot_block(Body *P, const BodyK *pk, Body &C, const BodyK &ck, V3 rot)
{
    if (ARC_UNLIKELY(len2(rot) < 1e-24f))
        return;
    S3 K = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    if (PK >= 0)
        iw_accum<PK < 0 ? KGEN : PK>(K, *P, *pk);
    iw_accum<CK>(K, C, ck);
    V3 p;
    if (!sym3(K, rot, p))
        return;
    if (PK >= 0)
        push_rot<PK < 0 ? KGEN : PK>(*P, *pk, p);
    push_rot<CK>(C, ck, neg(p));
}

/* rigid point-to-point block, arc_phys.c:308-320 */
template <int PK, int CK>
XD void point_block(Body *P, const BodyK *pk, V3 r1, Body &C, const BodyK &ck, V3 r2, V3 d)
{
    if (ARC_UNLIKELY(len2(d) < 1e-24f))
        return;
    S3 K = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    if (PK >= 0)
        point_mass<PK < 0 ? KGEN : PK>(K, *P, *pk, r1);
    point_mass<CK>(K, C, ck, r2);
    V3 p;
    if (!sym3(K, d, p))
        return;
    if (PK >= 0)
        push_pos<PK < 0 ? KGEN : PK>(*P, *pk, r1, p);
    push_pos<CK>(C, ck, r2, neg(p));
}

/* Position pass of one hinge, arc_phys.c:413-457.  PK = -1 attaches the
 * joint to the static world (P and pk are then unused). */
template <int PK, int CK>
XD void joint_positions(Body *P, const BodyK *pk, Body &C, const BodyK &ck, const JointK &j,
                        float target, float hh, float &tacc)
{
    constexpr int PKK = PK < 0 ? KGEN : PK;
    const bool HP = PK >= 0;
    if (j.kp > 0.0f) {
        /* PD drive: compliant, torque-limited rotation of the parent by
         * (phi - target) relative to the child (solve_rot, arc_phys.c:353-371) */
        V3 a1 = HP ? brot(*P, j.a1) : j.a1;
        V3 e1 = HP ? brot(*P, j.b1) : j.b1;
        V3 e2 = brot(C, j.b2);
        float s = dot(cross(e1, e2), a1), c = dot(e1, e2);
        float phi = f_atan2(s, c);
        V3 rot = scl(a1, phi - target);
        V3 p = zero3();
        float th = len(rot);
        if (ARC_LIKELY(!(th < 1e-12f))) {
            V3 n = scl(rot, recipf(th));
            float den = (HP ? wrot<PKK>(*P, *pk, n) : 0.0f) + wrot<CK>(C, ck, n);
            den = den + j.soft;
            if (ARC_LIKELY(!(den <= 0.0f))) {
                float dl = th / den;
                if (dl > j.cap_pos)
                    dl = j.cap_pos;
                p = scl(n, dl);
                if (HP)
                    push_rot<PKK>(*P, *pk, p);
                push_rot<CK>(C, ck, neg(p));
            }
        }
        tacc -= dot(p, a1) / hh;
    }
    if (j.has_limits) { /* limits, arc_phys.c:434-445 */
        V3 a1 = HP ? brot(*P, j.a1) : j.a1;
        V3 e1 = HP ? brot(*P, j.b1) : j.b1;
        V3 e2 = brot(C, j.b2);
        float s = dot(cross(e1, e2), a1), c = dot(e1, e2);
        if (ARC_UNLIKELY(j.wide || s * j.lim[0] - c * j.lim[1] < 0.0f || j.lim[3] * c - j.lim[2] * s < 0.0f)) {
            float phi = f_atan2(s, c);
            if (phi < j.lo)
                rot_block<PK, CK>(P, pk, C, ck, scl(a1, phi - j.lo));
            else if (phi > j.hi)
                rot_block<PK, CK>(P, pk, C, ck, scl(a1, phi - j.hi));
        }
    }
    { /* hinge axis alignment, arc_phys.c:448-449 */
        V3 a1 = HP ? brot(*P, j.a1) : j.a1;
        rot_block<PK, CK>(P, pk, C, ck, cross(a1, brot(C, j.a2)));
    }
    { /* attachment, arc_phys.c:452-456 */
        V3 r1 = HP ? brot(*P, j.r1) : j.r1;
        V3 r2 = brot(C, j.r2);
        V3 p1 = HP ? add(P->x, r1) : r1;
        V3 p2 = add(C.x, r2);
        point_block<PK, CK>(P, pk, r1, C, ck, r2, sub(p2, p1));
    }
}

/* Velocity pass of one hinge: PD damping, arc_phys.c:459-481 */
template <int PK, int CK>
XD void joint_velocities(Body *P, const BodyK *pk, Body &C, const BodyK &ck, const JointK &j,
                         float h, float &tacc)
{
    constexpr int PKK = PK < 0 ? KGEN : PK;
    const bool HP = PK >= 0;
    if (j.kd <= 0.0f)
        return;
    V3 a = HP ? brot(*P, j.a1) : j.a1;
    V3 wp = HP ? P->w : zero3();
    float rel = dot(sub(C.w, wp), a);
    float ws = (HP ? wrot<PKK>(*P, *pk, a) : 0.0f) + wrot<CK>(C, ck, a);
    if (ws <= 0.0f)
        return;
    float imp = -j.kd * rel * h;
    float cap = minf_(fabsf(rel) / ws, j.tmax_h);
    imp = clampf_(imp, -cap, cap);
    if (HP)
        P->w = sub(P->w, iw_mul<PKK>(*P, *pk, scl(a, imp)));
    C.w = add(C.w, iw_mul<CK>(C, ck, scl(a, imp)));
    tacc += imp / h;
}

/* joint outputs, arc_phys.c:386-400 (from q, not from the cached R) */
template <bool HP>
XD float joint_angle(const Body *P, const Body &C, const JointK &j)
{
    Q4 qp = HP ? P->q : qid();
    V3 a = qrot(qp, j.a1), e1 = qrot(qp, j.b1), e2 = qrot(C.q, j.b2);
    return f_atan2(dot(cross(e1, e2), a), dot(e1, e2));
}
template <bool HP>
XD float joint_rate(const Body *P, const Body &C, const JointK &j)
{
    V3 a = qrot(HP ? P->q : qid(), j.a1);
    V3 wp = HP ? P->w : zero3();
    return dot(sub(C.w, wp), a);
}

/* ------------------------------------------------------------------ */
/* spheres and the block terrain                                        */

/* Sphere offsets in the body frame: AT_GEN any, AT_ZERO the origin,
 * AT_Z on the body z axis (0, 0, z).  The special forms drop products with
 * exact zeros only. */
enum { AT_GEN = 0, AT_ZERO = 1, AT_Z = 2 };

template <int AT> XD V3 at_world(const Body &b, V3 at)
{
    if (AT == AT_ZERO)
        return b.x;
    if (AT == AT_Z)
        return add(b.x, v3(b.R.e[0][2] * at.z, b.R.e[1][2] * at.z, b.R.e[2][2] * at.z));
    return add(b.x, brot(b, at));
}
/* x0 + q0 (at), the sphere centre at the substep start */
template <int AT> XD V3 at_world0(const Pose0 &p, V3 at)
{
    if (AT == AT_ZERO)
        return p.x0;
    if (AT == AT_Z) {
        const Q4 q = p.q0;
        float z = at.z;
        float tx = (q.y * z) * 2.0f, ty = (-(q.x * z)) * 2.0f;
        return add(p.x0, v3(tx * q.w - q.z * ty, ty * q.w + q.z * tx, z + (q.x * ty - q.y * tx)));
    }
    return add(p.x0, qrot(p.q0, at));
}

/* friction anchor of one sphere held in plain variables */
struct AnchorRef {
    V3 &a, &n;
    XD V3 point() const { return a; }
    XD V3 normal() const { return n; }
    XD void set_point(V3 v) { a = v; }
    XD void set_normal(V3 v) { n = v; }
};

struct SphereK {
    V3 at;
    float rad;
    int body;
};

/* Per-world block terrain.  Every height the engine produces is a whole
 * number F of 3 cm steps shifted by the world's base (csrc/arc_terrain.c:62-70:
 * h = 0.03f * F, then h -= base), so a column is stored as F in one byte and
 * read back as 0.03f * F - base, bit for bit the same binary32 value.  The grid
 * carries a border of TPAD columns marked Q_OUT on every side; columns outside
 * the grid read as the world's `outside` height. */
constexpr int TPAD = 2;
constexpr int Q_OUT = -128;
/* largest terrain difficulty whose columns fit the int8 encoding: a column is
 * F 0.03 m steps with |F| <= 21 d + 1 (hills 0.45 d, bumps 0.12 d, jitter
 * 0.06 d, csrc/arc_terrain.c:54-71), so d <= 6 keeps |F| <= 127 */
constexpr double MAX_DIFFICULTY = 6.0;
struct TerrainRef {
    const int8_t *q;  /* column (0, 0); row pitch `pitch` */
    int pitch;        /* nx + 2 * TPAD */
    float base, outside;
};
struct TerrainK {
    int nx, ny;
    float cell, ox, oy;
    float inv_cell; /* RN(1 / cell) */
    int fast_cell;  /* column indices through qdiv (checked exact, numerators in its domain) */
};
XD float dequant(const TerrainRef &r, int k) { return k == Q_OUT ? r.outside : 0.03f * (float)k - r.base; }
XD float terrain_cell(const TerrainK &t, const TerrainRef &r, int ix, int iy)
{
    if (ix < 0 || iy < 0 || ix >= t.nx || iy >= t.ny)
        return r.outside;
    return dequant(r, r.q[ix + iy * r.pitch]);
}
/* height of the column containing (x, y), arc_terrain.h:42-50 */
XD float terrain_height(const TerrainK &t, const TerrainRef &r, float x, float y)
{
    float fx = floorf((x - t.ox) / t.cell), fy = floorf((y - t.oy) / t.cell);
    if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)t.nx && fy < (float)t.ny))
        return r.outside;
    return dequant(r, r.q[(int)fx + (int)fy * r.pitch]);
}

/* the columns a sphere's box overlaps.  Callers have checked |c.x|, |c.y| <
 * 1e6, and fast_cell is set only when every radius lies in [2^-60, 2^90) and
 * ox, oy are zero or lie in the same range: then each numerator (c -+ rad) - o
 * is zero or at least 2^-84 (a difference of floats of magnitude >= 2^-61 is a
 * multiple of 2^-84) and below 2^100, inside qdiv's checked domain */
XD void sphere_window(const TerrainK &t, V3 c, float rad, int &ix0, int &ix1, int &iy0, int &iy1)
{
    const Divisor D = { t.cell, t.inv_cell, t.fast_cell };
    ix0 = (int)floorf(qdiv(c.x - rad - t.ox, D));
    ix1 = (int)floorf(qdiv(c.x + rad - t.ox, D));
    iy0 = (int)floorf(qdiv(c.y - rad - t.oy, D));
    iy1 = (int)floorf(qdiv(c.y + rad - t.oy, D));
}

/* column heights straight from the world's grid */
struct GridCells {
    const TerrainK &t;
    const TerrainRef &r;
    XD float operator()(int ix, int iy) const { return terrain_cell(t, r, ix, iy); }
};

/* Column sources that can tell at once that no column of a sphere's window
 * reaches a height: clear(ix0, ix1, iy0, iy1, lo) is true only if no column
 * of the window has its top above lo (false when it cannot tell). */
template <class C, class = void> struct has_window_test : std::false_type {};
template <class C>
struct has_window_test<C, std::void_t<decltype(std::declval<const C &>().clear(0, 0, 0, 0, 0.0f))>> : std::true_type {};

/* anchored static / kinetic friction, arc_phys.c:512-539.  The anchor is
 * read and written through `an` (only when the sphere touches), the
 * substep-start pose fetched through pose0() only at a touch-down. */
template <int KIND, int AT, class Anchor, class PoseFn>
XD void friction(Body &b, const BodyK &k, V3 at, float rad, int &anchored, Anchor &an, const PoseFn &pose0,
                 V3 n, float ln, float mus, float mud)
{
    V3 c = at_world<AT>(b, at);
    V3 p = mad(c, n, -rad);
    float frac = 1.0f;
    V3 anc;
    if (ARC_UNLIKELY(!anchored || dot(n, an.normal()) < 0.9f)) {
        /* touch down: anchor where the contact point was at the substep start */
        V3 c0 = at_world0<AT>(pose0(), at);
        anc = mad(c0, n, -rad);
        an.set_point(anc);
        an.set_normal(n);
        anchored = 1;
    } else {
        anc = an.point();
    }
    V3 e = sub(p, anc);
    e = mad(e, n, -dot(e, n));
    float l = len(e);
    if (ARC_UNLIKELY(l < 1e-12f))
        return;
    V3 r = sub(p, b.x);
    float wt = wpos<KIND>(b, k, r, scl(e, recipf(l)));
    float lt = l / wt;
    if (lt > mus * ln) {
        frac = minf_(mud * ln / lt, 1.0f);
        an.set_point(mad(anc, e, 1.0f - frac));
    }
    pos_static<KIND>(b, k, r, scl(e, -frac));
}

/* One sphere against the column grid, each contact solved at once and
 * recorded through rec(local, n) for the velocity pass; arc_phys.c:542-610.
 * cell(ix, iy) gives a column's height (any source holding the same int8
 * grid).  Updates touch / anchored / anchor state. */
template <int KIND, int AT, class Rec, class Anchor, class PoseFn, class Cells>
XD void collide(Body &b, const BodyK &k, V3 at, float rad, int &touch, int &anchored, Anchor &an,
                const PoseFn &pose0, const TerrainK &t, const Cells &cell, Rec &rec, float mus, float mud)
{
    float ln = 0.0f, dmax = 0.0f;
    V3 c = at_world<AT>(b, at), nmax = v3(0.0f, 0.0f, 1.0f);
    int hit = 0;
    if (ARC_UNLIKELY(!(fabsf(c.x) < 1e6f && fabsf(c.y) < 1e6f))) {
        anchored = 0;
        return;
    }
    int ix0, ix1, iy0, iy1;
    sphere_window(t, c, rad, ix0, ix1, iy0, iy1);
    if constexpr (has_window_test<Cells>::value) {
        /* the loop below resolves a contact only with a column whose top is
         * above the sphere's lowest point (or when that point is NaN); with
         * none in the window it changes nothing but the anchor */
        if (cell.clear(ix0, ix1, iy0, iy1, c.z - rad)) {
            anchored = 0;
            return;
        }
    }
    for (int iy = iy0; iy <= iy1; ++iy) {
        for (int ix = ix0; ix <= ix1; ++ix) {
            float top = cell(ix, iy);
            if (c.z - rad >= top)
                continue;
            float x0 = t.ox + (float)ix * t.cell, x1 = x0 + t.cell;
            float y0 = t.oy + (float)iy * t.cell, y1 = y0 + t.cell;
            int sx = c.x > x1 ? 1 : (c.x < x0 ? -1 : 0);
            int sy = c.y > y1 ? 1 : (c.y < y0 ? -1 : 0);
            if (sx && cell(ix + sx, iy) >= top)
                continue;
            if (sy && cell(ix, iy + sy) >= top)
                continue;
            if (sx && sy && cell(ix + sx, iy + sy) >= top)
                continue;
            V3 q = v3(clampf_(c.x, x0, x1), clampf_(c.y, y0, y1), minf_(c.z, top));
            V3 diff = sub(c, q);
            float dist = len(diff), depth;
            V3 n;
            if (dist > 1e-9f) {
                if (dist >= rad)
                    continue;
                n = scl(diff, recipf(dist));
                depth = rad - dist;
            } else {
                n = v3(0.0f, 0.0f, 1.0f);
                depth = top - c.z + rad;
            }
            /* normal contact, arc_phys.c:488-500 */
            V3 r = sub(mad(c, n, -rad), b.x);
            float lni = pos_static<KIND>(b, k, r, scl(n, depth));
            rec(mtmul(b.R, r), n);
            ln += lni;
            c = at_world<AT>(b, at);
            if (depth > dmax) {
                dmax = depth;
                nmax = n;
            }
            hit = 1;
        }
    }
    if (hit)
        friction<KIND, AT>(b, k, at, rad, anchored, an, pose0, nmax, ln, mus, mud);
    else
        anchored = 0;
    touch += hit;
}

/* restitution-0 velocity correction of one recorded contact, arc_phys.c:612-626 */
template <int KIND>
XD void contact_velocity(Body &b, const BodyK &k, V3 local, V3 n)
{
    V3 r = brot(b, local);
    V3 vc = add(b.v, cross(b.w, r));
    float vn = dot(n, vc);
    push_vel<KIND>(b, k, r, scl(n, -vn / wpos<KIND>(b, k, r, n)));
}

/* ------------------------------------------------------------------ */
/* integration                                                          */

/* Implicit gyroscopic step in the body frame, arc_phys.c:636-656, with the
 * Jacobian columns written out (products with the unit vectors' zeros
 * dropped). */
XD V3 gyro(const BodyK &k, V3 w0, float h)
{
    V3 I = k.I;
    V3 iw = had(w0, I);
    V3 f = scl(cross(w0, iw), h);
    float J00 = I.x, J11 = I.y, J22 = I.z;
    float J10 = (w0.z * I.x - iw.z) * h, J20 = (iw.y - w0.y * I.x) * h;
    float J01 = (iw.z - w0.z * I.y) * h, J21 = (w0.x * I.y - iw.x) * h;
    float J02 = (w0.y * I.z - iw.y) * h, J12 = (iw.x - w0.x * I.z) * h;
    /* Cramer, arc_phys.c:245-262 (threshold 1e-300 is zero in binary32) */
    float c00 = J11 * J22 - J12 * J21;
    float c01 = J12 * J20 - J10 * J22;
    float c02 = J10 * J21 - J11 * J20;
    float det = J00 * c00 + J01 * c01 + J02 * c02;
    if (fabsf(det) < 0.0f)
        return w0;
    float inv = recipf(det);
    V3 k0 = v3(J00, J10, J20), k1 = v3(J01, J11, J21), k2 = v3(J02, J12, J22);
    V3 dw;
    dw.x = dot(f, cross(k1, k2)) * inv;
    dw.y = dot(k0, cross(f, k2)) * inv;
    dw.z = dot(k0, cross(k1, f)) * inv;
    return sub(w0, dw);
}

/* arc_phys.c:658-672; returns the substep-start pose */
template <int KIND>
XD Pose0 integrate(Body &b, const BodyK &k, V3 g, float h)
{
    Pose0 p0;
    p0.x0 = b.x;
    p0.q0 = b.q;
    b.v = mad(b.v, g, h);
    b.x = mad(b.x, b.v, h);
    if (!iso_in<KIND>(k))
        b.w = mmul(b.R, gyro(k, mtmul(b.R, b.w), h));
    b.q = qturn(b.q, scl(b.w, h));
    bsync<KIND>(b, k);
    return p0;
}

/* arc_phys.c:674-687; R and Iw are deliberately not refreshed here */
XD void velocities(Body &b, const Pose0 &p0, float inv_h, float two_h)
{
    b.q = qunit(b.q);
    Q4 dq = qmul(b.q, conj(p0.q0));
    b.v = scl(sub(b.x, p0.x0), inv_h);
    b.w = scl(v3(dq.x, dq.y, dq.z), two_h);
    if (dq.w < 0.0f)
        b.w = neg(b.w);
}

XD bool finite_(float v) { return v == v && v < INFINITY && v > -INFINITY; }

/* arc_phys.c:694-705 */
XD bool body_unstable(const Body &b)
{
    float v2 = len2(b.v), w2 = len2(b.w);
    return !finite_(v2) || !finite_(w2) || !(len2(b.x) < 1e6f * 1e6f) || v2 > 60.0f * 60.0f
           || w2 > 400.0f * 400.0f;
}

} /* namespace xpbd */
