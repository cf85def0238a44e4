/*
 * root_split.h -- the position pass of a root (hip) joint solved by two lanes.
 *
 * A root joint couples the torso (parent) with a coxa (child); both bodies
 * are isotropic.  joint_positions<KISO, KISO> (xpbd_core.h, after
 * csrc/arc_phys.c:413-457) applies every update to the two bodies in turn;
 * here the torso side runs in one lane and the coxa side in another, each on
 * its own body, with the same operations in the same order.  What one side
 * computes and the other needs (rotated axes and anchors, anchor positions,
 * each body's share of the generalised inverse mass) is passed between the
 * lanes by DPP; what both need (the drive angle, normals, the 3x3 solves, the
 * impulses) both compute from the same values, so both lanes hold the same
 * bits.  The bodies end bit for bit as the one-lane solve leaves them.
 */
#pragma once
#include "lane_engine.h"

namespace xpbd {

/* the value of lane CL (the coxa side) or PL (the torso side) of each quad,
 * in every lane of the quad: DPP quad_perm */
template <int CL, int PL> struct QuadX {
    static constexpr int perm(int l) { return l | (l << 2) | (l << 4) | (l << 6); }
    __device__ __forceinline__ float P(float v) const
    {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), perm(PL), 0xf, 0xf, false));
    }
    __device__ __forceinline__ float C(float v) const
    {
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), perm(CL), 0xf, 0xf, false));
    }
    __device__ __forceinline__ V3 P(V3 v) const { return v3(P(v.x), P(v.y), P(v.z)); }
    __device__ __forceinline__ V3 C(V3 v) const { return v3(C(v.x), C(v.y), C(v.z)); }
};

/* a or b, component by component (a conditional over whole structs would
 * take their addresses and move them to memory) */
__device__ __forceinline__ V3 sel3(bool c, V3 a, V3 b) { return v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }

/* push_rot<KISO> and push_pos<KISO> on the lane's own body */
__device__ __forceinline__ void push_rot_iso(Body &B, float invI, V3 p)
{
    B.q = qturn(B.q, scl(p, invI));
    m_from_q(B.R, B.q);
}
__device__ __forceinline__ void push_pos_iso(Body &B, float invm, float invI, V3 r, V3 p)
{
    B.x = mad(B.x, p, invm);
    B.q = qturn(B.q, scl(cross(r, p), invI));
    m_from_q(B.R, B.q);
}

/* rot_block<KISO, KISO>: the parent takes p, the child -p */
__device__ __forceinline__ void rot_block_pair(Body &B, bool child, float invI, V3 rot)
{
    if (ARC_UNLIKELY(len2(rot) < 1e-24f))
        return;
    constexpr float iP = robot::body[0].invI.x, iC = robot::body[1].invI.x;
    S3 K = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    K.a += iP; K.e += iP; K.g += iP; /* iw_accum<KISO>(parent) */
    K.a += iC; K.e += iC; K.g += iC; /* iw_accum<KISO>(child) */
    V3 p;
    if (!sym3(K, rot, p))
        return;
    push_rot_iso(B, invI, sel3(child, neg(p), p));
}

/* point_block<KISO, KISO> with the parent's anchor r1 and the child's r2 */
__device__ __forceinline__ void point_block_pair(Body &B, bool child, float invm, float invI, V3 r1, V3 r2, V3 d)
{
    if (ARC_UNLIKELY(len2(d) < 1e-24f))
        return;
    S3 K = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
    point_mass<KISO>(K, B, robot::body[0], r1); /* the isotropic form reads only r and the constants */
    point_mass<KISO>(K, B, robot::body[1], r2);
    V3 p;
    if (!sym3(K, d, p))
        return;
    push_pos_iso(B, invm, invI, sel3(child, r2, r1), sel3(child, neg(p), p));
}

/* joint_positions<KISO, KISO> of root joint JI by the two lanes that X
 * names: B is the torso in the parent lane, the coxa in the child lane (other
 * lanes compute on whatever B holds and their results are not used).  Returns
 * the drive's term dot(p, a1) / hh, the same in both lanes, which the joint's
 * torque sum loses. */
template <int JI, class X>
__device__ __forceinline__ float root_positions_pair(Body &B, bool child, float target, float hh, const Model &M)
{
    ROBOT_JOINT(JI, M);
    const X x;
    const float invm = child ? robot::body[1].invm : robot::body[0].invm;
    const float invI = child ? robot::body[1].invI.x : robot::body[0].invI.x;
    float dt = 0.0f;
    if (J.kp > 0.0f) { /* PD drive, csrc/arc_phys.c:353-371 */
        const V3 u = brot(B, sel3(child, J.b2, J.a1)), v = brot(B, J.b1);
        const V3 a1 = x.P(u), e1 = x.P(v), e2 = x.C(u);
        float s = dot(cross(e1, e2), a1), c = dot(e1, e2);
        float phi = f_atan2(s, c);
        V3 rot = scl(a1, phi - target);
        V3 p = zero3();
        float th = len(rot);
        if (ARC_LIKELY(!(th < 1e-12f))) {
            V3 n = scl(rot, recipf(th));
            const float w = dot(n, scl(n, invI)); /* wrot<KISO> of the lane's body */
            float den = x.P(w) + x.C(w);
            den = den + J.soft;
            if (ARC_LIKELY(!(den <= 0.0f))) {
                float dl = th / den;
                if (dl > J.cap_pos)
                    dl = J.cap_pos;
                p = scl(n, dl);
                push_rot_iso(B, invI, sel3(child, neg(p), p));
            }
        }
        dt = dot(p, a1) / hh;
    }
    if (J.has_limits) { /* limits, csrc/arc_phys.c:434-445 */
        const V3 u = brot(B, sel3(child, J.b2, J.a1)), v = brot(B, J.b1);
        const V3 a1 = x.P(u), e1 = x.P(v), e2 = x.C(u);
        float s = dot(cross(e1, e2), a1), c = dot(e1, e2);
        if (ARC_UNLIKELY(J.wide || s * J.lim[0] - c * J.lim[1] < 0.0f || J.lim[3] * c - J.lim[2] * s < 0.0f)) {
            float phi = f_atan2(s, c);
            if (phi < J.lo)
                rot_block_pair(B, child, invI, scl(a1, phi - J.lo));
            else if (phi > J.hi)
                rot_block_pair(B, child, invI, scl(a1, phi - J.hi));
        }
    }
    { /* hinge axis alignment, csrc/arc_phys.c:448-449 */
        const V3 u = brot(B, sel3(child, J.a2, J.a1));
        rot_block_pair(B, child, invI, cross(x.P(u), x.C(u)));
    }
    { /* attachment, csrc/arc_phys.c:452-456 */
        const V3 u = brot(B, sel3(child, J.r2, J.r1));
        const V3 pw = add(B.x, u);
        point_block_pair(B, child, invm, invI, x.P(u), x.C(u), sub(x.C(pw), x.P(pw)));
    }
    return dt;
}

} /* namespace xpbd */
