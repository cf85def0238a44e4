/*
 * leg_engine.h -- one control step of one quadruped world run by the four
 * lanes of a quad: lane q holds leg q (coxa, femur, tibia) and a replica of
 * the torso, which every lane keeps identical.
 *
 * The reference projects the joints in the order 0..11 (csrc/arc_phys.c:
 * 726-737, Gauss-Seidel).  Joints 3l (hip yaw) couple the torso with leg l;
 * joints 3l+1 and 3l+2 touch only leg l's own bodies.  So the order
 *     root 0, root 1, root 2, root 3, then every leg's joint 1, then joint 2
 * applies the same sequence of updates to every body (operations on disjoint
 * bodies commute exactly) and gives bit-identical results.  The four roots
 * run one after another, each in the lane of its leg, and the torso it
 * produces is broadcast to the quad (DPP) before the next; the mid and distal
 * joints of the four legs run at once.  Spheres, integration and the
 * velocity passes act on one body at a time: each lane does its leg's bodies
 * and the torso's (the same arithmetic in all four lanes).
 *
 * Per lane this is 4 bodies, 6 spheres and 6 joint solves per substep
 * against 13, 21 and 12 in the one-world-per-lane form (lane_engine.h):
 * a shorter dependent chain per world, for latency-bound batches.
 */
#pragma once
#include "lane_engine.h"

namespace xpbd {

/* The form relies on the legs differing only in joint geometry: every leg's
 * bodies, spheres and joint gains / limits equal leg 0's (checked here on
 * the generated constants). */
constexpr bool same_v3(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
constexpr bool legs_alike()
{
    for (int l = 1; l < LEGS; ++l) {
        for (int m = 0; m < 3; ++m) {
            const BodyK &a = robot::body[1 + m], &b = robot::body[1 + 3 * l + m];
            if (!(a.mass == b.mass && a.invm == b.invm && same_v3(a.I, b.I) && same_v3(a.invI, b.invI)
                  && a.iso_I == b.iso_I && a.iso_invI == b.iso_invI))
                return false;
            const JointK &p = robot::joint[m], &q = robot::joint[3 * l + m];
            if (!(p.lo == q.lo && p.hi == q.hi && p.kp == q.kp && p.kd == q.kd && p.tmax == q.tmax
                  && p.wide == q.wide && p.has_limits == q.has_limits && p.lim[0] == q.lim[0]
                  && p.lim[1] == q.lim[1] && p.lim[2] == q.lim[2] && p.lim[3] == q.lim[3]))
                return false;
        }
        for (int k = 1; k <= 5; ++k)
            if (!(same_v3(robot::sphere[k].at, robot::sphere[5 * l + k].at)
                  && robot::sphere[k].rad == robot::sphere[5 * l + k].rad))
                return false;
    }
    return true;
}
static_assert(legs_alike(), "the leg form needs legs that differ only in joint geometry");

/* value of lane T of each quad, in all four lanes (DPP quad_perm) */
template <int T> __device__ __forceinline__ float qbcast(float v)
{
    constexpr int ctrl = T | (T << 2) | (T << 4) | (T << 6);
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), ctrl, 0xf, 0xf, false));
}
__device__ __forceinline__ int qor(int v)
{
    v |= __builtin_amdgcn_update_dpp(v, v, 0xb1, 0xf, 0xf, false); /* quad_perm 1,0,3,2 */
    v |= __builtin_amdgcn_update_dpp(v, v, 0x4e, 0xf, 0xf, false); /* quad_perm 2,3,0,1 */
    return v;
}
template <int T> __device__ __forceinline__ void qbcast3(V3 &a) { a.x = qbcast<T>(a.x); a.y = qbcast<T>(a.y); a.z = qbcast<T>(a.z); }

/* one of four per-leg values by the lane's leg */
__device__ __forceinline__ float sel4(int q, float a, float b, float c, float d) { return q == 0 ? a : q == 1 ? b : q == 2 ? c : d; }
__device__ __forceinline__ V3 sel4(int q, V3 a, V3 b, V3 c, V3 d)
{
    return v3(sel4(q, a.x, b.x, c.x, d.x), sel4(q, a.y, b.y, c.y, d.y), sel4(q, a.z, b.z, c.z, d.z));
}
__device__ __forceinline__ Q4 sel4(int q, Q4 a, Q4 b, Q4 c, Q4 d)
{
    return q4(sel4(q, a.w, b.w, c.w, d.w), sel4(q, a.x, b.x, c.x, d.x), sel4(q, a.y, b.y, c.y, d.y),
              sel4(q, a.z, b.z, c.z, d.z));
}

/* constants of joint 3q + m for the lane's leg q (gains and limits are those
 * of every leg's joint m; the geometry differs per leg) */
template <int M_>
__device__ __forceinline__ JointK leg_joint(int q, const Model &M)
{
    constexpr JointK a = robot::joint[M_], b = robot::joint[3 + M_], c = robot::joint[6 + M_], d = robot::joint[9 + M_];
    JointK J = a;
    J.r1 = sel4(q, a.r1, b.r1, c.r1, d.r1);
    J.r2 = sel4(q, a.r2, b.r2, c.r2, d.r2);
    J.a1 = sel4(q, a.a1, b.a1, c.a1, d.a1);
    J.a2 = sel4(q, a.a2, b.a2, c.a2, d.a2);
    J.b1 = sel4(q, a.b1, b.b1, c.b1, d.b1);
    J.b2 = sel4(q, a.b2, b.b2, c.b2, d.b2);
    J.qrel0 = sel4(q, a.qrel0, b.qrel0, c.qrel0, d.qrel0);
    J.soft = M.joint[M_].soft;
    J.cap_pos = M.joint[M_].cap_pos;
    J.tmax_h = M.joint[M_].tmax_h;
    return J;
}

/* State of one world in one lane of its quad: the torso replica, the leg's
 * three bodies (0 coxa, 1 femur, 2 tibia) and its three joints; spheres in
 * lane-local order 0 torso, 1 femur, 2..5 tibia. */
struct LegWorld {
    Body T;
    Body b[3];
    float target[3], tacc[3], angle[3], rate[3], torque[3];
    uint32_t touchp[2]; /* byte k of word k / 4: substeps local sphere k touched */
    uint32_t anchored;  /* bit k: local sphere k has a valid friction anchor */
    int unstable;
    float time;
};
__device__ __forceinline__ int leg_touch(const LegWorld &W, int k) { return (int)((W.touchp[k >> 2] >> (8 * (k & 3))) & 0xffu); }
/* global sphere index of local sphere k of leg q */
__device__ __forceinline__ int leg_sphere(int q, int k) { return k == 0 ? 0 : 5 * q + k; }

/* contact slots of one lane: torso, femur, tibia ranges of sizes ccap[0],
 * ccap[1], ccap[5] (every leg's femur and tibia have the same spheres) */
__device__ __forceinline__ int leg_slot(const Model &M, int lb)
{
    return lb == 0 ? 0 : lb == 1 ? M.ccap[0] : M.ccap[0] + M.ccap[1];
}
__device__ __forceinline__ int leg_cap(const Model &M, int lb) { return lb == 0 ? M.ccap[0] : lb == 1 ? M.ccap[1] : M.ccap[5]; }
__device__ __forceinline__ int leg_slots(const Model &M) { return M.ccap[0] + M.ccap[1] + M.ccap[5]; }

/* the torso's pose / velocity after lane T's root joint, in all four lanes */
template <int T> __device__ __forceinline__ void bcast_pose(Body &B)
{
    qbcast3<T>(B.x);
    B.q.w = qbcast<T>(B.q.w); B.q.x = qbcast<T>(B.q.x); B.q.y = qbcast<T>(B.q.y); B.q.z = qbcast<T>(B.q.z);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            B.R.e[r][c] = qbcast<T>(B.R.e[r][c]);
}
template <int T> __device__ __forceinline__ void bcast_vel(Body &B)
{
    qbcast3<T>(B.v);
    qbcast3<T>(B.w);
}

/* One control step of the quad's world (lane q).  Anchors: global sphere
 * indices; Poses: local bodies 0 torso, 1 coxa, 2 femur, 3 tibia; the
 * contact store addresses this lane's own list.  Returns this lane's
 * contact-overflow bits (bit per global contact body, as the lane form). */
template <class Anchors, class Poses, class Tile>
__device__ __forceinline__ int leg_step(LegWorld &W, int q0, Anchors &A0, Poses &P0, const Model &M,
                                        const TerrainRef &tr0, const ContactStore &cs0, const Tile &T0)
{
    int overflow = 0;
    bsync<KISO>(W.T, robot::body[0]);
    bsync<KISO>(W.b[0], robot::body[1]);
    bsync<KANI>(W.b[1], robot::body[2]);
    bsync<KANI>(W.b[2], robot::body[3]);
    W.touchp[0] = W.touchp[1] = 0u;
    W.tacc[0] = W.tacc[1] = W.tacc[2] = 0.0f;

    for (int sub0 = 0; sub0 < M.substeps; ++sub0) {
        /* the leg index, opaque per substep: the per-leg joint constants are
         * selected where used instead of being held across the step */
        const int q = (int)launder((size_t)q0);
        auto &&A = A0.fresh();
        ContactStore cs = cs0;
        cs.world = launder(cs0.world);
        TerrainRef tr = tr0;
        tr.q = launder(tr0.q);
        P0.store(0, integrate<KISO>(W.T, robot::body[0], M.g, M.h));
        P0.store(1, integrate<KISO>(W.b[0], robot::body[1], M.g, M.h));
        P0.store(2, integrate<KANI>(W.b[1], robot::body[2], M.g, M.h));
        P0.store(3, integrate<KANI>(W.b[2], robot::body[3], M.g, M.h));

        /* joint positions: the four roots in turn, then mids, then distals */
        static_for<0, LEGS>([&](auto t) {
            if (q == t) {
                ROBOT_JOINT(3 * t, M);
                joint_positions<KISO, KISO>(&W.T, &robot::body[0], W.b[0], robot::body[1], J, W.target[0], M.hh,
                                            W.tacc[0]);
            }
            bcast_pose<t>(W.T);
        });
        {
            const JointK J = leg_joint<1>(q, M);
            joint_positions<KISO, KANI>(&W.b[0], &robot::body[1], W.b[1], robot::body[2], J, W.target[1], M.hh,
                                        W.tacc[1]);
        }
        {
            const JointK J = leg_joint<2>(q, M);
            joint_positions<KANI, KANI>(&W.b[1], &robot::body[2], W.b[2], robot::body[3], J, W.target[2], M.hh,
                                        W.tacc[2]);
        }

        /* spheres: the torso's (every lane), then the leg's five in order */
        int ncount[3] = { 0, 0, 0 };
        const auto cells = T0.cells(M.terr, tr);
        static_for<0, 6>([&](auto k) {
            constexpr int LB = k == 0 ? 0 : (k == 1 ? 1 : 2); /* 0 torso, 1 femur, 2 tibia (contact lists) */
            constexpr int GS = k == 0 ? 0 : (int)k;           /* the sphere's constants: leg 0's */
            int anch = (W.anchored >> k) & 1u;
            int hit = 0;
            auto rec = [&](V3 l, V3 n) {
                if (ncount[LB] < leg_cap(M, LB))
                    cs.put(leg_slot(M, LB) + ncount[LB], l, n);
                else
                    overflow |= 1 << (LB == 0 ? 0 : LB == 1 ? 1 + q0 : 5 + q0);
                ncount[LB]++;
            };
            auto an = A.at(leg_sphere(q, k));
            if constexpr (k == 0) {
                auto pose0 = [&]() { return P0.load(0); };
                collide<KISO, sphere_at(0)>(W.T, robot::body[0], robot::sphere[0].at, robot::sphere[0].rad, hit,
                                            anch, an, pose0, M.terr, cells, rec, M.mus, M.mud);
            } else {
                auto pose0 = [&]() { return P0.load(LB + 1); };
                collide<KANI, sphere_at(GS)>(W.b[LB], robot::body[LB + 1], robot::sphere[GS].at,
                                             robot::sphere[GS].rad, hit, anch, an, pose0, M.terr, cells, rec,
                                             M.mus, M.mud);
            }
            W.anchored = (W.anchored & ~(1u << k)) | ((uint32_t)anch << k);
            W.touchp[k >> 2] += (uint32_t)hit << (8 * (k & 3));
        });

        velocities(W.T, P0.load(0), M.inv_h, M.two_h);
        velocities(W.b[0], P0.load(1), M.inv_h, M.two_h);
        velocities(W.b[1], P0.load(2), M.inv_h, M.two_h);
        velocities(W.b[2], P0.load(3), M.inv_h, M.two_h);

        static_for<0, LEGS>([&](auto t) {
            if (q == t) {
                ROBOT_JOINT(3 * t, M);
                joint_velocities<KISO, KISO>(&W.T, &robot::body[0], W.b[0], robot::body[1], J, M.h, W.tacc[0]);
            }
            bcast_vel<t>(W.T);
        });
        {
            const JointK J = leg_joint<1>(q, M);
            joint_velocities<KISO, KANI>(&W.b[0], &robot::body[1], W.b[1], robot::body[2], J, M.h, W.tacc[1]);
        }
        {
            const JointK J = leg_joint<2>(q, M);
            joint_velocities<KANI, KANI>(&W.b[1], &robot::body[2], W.b[2], robot::body[3], J, M.h, W.tacc[2]);
        }

        /* recorded contacts, per body in record order */
        static_for<0, 3>([&](auto lb) {
            int n = ncount[lb] < leg_cap(M, lb) ? ncount[lb] : leg_cap(M, lb);
            for (int k = 0; k < n; ++k) {
                V3 l, nn;
                cs.get(leg_slot(M, lb) + k, l, nn);
                if constexpr (lb == 0)
                    contact_velocity<KISO>(W.T, robot::body[0], l, nn);
                else
                    contact_velocity<KANI>(W.b[lb], robot::body[lb + 1], l, nn);
            }
        });
    }

    /* epilogue: the lane's three joints, time, stability of the world */
    const int q = q0;
    {
        const JointK J0 = leg_joint<0>(q, M), J1 = leg_joint<1>(q, M), J2 = leg_joint<2>(q, M);
        W.angle[0] = joint_angle<true>(&W.T, W.b[0], J0);
        W.rate[0] = joint_rate<true>(&W.T, W.b[0], J0);
        W.angle[1] = joint_angle<true>(&W.b[0], W.b[1], J1);
        W.rate[1] = joint_rate<true>(&W.b[0], W.b[1], J1);
        W.angle[2] = joint_angle<true>(&W.b[1], W.b[2], J2);
        W.rate[2] = joint_rate<true>(&W.b[1], W.b[2], J2);
        W.torque[0] = W.tacc[0] / (float)M.substeps;
        W.torque[1] = W.tacc[1] / (float)M.substeps;
        W.torque[2] = W.tacc[2] / (float)M.substeps;
    }
    W.time += M.dt;
    int bad = body_unstable(W.T) | body_unstable(W.b[0]) | body_unstable(W.b[1]) | body_unstable(W.b[2]);
    if (qor(bad))
        W.unstable = 1;
    return overflow;
}

} /* namespace xpbd */
