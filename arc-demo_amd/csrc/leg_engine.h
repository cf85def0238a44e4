/*
 * leg_engine.h -- one control step of one quadruped world run by a group of
 * P lanes: each lane holds K = 4 / P legs (coxa, femur, tibia) and a replica
 * of the torso, which every lane of the group keeps identical.  P = 4 (a
 * quad, one leg per lane) and P = 2 (a pair, two legs per lane).
 *
 * The reference projects the joints in the order 0..11 (csrc/arc_phys.c:
 * 726-737, Gauss-Seidel).  Joints 3l (hip yaw) couple the torso with leg l;
 * joints 3l+1 and 3l+2 touch only leg l's own bodies.  So the order
 *     root 0, root 1, root 2, root 3, then every leg's joint 1 and joint 2
 * applies the same sequence of updates to every body (operations on disjoint
 * bodies commute exactly) and gives bit-identical results.  The four roots
 * run one after another (with P = 4 each by two lanes of the quad, the coxa's
 * and a partner holding the torso, RootsPair; with P = 2 in the lane holding
 * its leg) and the torso is broadcast to the group (DPP) after each; the mid
 * and distal joints of the group's legs run at once.  Spheres,
 * integration and the velocity passes act on one body at a time: each lane
 * does its legs' bodies and the torso's (the same arithmetic in every lane).
 *
 * With P = 4 a lane does 4 bodies, 6 spheres and 6 joint solves per substep
 * against 13, 21 and 12 in the one-world-per-lane form (lane_engine.h): a
 * 2.5x shorter dependent chain per world, for latency-bound batches and, at
 * two wavefronts per SIMD, the best throughput too.  P = 2 halves the
 * redundant work but holds 7 bodies per lane; it spills registers and is
 * slower on gfx950 (kept as a tuned alternative).
 */
#pragma once
#include "lane_engine.h"
#include "root_split.h"

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

/* Cross-lane helpers of a world's group of P lanes (P = 4: a quad; P = 2: a
 * pair of a quad), DPP quad_perm within each quad. */
template <int P, int T> __device__ __forceinline__ float gbcast(float v)
{
    /* lane T of the group, in every lane of the group */
    constexpr int ctrl = P == 4 ? (T | (T << 2) | (T << 4) | (T << 6))
                                : (T | (T << 2) | ((2 + T) << 4) | ((2 + T) << 6));
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), ctrl, 0xf, 0xf, false));
}
template <int P> __device__ __forceinline__ int gor(int v)
{
    v |= __builtin_amdgcn_update_dpp(v, v, 0xb1, 0xf, 0xf, false); /* quad_perm 1,0,3,2 */
    if (P == 4)
        v |= __builtin_amdgcn_update_dpp(v, v, 0x4e, 0xf, 0xf, false); /* quad_perm 2,3,0,1 */
    return v;
}
template <int P, int T> __device__ __forceinline__ void gbcast3(V3 &a)
{
    a.x = gbcast<P, T>(a.x);
    a.y = gbcast<P, T>(a.y);
    a.z = gbcast<P, T>(a.z);
}

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
    J.soft = M.joint[0].soft; /* the same for every drive (drives_alike) */
    J.cap_pos = M.joint[0].cap_pos;
    J.tmax_h = M.joint[0].tmax_h;
    return J;
}

/* Where a lane finds the constants of its legs' joints: selected from the
 * compile-time constants of the four legs at each use (SelJoints), or, for the
 * mid and distal joints, read from a table of the four legs' geometry in LDS
 * (LdsJoints: [m - 1][leg][r1, r2, a1, a2, b1, b2], filled by fill_joint_table),
 * which costs LDS reads instead of three selects per value.  Same values. */
constexpr int JGEO = 18;
struct SelJoints {
    template <int M_> __device__ __forceinline__ JointK get(int leg, const Model &M) const { return leg_joint<M_>(leg, M); }
};
struct LdsJoints {
    const float *g;
    template <int M_> __device__ __forceinline__ JointK get(int leg, const Model &M) const
    {
        if constexpr (M_ == 0) {
            return leg_joint<0>(leg, M);
        } else {
            JointK J = leg_joint<M_>(0, M); /* gains, limits and the drive's constants: every leg's */
            const float *p = g + ((M_ - 1) * LEGS + leg) * JGEO;
            J.r1 = v3(p[0], p[1], p[2]);
            J.r2 = v3(p[3], p[4], p[5]);
            J.a1 = v3(p[6], p[7], p[8]);
            J.a2 = v3(p[9], p[10], p[11]);
            J.b1 = v3(p[12], p[13], p[14]);
            J.b2 = v3(p[15], p[16], p[17]);
            return J;
        }
    }
};
/* thread 0 of the workgroup writes the table (2 x 4 x 18 values); the caller
 * synchronises the workgroup before its first read */
__device__ __forceinline__ void fill_joint_table(float *g)
{
    if (threadIdx.x != 0)
        return;
    static_for<1, 3>([&](auto m) {
        static_for<0, LEGS>([&](auto l) {
            constexpr JointK J = robot::joint[3 * l + m];
            float *p = g + ((m - 1) * LEGS + l) * JGEO;
            const V3 v[6] = { J.r1, J.r2, J.a1, J.a2, J.b1, J.b2 };
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                p[3 * k] = v[k].x;
                p[3 * k + 1] = v[k].y;
                p[3 * k + 2] = v[k].z;
            }
        });
    });
}

/* State of one world in one lane of its group: the torso replica, the
 * lane's K legs (bodies 0 coxa, 1 femur, 2 tibia) and their joints; spheres
 * in lane-local order 0 torso, then 1 + 5 l + m for sphere m of local leg l. */
template <int K> struct LegWorld {
    Body T;
    Body b[K][3];
    float target[3 * K], tacc[3 * K], angle[3 * K], rate[3 * K], torque[3 * K];
    uint32_t touchp[(1 + 5 * K + 3) / 4]; /* byte k of word k / 4: substeps local sphere k touched */
    uint32_t anchored;                    /* bit k: local sphere k has a valid friction anchor */
    int unstable;
    float time;
};
template <int K> __device__ __forceinline__ int leg_touch(const LegWorld<K> &W, int k)
{
    return (int)((W.touchp[k >> 2] >> (8 * (k & 3))) & 0xffu);
}

/* contact slots of one lane: the torso's range, then its legs' femur
 * ranges, then their tibia ranges (every leg's femur and tibia carry the
 * same spheres: ranges of ccap[1] and ccap[5]) */
template <int K> __device__ __forceinline__ int leg_slot(const Model &M, int kind, int l)
{
    return kind == 0 ? 0 : kind == 1 ? M.ccap[0] + l * M.ccap[1] : M.ccap[0] + K * M.ccap[1] + l * M.ccap[5];
}
__device__ __forceinline__ int leg_cap(const Model &M, int kind) { return kind == 0 ? M.ccap[0] : kind == 1 ? M.ccap[1] : M.ccap[5]; }

/* the torso's pose / velocity after lane T's root joints, in the group */
template <int P, int T> __device__ __forceinline__ void bcast_pose(Body &B)
{
    gbcast3<P, T>(B.x);
    B.q.w = gbcast<P, T>(B.q.w); B.q.x = gbcast<P, T>(B.q.x); B.q.y = gbcast<P, T>(B.q.y); B.q.z = gbcast<P, T>(B.q.z);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            B.R.e[r][c] = gbcast<P, T>(B.R.e[r][c]);
}
template <int P, int T> __device__ __forceinline__ void bcast_vel(Body &B)
{
    gbcast3<P, T>(B.v);
    gbcast3<P, T>(B.w);
}

/* The four root (hip) joints, which all move the torso: solved in turn, each
 * by the lane holding its leg, the torso broadcast to the group after each
 * owner's last root (positions, then velocities). */
template <int K> struct RootsLocal {
    __device__ __forceinline__ void positions(LegWorld<K> &W, int q, const Model &M)
    {
        constexpr int P = LEGS / K;
        static_for<0, LEGS>([&](auto t) {
            constexpr int owner = t / K, l = t % K;
            if (q == owner) {
                ROBOT_JOINT(3 * t, M);
                joint_positions<KISO, KISO>(&W.T, &robot::body[0], W.b[l][0], robot::body[1], J, W.target[3 * l],
                                            M.hh, W.tacc[3 * l]);
            }
            if constexpr (l == K - 1)
                bcast_pose<P, owner>(W.T);
        });
    }
    __device__ __forceinline__ void velocities(LegWorld<K> &W, int q, const Model &M)
    {
        constexpr int P = LEGS / K;
        static_for<0, LEGS>([&](auto t) {
            constexpr int owner = t / K, l = t % K;
            if (q == owner) {
                ROBOT_JOINT(3 * t, M);
                joint_velocities<KISO, KISO>(&W.T, &robot::body[0], W.b[l][0], robot::body[1], J, M.h,
                                             W.tacc[3 * l]);
            }
            if constexpr (l == K - 1)
                bcast_vel<P, owner>(W.T);
        });
    }
};

/* The four root joints of the quad's world (one leg per lane, K = 1), each
 * solved by two lanes (root_split.h): root t by lane t (its coxa, the child
 * side) and lane t ^ 1 (the torso, the parent side), so half the quad works
 * on the roots instead of a quarter.  Lane t moves its coxa into its torso
 * registers for the solve and takes it back afterwards; every lane's torso is
 * then replaced by the parent side's result (DPP), which also restores lane
 * t's.  The velocity pass stays with the owner lane (RootsLocal). */
struct RootsPair {
    __device__ __forceinline__ static void swap_in(bool child, Body &T, const Body &C)
    {
        T.x = sel3(child, C.x, T.x);
        T.q = q4(child ? C.q.w : T.q.w, child ? C.q.x : T.q.x, child ? C.q.y : T.q.y, child ? C.q.z : T.q.z);
#pragma unroll
        for (int e = 0; e < 9; ++e)
            T.R.e[e / 3][e % 3] = child ? C.R.e[e / 3][e % 3] : T.R.e[e / 3][e % 3];
    }
    __device__ __forceinline__ void positions(LegWorld<1> &W, int q, const Model &M)
    {
        static_for<0, LEGS>([&](auto t) {
            constexpr int h = t ^ 1;
            const bool child = q == t;
            Body &C = W.b[0][0];
            swap_in(child, W.T, C);
            const float tg = gbcast<4, t>(W.target[0]);
            const float d = root_positions_pair<3 * t, QuadX<t, h>>(W.T, child, tg, M.hh, M);
            if (child) {
                W.tacc[0] = W.tacc[0] - d;
                C.x = W.T.x;
                C.q = W.T.q;
                C.R = W.T.R;
            }
            bcast_pose<4, h>(W.T);
        });
    }
    __device__ __forceinline__ void velocities(LegWorld<1> &W, int q, const Model &M)
    {
        RootsLocal<1> local;
        local.velocities(W, q, M);
    }
};

/* The root joints of a workgroup's WPB worlds (one leg per lane, K = 1)
 * solved by WPB lanes, one world each, through LDS: every lane publishes its
 * coxa (and lane 0 of the group the torso), lane r of the block solves the
 * four roots of world r in joint order, and every lane takes the torso and its
 * coxa back.  The roots then use all lanes of one wavefront instead of one lane
 * in four of every wavefront.  xs: [field][WPB] floats, RS_FIELDS fields. */
constexpr int RS_FIELDS = 88, RS_VFIELDS = 28; /* the pose and the velocity exchange */
template <int WPB> struct RootsShared {
    float *xs; /* pose exchange: torso, coxas (x, q, R), targets, torque sums */
    float *xv; /* velocity exchange: torso (w, R), coxas (w), torque sums; a buffer of its own, so that
                  each exchange needs no barrier before it is written (the other's barriers order it) */
    int g;        /* the lane's world slot in the block */
    int turn = 0; /* the wavefront whose first WPB lanes solve the roots: turns by substep, spreading the
                     chains over the SIMDs */
    __device__ __forceinline__ bool solver() const { return (int)(threadIdx.x / 64) == turn && (int)(threadIdx.x % 64) < WPB; }
    static constexpr int F_T = 0, F_C = 16, F_TG = 80, F_TA = 84; /* torso, coxas, targets, torque sums */
    __device__ __forceinline__ float &at(int f, int w) const { return xs[f * WPB + w]; }
    __device__ __forceinline__ void put_pose(int f, int w, const Body &b) const
    {
        at(f + 0, w) = b.x.x; at(f + 1, w) = b.x.y; at(f + 2, w) = b.x.z;
        at(f + 3, w) = b.q.w; at(f + 4, w) = b.q.x; at(f + 5, w) = b.q.y; at(f + 6, w) = b.q.z;
#pragma unroll
        for (int e = 0; e < 9; ++e)
            at(f + 7 + e, w) = b.R.e[e / 3][e % 3];
    }
    __device__ __forceinline__ void get_pose(int f, int w, Body &b) const
    {
        b.x = v3(at(f + 0, w), at(f + 1, w), at(f + 2, w));
        b.q = q4(at(f + 3, w), at(f + 4, w), at(f + 5, w), at(f + 6, w));
#pragma unroll
        for (int e = 0; e < 9; ++e)
            b.R.e[e / 3][e % 3] = at(f + 7 + e, w);
    }
    static constexpr int V_T = 0, V_C = 12, V_TA = 24; /* torso w, R; coxa w; torque sums */
    __device__ __forceinline__ float &atv(int f, int w) const { return xv[f * WPB + w]; }
    __device__ __forceinline__ void positions(LegWorld<1> &W, int q, const Model &M)
    {
        put_pose(F_C + 16 * q, g, W.b[0][0]);
        if (q == 0)
            put_pose(F_T, g, W.T);
        at(F_TG + q, g) = W.target[0];
        at(F_TA + q, g) = W.tacc[0];
        __syncthreads();
        if (solver()) {
            /* one coxa in registers at a time beside the torso */
            const int r = threadIdx.x % 64;
            Body T;
            get_pose(F_T, r, T);
            static_for<0, LEGS>([&](auto t) {
                Body C;
                get_pose(F_C + 16 * t, r, C);
                float ta = at(F_TA + t, r);
                ROBOT_JOINT(3 * t, M);
                joint_positions<KISO, KISO>(&T, &robot::body[0], C, robot::body[1], J, at(F_TG + t, r), M.hh, ta);
                put_pose(F_C + 16 * t, r, C);
                at(F_TA + t, r) = ta;
            });
            put_pose(F_T, r, T);
        }
        __syncthreads();
        get_pose(F_T, g, W.T);
        get_pose(F_C + 16 * q, g, W.b[0][0]);
        W.tacc[0] = at(F_TA + q, g);
    }
    __device__ __forceinline__ void velocities(LegWorld<1> &W, int q, const Model &M)
    {
        atv(V_C + 3 * q + 0, g) = W.b[0][0].w.x;
        atv(V_C + 3 * q + 1, g) = W.b[0][0].w.y;
        atv(V_C + 3 * q + 2, g) = W.b[0][0].w.z;
        if (q == 0) {
            atv(V_T + 0, g) = W.T.w.x; atv(V_T + 1, g) = W.T.w.y; atv(V_T + 2, g) = W.T.w.z;
#pragma unroll
            for (int e = 0; e < 9; ++e)
                atv(V_T + 3 + e, g) = W.T.R.e[e / 3][e % 3];
        }
        atv(V_TA + q, g) = W.tacc[0];
        __syncthreads();
        if (solver()) {
            const int r = threadIdx.x % 64;
            Body T;
            T.w = v3(atv(V_T + 0, r), atv(V_T + 1, r), atv(V_T + 2, r));
#pragma unroll
            for (int e = 0; e < 9; ++e)
                T.R.e[e / 3][e % 3] = atv(V_T + 3 + e, r);
            static_for<0, LEGS>([&](auto t) {
                Body C;
                C.w = v3(atv(V_C + 3 * t + 0, r), atv(V_C + 3 * t + 1, r), atv(V_C + 3 * t + 2, r));
                float ta = atv(V_TA + t, r);
                ROBOT_JOINT(3 * t, M);
                joint_velocities<KISO, KISO>(&T, &robot::body[0], C, robot::body[1], J, M.h, ta);
                atv(V_C + 3 * t + 0, r) = C.w.x;
                atv(V_C + 3 * t + 1, r) = C.w.y;
                atv(V_C + 3 * t + 2, r) = C.w.z;
                atv(V_TA + t, r) = ta;
            });
            atv(V_T + 0, r) = T.w.x; atv(V_T + 1, r) = T.w.y; atv(V_T + 2, r) = T.w.z;
        }
        __syncthreads();
        W.T.w = v3(atv(V_T + 0, g), atv(V_T + 1, g), atv(V_T + 2, g));
        W.b[0][0].w = v3(atv(V_C + 3 * q + 0, g), atv(V_C + 3 * q + 1, g), atv(V_C + 3 * q + 2, g));
        W.tacc[0] = atv(V_TA + q, g);
        turn = (turn + 1) % (4 * WPB / 64); /* the block's wavefronts */
    }
};

/* One control step of the group's world by lane q (legs q K .. q K + K - 1).
 * Anchors: global sphere indices; Poses: local bodies 0 torso, 1 + 3 l + m;
 * the contact store addresses this lane's own list.  Returns this lane's
 * contact-overflow bits (bit per global contact body, as the lane form). */
template <int K, class Anchors, class Poses, class Tile, class Roots, class JG = SelJoints>
__device__ __forceinline__ int leg_step(LegWorld<K> &W, int q0, Anchors &A0, Poses &P0, const Model &M,
                                        const TerrainRef &tr0, const ContactStore &cs0, const Tile &T0, Roots &RJ,
                                        const JG &jg = JG())
{
    constexpr int P = LEGS / K;
    int overflow = 0;
    bsync<KISO>(W.T, robot::body[0]);
    static_for<0, K>([&](auto l) {
        bsync<KISO>(W.b[l][0], robot::body[1]);
        bsync<KANI>(W.b[l][1], robot::body[2]);
        bsync<KANI>(W.b[l][2], robot::body[3]);
    });
    static_for<0, (1 + 5 * K + 3) / 4>([&](auto w) { W.touchp[w] = 0u; });
    static_for<0, 3 * K>([&](auto j) { W.tacc[j] = 0.0f; });

    for (int sub0 = 0; sub0 < M.substeps; ++sub0) {
        /* the lane index, opaque per substep: the per-leg joint constants are
         * selected where used instead of being held across the step */
        const int q = (int)launder((size_t)q0);
        auto &&A = A0.fresh();
        ContactStore cs = cs0;
        cs.world = launder(cs0.world);
        TerrainRef tr = tr0;
        tr.q = launder(tr0.q);
        P0.store(0, integrate<KISO>(W.T, robot::body[0], M.g, M.h));
        static_for<0, K>([&](auto l) {
            P0.store(1 + 3 * l, integrate<KISO>(W.b[l][0], robot::body[1], M.g, M.h));
            P0.store(2 + 3 * l, integrate<KANI>(W.b[l][1], robot::body[2], M.g, M.h));
            P0.store(3 + 3 * l, integrate<KANI>(W.b[l][2], robot::body[3], M.g, M.h));
        });

        /* joint positions: the four roots in turn (leg t's in lane t / K),
         * the torso broadcast when the next root is another lane's; then
         * each leg's mid and distal joints */
        RJ.positions(W, q, M);
        static_for<0, K>([&](auto l) {
            {
                const JointK J = jg.template get<1>(q * K + l, M);
                joint_positions<KISO, KANI>(&W.b[l][0], &robot::body[1], W.b[l][1], robot::body[2], J,
                                            W.target[3 * l + 1], M.hh, W.tacc[3 * l + 1]);
            }
            {
                const JointK J = jg.template get<2>(q * K + l, M);
                joint_positions<KANI, KANI>(&W.b[l][1], &robot::body[2], W.b[l][2], robot::body[3], J,
                                            W.target[3 * l + 2], M.hh, W.tacc[3 * l + 2]);
            }
        });

        /* spheres: the torso's (every lane), then each leg's five in order */
        int ncount[1 + 2 * K];
        static_for<0, 1 + 2 * K>([&](auto c) { ncount[c] = 0; });
        const auto cells = T0.cells(M.terr, tr);
        static_for<0, 1 + 5 * K>([&](auto k) {
            constexpr int l = k == 0 ? 0 : (k - 1) / 5;        /* local leg */
            constexpr int m = k == 0 ? 0 : (k - 1) % 5;        /* the leg's sphere */
            constexpr int KIND = k == 0 ? 0 : (m == 0 ? 1 : 2); /* 0 torso, 1 femur, 2 tibia */
            constexpr int NC = k == 0 ? 0 : (KIND == 1 ? 1 + l : 1 + K + l);
            constexpr int GS = k == 0 ? 0 : 1 + m;              /* the sphere's constants: leg 0's */
            int anch = (W.anchored >> k) & 1u;
            int hit = 0;
            auto rec = [&](V3 lp, V3 n) {
                /* a lane's first tibia records live in LDS; HBM gets the others
                 * and, in the last substep, all of them (the host snapshot's) */
                bool hbm = true;
                if constexpr (KIND == 2 && K == 1)
                    if (cs.mirror && ncount[NC] < CMIRROR) {
                        cs.mput(ncount[NC], lp, n);
                        hbm = sub0 == M.substeps - 1;
                    }
                if (ARC_LIKELY(ncount[NC] < leg_cap(M, KIND))) {
                    if (hbm)
                        cs.put(leg_slot<K>(M, KIND, l) + ncount[NC], lp, n);
                } else {
                    overflow |= 1 << (KIND == 0 ? 0 : KIND == 1 ? 1 + q0 * K + l : 5 + q0 * K + l);
                }
                ncount[NC]++;
            };
            auto an = A.at(k == 0 ? 0 : 5 * (q * K + l) + 1 + m);
            if constexpr (k == 0) {
                auto pose0 = [&]() { return P0.load(0); };
                collide<KISO, sphere_at(0)>(W.T, robot::body[0], robot::sphere[0].at, robot::sphere[0].rad, hit,
                                            anch, an, pose0, M.terr, cells, rec, M.mus, M.mud);
            } else {
                auto pose0 = [&]() { return P0.load(1 + 3 * l + KIND); };
                collide<KANI, sphere_at(GS)>(W.b[l][KIND], robot::body[1 + KIND], robot::sphere[GS].at,
                                             robot::sphere[GS].rad, hit, anch, an, pose0, M.terr, cells, rec,
                                             M.mus, M.mud);
            }
            W.anchored = (W.anchored & ~(1u << k)) | ((uint32_t)anch << k);
            W.touchp[k >> 2] += (uint32_t)hit << (8 * (k & 3));
        });

        velocities(W.T, P0.load(0), M.inv_h, M.two_h);
        static_for<0, K>([&](auto l) {
            velocities(W.b[l][0], P0.load(1 + 3 * l), M.inv_h, M.two_h);
            velocities(W.b[l][1], P0.load(2 + 3 * l), M.inv_h, M.two_h);
            velocities(W.b[l][2], P0.load(3 + 3 * l), M.inv_h, M.two_h);
        });

        RJ.velocities(W, q, M);
        static_for<0, K>([&](auto l) {
            {
                const JointK J = jg.template get<1>(q * K + l, M);
                joint_velocities<KISO, KANI>(&W.b[l][0], &robot::body[1], W.b[l][1], robot::body[2], J, M.h,
                                             W.tacc[3 * l + 1]);
            }
            {
                const JointK J = jg.template get<2>(q * K + l, M);
                joint_velocities<KANI, KANI>(&W.b[l][1], &robot::body[2], W.b[l][2], robot::body[3], J, M.h,
                                             W.tacc[3 * l + 2]);
            }
        });

        /* recorded contacts, per body in record order */
        static_for<0, 1 + 2 * K>([&](auto c) {
            constexpr int KIND = c == 0 ? 0 : (c <= K ? 1 : 2);
            constexpr int l = c == 0 ? 0 : (c <= K ? c - 1 : c - 1 - K);
            int n = ncount[c] < leg_cap(M, KIND) ? ncount[c] : leg_cap(M, KIND);
            for (int k = 0; k < n; ++k) {
                V3 lp, nn;
                bool lds = false;
                if constexpr (KIND == 2 && K == 1)
                    lds = cs.mirror && k < CMIRROR;
                if (lds)
                    cs.mget(k, lp, nn);
                else
                    cs.get(leg_slot<K>(M, KIND, l) + k, lp, nn);
                if constexpr (KIND == 0)
                    contact_velocity<KISO>(W.T, robot::body[0], lp, nn);
                else
                    contact_velocity<KANI>(W.b[l][KIND], robot::body[1 + KIND], lp, nn);
            }
            /* the last substep's records, for the host snapshot: contact body
             * 0 the torso (lane 0 of the group), 1 + g leg g's femur, 5 + g its
             * tibia */
            if (sub0 == M.substeps - 1 && (KIND != 0 || q0 == 0))
                cs.count(KIND == 0 ? 0 : KIND == 1 ? 1 + q0 * K + l : 5 + q0 * K + l, n);
        });
    }

    /* epilogue: the lane's joints, time, stability of the world */
    const int q = q0;
    static_for<0, K>([&](auto l) {
        const JointK J0 = leg_joint<0>(q * K + l, M), J1 = leg_joint<1>(q * K + l, M), J2 = leg_joint<2>(q * K + l, M);
        W.angle[3 * l] = joint_angle<true>(&W.T, W.b[l][0], J0);
        W.rate[3 * l] = joint_rate<true>(&W.T, W.b[l][0], J0);
        W.angle[3 * l + 1] = joint_angle<true>(&W.b[l][0], W.b[l][1], J1);
        W.rate[3 * l + 1] = joint_rate<true>(&W.b[l][0], W.b[l][1], J1);
        W.angle[3 * l + 2] = joint_angle<true>(&W.b[l][1], W.b[l][2], J2);
        W.rate[3 * l + 2] = joint_rate<true>(&W.b[l][1], W.b[l][2], J2);
        W.torque[3 * l] = W.tacc[3 * l] / (float)M.substeps;
        W.torque[3 * l + 1] = W.tacc[3 * l + 1] / (float)M.substeps;
        W.torque[3 * l + 2] = W.tacc[3 * l + 2] / (float)M.substeps;
    });
    W.time += M.dt;
    int bad = body_unstable(W.T);
    static_for<0, K>([&](auto l) {
        bad |= body_unstable(W.b[l][0]) | body_unstable(W.b[l][1]) | body_unstable(W.b[l][2]);
    });
    if (gor<P>(bad))
        W.unstable = 1;
    return overflow;
}

} /* namespace xpbd */
