/*
 * lane_engine.h -- one control step of one quadruped world, run by a single
 * lane with the whole world state in registers.
 *
 * The robot's topology is fixed (csrc/arc_robot.c:180-229): 13 bodies
 * (torso, then coxa / femur / tibia per leg), 12 hinges (joint 3l+k drives
 * body 3l+k+1), 21 spheres (torso; per leg a femur sphere and four tibia
 * spheres).  Every loop over bodies, joints and spheres is unrolled at compile
 * time, so all indices are constants and the solver runs in exactly the
 * reference's Gauss-Seidel order (csrc/arc_phys.c:707-747) with each world in
 * its own lane.  Numeric constants come from the Model (device constant
 * memory in the kernels).
 *
 * The same code compiles for the host: the CPU parity harness of the tests
 * runs it against the oracle without a GPU.
 */
#pragma once
#include <type_traits>
#include "xpbd_core.h"
#include "robot_consts.h"  /* generated at build time by gen_robot.cpp */

namespace xpbd {

constexpr int NB = 13, NJ = 12, NS = 21, LEGS = 4;
constexpr int NCH = LEGS + 2;          /* contact channels: feet, torso, other */
constexpr int MAXC = 128;              /* contact record cap, arc_phys.h:22 */

constexpr int body_kind(int b) { return (b == 0 || (b - 1) % 3 == 0) ? KISO : KANI; }
constexpr int joint_parent(int j) { return j % 3 == 0 ? 0 : j; }
constexpr int joint_child(int j) { return j + 1; }
constexpr int sphere_body(int s) { return s == 0 ? 0 : ((s - 1) % 5 == 0 ? 2 + 3 * ((s - 1) / 5) : 3 + 3 * ((s - 1) / 5)); }
constexpr int sphere_at(int s) { return (s == 0 || (s - 1) % 5 == 0) ? AT_ZERO : AT_Z; }
constexpr int sphere_channel(int s) { return s == 0 ? LEGS : ((s - 1) % 5 == 4 ? (s - 1) / 5 : LEGS + 1); }
/* bodies carrying spheres, in contact-list slot order */
constexpr int NCB = 9;
constexpr int cbody(int i) { return i == 0 ? 0 : (i <= 4 ? 2 + 3 * (i - 1) : 3 + 3 * (i - 5)); }
constexpr int cbody_index(int b) { return b == 0 ? 0 : ((b - 2) % 3 == 0 ? 1 + (b - 2) / 3 : 5 + (b - 3) / 3); }

/* Values the optimiser cannot see through.  Addresses derived from them are
 * recomputed where used instead of being hoisted out of the substep loop,
 * which would keep hundreds of loop-invariant 64-bit addresses live. */
XD size_t launder(size_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(v));
#endif
    return v;
}
template <class T> XD T *launder(T *p) { return (T *)launder((size_t)p); }
/* A uniform value as a scalar register pair of its own.  Values taken from the
 * kernel's argument structure otherwise stay parts of the wide scalar tuples
 * they were loaded in, and a tuple spilled under register pressure is
 * reloaded whole (one v_readlane per register) wherever one member is used. */
XD size_t own_sgpr(size_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(v));
#endif
    return v;
}
template <class T> XD T *own_sgpr(T *p) { return (T *)own_sgpr((size_t)p); }

template <int I, int N, class F> XD void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

static_assert(robot::NBODY == NB && robot::NJOINT == NJ && robot::NSPHERE == NS, "robot topology");

/* Every drive has the same gain and torque limit (csrc/arc_robot.c:180-195),
 * so the three products with the substep length are the same for all joints:
 * the kernels read joint 0's, three values instead of thirty-six. */
constexpr bool drives_alike()
{
    for (int j = 1; j < NJ; ++j)
        if (!(robot::joint[j].kp == robot::joint[0].kp && robot::joint[j].tmax == robot::joint[0].tmax))
            return false;
    return true;
}
static_assert(drives_alike(), "the kernels share joint 0's substep products among all drives");

/* joint constants as the kernel sees them: geometry and gains compiled in
 * (robot::joint), the three products with the substep length from the Model */
#define ROBOT_JOINT(ji, M)                            \
    JointK J = robot::joint[ji];                      \
    J.soft = (M).joint[0].soft;                       \
    J.cap_pos = (M).joint[0].cap_pos;                 \
    J.tmax_h = (M).joint[0].tmax_h

struct Model {
    BodyK body[NB];
    JointK joint[NJ];
    SphereK sphere[NS];
    TerrainK terr;
    V3 g;
    float mus, mud;
    float dt, h, hh, inv_h, two_h, inv_sub;
    int substeps;
    /* contact list: per sphere-carrying body a slot range of the per-world
     * scratch list, sized from the spheres' cell windows */
    int cslot[NCB], ccap[NCB];
    int cslots;
};

/* Register-resident state of one world during a step.  The spheres'
 * friction anchors live in an anchor store (HBM on the device: they are
 * touched only while a sphere is in contact), the substep-start poses in a
 * pose store (LDS on the device), and the per-sphere contact counts of the
 * step are packed four to a register. */
constexpr int NTP = (NS + 3) / 4;
struct LaneWorld {
    Body b[NB];
    float target[NJ], tacc[NJ];
    float angle[NJ], rate[NJ], torque[NJ];
    uint32_t touchp[NTP]; /* byte s % 4 of word s / 4: substeps sphere s touched */
    uint32_t anchored;    /* bit s: sphere s has a valid friction anchor */
    int unstable;
    float time;
};
XD int touch_of(const LaneWorld &W, int s) { return (int)((W.touchp[s >> 2] >> (8 * (s & 3))) & 0xffu); }

/* anchors and substep-start poses of the host build: plain arrays */
struct ArrayAnchors {
    V3 a[NS], n[NS];
    XD AnchorRef at(int s) { return AnchorRef{ a[s], n[s] }; }
    XD ArrayAnchors &fresh() { return *this; }
};
struct ArrayPoses {
    Pose0 p[NB];
    XD void store(int b, const Pose0 &v) { p[b] = v; }
    XD Pose0 load(int b) const { return p[b]; }
};

/* Contact list storage: slot k field f of this world. */
struct ContactStore {
    float *base;   /* [slot * 6 + field] * stride + world */
    size_t stride;
    size_t world;
    /* contact counts of the last substep, [contact body][cstride] + cworld
     * (null: not kept) */
    uint8_t *counts = nullptr;
    size_t cstride = 0, cworld = 0;
    XD void count(int ci, int n) const
    {
        if (counts)
            counts[(size_t)ci * cstride + cworld] = (uint8_t)n;
    }
    XD void put(int slot, V3 l, V3 n) const
    {
        float *p = base + (size_t)slot * 6 * stride + world;
        p[0] = l.x; p[stride] = l.y; p[2 * stride] = l.z;
        p[3 * stride] = n.x; p[4 * stride] = n.y; p[5 * stride] = n.z;
    }
    XD void get(int slot, V3 &l, V3 &n) const
    {
        const float *p = base + (size_t)slot * 6 * stride + world;
        l = v3(p[0], p[stride], p[2 * stride]);
        n = v3(p[3 * stride], p[4 * stride], p[5 * stride]);
    }
    /* optional LDS home of the first CMIRROR records of one list (the list
     * most contacts land in): record k field f at mirror[(k * 6 + f) * mstride] */
    float *mirror = nullptr;
    int mstride = 0;
    XD void mput(int k, V3 l, V3 n) const
    {
        float *p = mirror + k * 6 * mstride;
        p[0] = l.x; p[mstride] = l.y; p[2 * mstride] = l.z;
        p[3 * mstride] = n.x; p[4 * mstride] = n.y; p[5 * mstride] = n.z;
    }
    XD void mget(int k, V3 &l, V3 &n) const
    {
        const float *p = mirror + k * 6 * mstride;
        l = v3(p[0], p[mstride], p[2 * mstride]);
        n = v3(p[3 * mstride], p[4 * mstride], p[5 * mstride]);
    }
};
constexpr int CMIRROR = 4;

/* One control step (arc_world_step, csrc/arc_phys.c:707-747).
 * Returns a bit mask of bodies whose contact list overflowed its slots
 * (never expected for the robot; reported as an error by the host). */
/* terrain source of a lane step: column heights read from the world's grid */
struct GridTile {
    XD GridCells cells(const TerrainK &t, const TerrainRef &r) const { return GridCells{ t, r }; }
};

template <class Anchors, class Poses, class Tile = GridTile>
XD int lane_step(LaneWorld &W, Anchors &A0, Poses &P0, const Model &M, const TerrainRef &tr0, const ContactStore &cs0,
                 const Tile &T0 = Tile())
{
    int overflow = 0;
    /* prologue: refresh caches, clear counters (lim is precomputed per joint) */
    static_for<0, NB>([&](auto bi) { bsync<body_kind(bi)>(W.b[bi], robot::body[bi]); });
    static_for<0, NTP>([&](auto k) { W.touchp[k] = 0u; });
    static_for<0, NJ>([&](auto ji) { W.tacc[ji] = 0.0f; });

    for (int sub = 0; sub < M.substeps; ++sub) {
        /* per-substep copies of the memory handles, see launder() */
        auto &&A = A0.fresh();
        ContactStore cs = cs0;
        cs.world = launder(cs0.world);
        TerrainRef tr = tr0;
        tr.q = launder(tr0.q);
        static_for<0, NB>([&](auto bi) { P0.store(bi, integrate<body_kind(bi)>(W.b[bi], robot::body[bi], M.g, M.h)); });

        /* joint positions, Gauss-Seidel in joint order */
        static_for<0, NJ>([&](auto ji) {
            constexpr int P = joint_parent(ji), C = joint_child(ji);
            ROBOT_JOINT(ji, M);
            joint_positions<body_kind(P), body_kind(C)>(&W.b[P], &robot::body[P], W.b[C], robot::body[C], J,
                                                        W.target[ji], M.hh, W.tacc[ji]);
        });

        /* spheres against the terrain, in sphere order; contacts recorded
         * per body (the velocity pass below only couples a contact with its
         * own body, so per-body order is the reference's order) */
        int ncount[NCB];
        int total = 0;
        static_for<0, NCB>([&](auto ci) { ncount[ci] = 0; });
        /* each sphere's columns are fetched while the previous sphere is solved */
        const auto cells = T0.cells(M.terr, tr);
        static_for<0, NS>([&](auto si) {
            constexpr int B = sphere_body(si);
            constexpr int CI = cbody_index(B);
            int anch = (W.anchored >> si) & 1u;
            int hit = 0;
            auto rec = [&](V3 l, V3 n) {
                if (total < MAXC) {
                    if (ncount[CI] < M.ccap[CI])
                        cs.put(M.cslot[CI] + ncount[CI], l, n);
                    else
                        overflow |= 1 << CI;
                    ncount[CI]++;
                    total++;
                }
            };
            auto an = A.at(si);
            auto pose0 = [&]() { return P0.load(B); };
            collide<body_kind(B), sphere_at(si)>(W.b[B], robot::body[B], robot::sphere[si].at,
                                                 robot::sphere[si].rad, hit, anch, an, pose0, M.terr, cells, rec,
                                                 M.mus, M.mud);
            W.anchored = (W.anchored & ~(1u << si)) | ((uint32_t)anch << si);
            W.touchp[si >> 2] += (uint32_t)hit << (8 * (si & 3));
        });

        static_for<0, NB>([&](auto bi) { velocities(W.b[bi], P0.load(bi), M.inv_h, M.two_h); });

        static_for<0, NJ>([&](auto ji) {
            constexpr int P = joint_parent(ji), C = joint_child(ji);
            ROBOT_JOINT(ji, M);
            joint_velocities<body_kind(P), body_kind(C)>(&W.b[P], &robot::body[P], W.b[C], robot::body[C], J, M.h,
                                                         W.tacc[ji]);
        });

        static_for<0, NCB>([&](auto ci) {
            constexpr int B = cbody(ci);
            int n = ncount[ci] < M.ccap[ci] ? ncount[ci] : M.ccap[ci];
            for (int k = 0; k < n; ++k) {
                V3 l, nn;
                cs.get(M.cslot[ci] + k, l, nn);
                contact_velocity<body_kind(B)>(W.b[B], robot::body[B], l, nn);
            }
            if (sub == M.substeps - 1)
                cs.count(ci, n); /* the last substep's records, for the host snapshot */
        });
    }

    /* epilogue: joint outputs, time, stability */
    static_for<0, NJ>([&](auto ji) {
        constexpr int P = joint_parent(ji), C = joint_child(ji);
        W.angle[ji] = joint_angle<true>(&W.b[P], W.b[C], robot::joint[ji]);
        W.rate[ji] = joint_rate<true>(&W.b[P], W.b[C], robot::joint[ji]);
        W.torque[ji] = W.tacc[ji] / (float)M.substeps;
    });
    W.time += M.dt;
    bool bad = false;
    static_for<0, NB>([&](auto bi) { bad |= body_unstable(W.b[bi]); });
    if (bad)
        W.unstable = 1;
    return overflow;
}

} /* namespace xpbd */
