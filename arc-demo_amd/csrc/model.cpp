/*
 * model.cpp -- binary32 builders for bodies, spheres and hinges and the
 * quadruped, evaluated in the reference's order so that the constants equal
 * those of the reference compiled in FP32:
 *   csrc/arc_phys.c:32-39 (world init), 84-143 (builders), 749-782 (pose),
 *   801-821 (inertia helpers); csrc/arc_robot.c:51-70, 98-144, 180-229
 *   (the robot, without its render-only visuals); csrc/arc_terrain.c:17-30.
 */
#include "model.h"
#include <string.h>

namespace xpbd {

static const float PI_F = 3.14159265358979323846f;

void terrain_flat(HostWorld &w, int nx, int ny, float cell, float height)
{
    if (nx > GRID_MAX) nx = GRID_MAX;
    if (ny > GRID_MAX) ny = GRID_MAX;
    w.t.nx = nx;
    w.t.ny = ny;
    w.t.cell = cell;
    w.t.ox = -0.5f * nx * cell;
    w.t.oy = -0.5f * ny * cell;
    w.t.inv_cell = 1.0f / cell;
    w.t.fast_cell = 0;
    w.outside = height;
    w.h.assign((size_t)nx * ny, height);
}

void world_init(HostWorld &w)
{
    w.g = v3(0.0f, 0.0f, -9.81f);
    w.mus = 1.0f;
    w.mud = 0.8f;
    w.nb = w.nj = w.ns = 0;
    memset(w.bk, 0, sizeof(w.bk));
    memset(w.b, 0, sizeof(w.b));
    memset(w.j, 0, sizeof(w.j));
    memset(w.s, 0, sizeof(w.s));
    w.unstable = 0;
    w.time = 0.0f;
    terrain_flat(w, 1, 1, 1.0f, 0.0f);
}

static void sync_generic(Body &b, const BodyK &k) { bsync<KGEN>(b, k); }

int add_body(HostWorld &w, float mass, V3 inertia, V3 pos, Q4 q)
{
    if (w.nb >= MAX_BODIES)
        return -1;
    Body &b = w.b[w.nb];
    BodyK &k = w.bk[w.nb];
    memset(&b, 0, sizeof(b));
    b.x = pos;
    b.q = qunit(q);
    k.mass = mass;
    k.invm = 1.0f / mass;
    k.I = inertia;
    k.invI = v3(1.0f / inertia.x, 1.0f / inertia.y, 1.0f / inertia.z);
    k.iso_I = inertia.x == inertia.y && inertia.y == inertia.z;
    k.iso_invI = k.invI.x == k.invI.y && k.invI.y == k.invI.z;
    sync_generic(b, k);
    return w.nb++;
}

int add_sphere(HostWorld &w, int body, V3 at, float rad)
{
    if (w.ns >= MAX_SPHERES)
        return -1;
    HostSphere &s = w.s[w.ns];
    memset(&s, 0, sizeof(s));
    s.body = body;
    s.at = at;
    s.rad = rad;
    return w.ns++;
}

int add_hinge(HostWorld &w, int parent, int child, V3 anchor, V3 axis_w)
{
    if (w.nj >= MAX_JOINTS || child <= parent)
        return -1;
    Q4 qp = qid(), qc;
    V3 xp = zero3();
    if (parent >= 0) {
        xp = w.b[parent].x;
        qp = w.b[parent].q;
    }
    qc = w.b[child].q;
    V3 axis = unit(axis_w);
    V3 ref = perp(axis);
    HostJoint &j = w.j[w.nj];
    memset(&j, 0, sizeof(j));
    j.parent = parent;
    j.child = child;
    j.r1 = qirot(qp, sub(anchor, xp));
    j.r2 = qirot(qc, sub(anchor, w.b[child].x));
    j.a1 = qirot(qp, axis);
    j.a2 = qirot(qc, axis);
    j.b1 = qirot(qp, ref);
    j.b2 = qirot(qc, ref);
    j.qrel0 = qmul(conj(qp), qc);
    j.lo = 1.0f;
    j.hi = -1.0f;
    return w.nj++;
}

void set_pose(HostWorld &w, V3 pos, Q4 q, const float *angles)
{
    w.b[0].x = pos;
    w.b[0].q = qunit(q);
    for (int i = 0; i < w.nj; ++i) {
        HostJoint &j = w.j[i];
        Body &c = w.b[j.child];
        V3 xp = j.parent >= 0 ? w.b[j.parent].x : zero3();
        Q4 qp = j.parent >= 0 ? w.b[j.parent].q : qid();
        V3 anchor = add(xp, qrot(qp, j.r1));
        c.q = qunit(qmul(qp, qmul(qaxis(j.a1, angles[i]), j.qrel0)));
        c.x = sub(anchor, qrot(c.q, j.r2));
        j.target = angles[i];
        j.angle = angles[i];
        j.rate = 0.0f;
        j.torque = 0.0f;
    }
    for (int i = 0; i < w.nb; ++i) {
        Body &b = w.b[i];
        b.v = zero3();
        b.w = zero3();
        sync_generic(b, w.bk[i]);
    }
    for (int i = 0; i < w.ns; ++i) {
        w.s[i].touch = 0;
        w.s[i].anchored = 0;
    }
    w.unstable = 0;
}

V3 inertia_box(float m, V3 he)
{
    float k = m / 3.0f;
    return v3(k * (he.y * he.y + he.z * he.z), k * (he.x * he.x + he.z * he.z), k * (he.x * he.x + he.y * he.y));
}
V3 inertia_capsule_z(float m, float r, float half)
{
    float len = 2.0f * (half + r);
    float ixx = m * (3.0f * r * r + len * len) / 12.0f;
    return v3(ixx, ixx, 0.5f * m * r * r);
}
V3 inertia_sphere(float m, float r)
{
    float i = 0.4f * m * r * r;
    return v3(i, i, i);
}

/* one leg: coxa, femur, tibia, three hinges with PD drives, five spheres */
static void robot_leg(HostWorld &w, RobotInfo &info, int leg, float sx, float sy)
{
    const float R0 = 0.2f, ELEV = 0.785f, SPLAY = 0.785f;
    const float LC = 0.06f, LF = 0.50f, LT = 0.95f;
    float psi = sx > 0.0f ? sy * SPLAY : sy * (PI_F - SPLAY);
    V3 d = v3(f_cos(psi), f_sin(psi), 0.0f);
    V3 up = v3(0.0f, 0.0f, 1.0f);
    V3 lift = cross(d, up);
    Q4 along = qaxis(unit(cross(up, d)), 0.5f * PI_F);
    /* the radius times a libm result is a binary64 product in the reference */
    V3 hip = add(scl(d, (float)((double)R0 * cos((double)ELEV))),
                 v3(0.0f, 0.0f, (float)((double)R0 * sin((double)ELEV))));
    V3 p1 = mad(hip, d, LC), p2 = mad(p1, d, LF);
    float fl = 0.5f * LF, tl = 0.5f * LT;

    int bc = add_body(w, 0.4f, v3(0.02f, 0.02f, 0.02f), mad(hip, d, 0.5f * LC), qid());
    int bf = add_body(w, 1.5f, inertia_capsule_z(1.5f, 0.06f, fl), mad(p1, d, fl), along);
    int bt = add_body(w, 1.0f, inertia_capsule_z(1.0f, 0.05f, tl), mad(p2, d, tl), along);

    static const float LIM[3][2] = { { -0.8f, 0.8f }, { -0.4f, 1.5f }, { -2.8f, -0.5f } };
    int j0 = add_hinge(w, 0, bc, hip, v3(0.0f, 0.0f, -sy));
    add_hinge(w, bc, bf, p1, lift);
    add_hinge(w, bf, bt, p2, lift);
    for (int k = 0; k < 3; ++k) {
        HostJoint &j = w.j[j0 + k];
        j.lo = LIM[k][0];
        j.hi = LIM[k][1];
        j.kp = 250.0f;
        j.kd = 6.0f;
        j.tmax = 80.0f;
    }
    info.def[3 * leg + 0] = 0.0f;
    info.def[3 * leg + 1] = 0.6f;
    info.def[3 * leg + 2] = -1.95f;

    add_sphere(w, bf, v3(0.0f, 0.0f, 0.0f), 0.06f);
    add_sphere(w, bt, v3(0.0f, 0.0f, -tl), 0.07f);
    add_sphere(w, bt, v3(0.0f, 0.0f, -0.3f * tl), 0.055f);
    add_sphere(w, bt, v3(0.0f, 0.0f, 0.4f * tl), 0.035f);
    info.foot[leg] = add_sphere(w, bt, v3(0.0f, 0.0f, tl), 0.045f);
}

void robot_build(HostWorld &w, RobotInfo &info)
{
    static const float SX[ROBOT_LEGS] = { 1.0f, 1.0f, -1.0f, -1.0f };
    static const float SY[ROBOT_LEGS] = { 1.0f, -1.0f, 1.0f, -1.0f };
    add_body(w, 8.0f, inertia_sphere(8.0f, 0.2f), zero3(), qid());
    add_sphere(w, 0, zero3(), 0.2f);
    info.n_torso_spheres = w.ns;
    for (int leg = 0; leg < ROBOT_LEGS; ++leg)
        robot_leg(w, info, leg, SX[leg], SY[leg]);
    for (int i = 0; i < ROBOT_JOINTS; ++i) {
        info.lo[i] = w.j[i].lo;
        info.hi[i] = w.j[i].hi;
    }
    info.mass = 0.0f;
    for (int i = 0; i < w.nb; ++i)
        info.mass += w.bk[i].mass;
    set_pose(w, zero3(), qid(), info.def);
    float lowest = 1e30f;
    info.footprint = 0.0f;
    for (int leg = 0; leg < ROBOT_LEGS; ++leg) {
        const HostSphere &s = w.s[info.foot[leg]];
        const Body &b = w.b[s.body];
        V3 c = add(b.x, qrot(b.q, s.at));
        lowest = minf_(lowest, c.z - s.rad);
        /* binary64 square root and accumulation, rounded, as in the reference */
        info.footprint = (float)((double)info.footprint + sqrt((double)(c.x * c.x + c.y * c.y)) / ROBOT_LEGS);
    }
    info.stand = -lowest;
    set_pose(w, v3(0.0f, 0.0f, info.stand), qid(), info.def);
}

JointK joint_constants(const HostJoint &j, float h)
{
    JointK k;
    k.r1 = j.r1; k.r2 = j.r2; k.a1 = j.a1; k.a2 = j.a2; k.b1 = j.b1; k.b2 = j.b2;
    k.qrel0 = j.qrel0;
    k.lo = j.lo;
    k.hi = j.hi;
    k.lim[0] = f_cos(j.lo);
    k.lim[1] = f_sin(j.lo);
    k.lim[2] = f_cos(j.hi);
    k.lim[3] = f_sin(j.hi);
    k.kp = j.kp;
    k.kd = j.kd;
    k.tmax = j.tmax;
    k.soft = j.kp > 0.0f ? (1.0f / j.kp) / (h * h) : 0.0f;
    k.cap_pos = j.tmax * h * h;
    k.tmax_h = j.tmax * h;
    k.wide = j.hi - j.lo > PI_F;
    k.has_limits = j.lo <= j.hi;
    return k;
}

} /* namespace xpbd */
