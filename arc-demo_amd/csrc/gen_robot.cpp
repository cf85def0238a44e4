/*
 * gen_robot.cpp -- build-time generator of robot_consts.h.
 *
 * Builds the quadruped with the library's own binary32 builder (model.cpp)
 * and prints its geometry, mass and drive constants as exact hexadecimal
 * float literals, so the step kernel compiles them in as immediates instead of
 * loading them from memory every substep.  Only constants that do not depend
 * on dt, substeps or the terrain are emitted; sim_setup() checks at run time
 * that they equal the model it builds, bit for bit.
 *
 * Usage: gen_robot > robot_consts.h
 */
#include <stdio.h>
#include "model.h"

using namespace xpbd;

static void f(float v) { printf("%af", (double)v); }
static void v3p(V3 v)
{
    printf("{");
    f(v.x); printf(", "); f(v.y); printf(", "); f(v.z);
    printf("}");
}
static void q4p(Q4 q)
{
    printf("{");
    f(q.w); printf(", "); f(q.x); printf(", "); f(q.y); printf(", "); f(q.z);
    printf("}");
}

int main()
{
    HostWorld w;
    RobotInfo info;
    world_init(w);
    robot_build(w, info);
    printf("/* generated by gen_robot.cpp from the binary32 robot builder; do not edit */\n");
    printf("#pragma once\n#include \"xpbd_core.h\"\nnamespace xpbd {\nnamespace robot {\n");
    printf("constexpr int NBODY = %d, NJOINT = %d, NSPHERE = %d;\n", w.nb, w.nj, w.ns);
    printf("constexpr BodyK body[%d] = {\n", w.nb);
    for (int b = 0; b < w.nb; ++b) {
        const BodyK &k = w.bk[b];
        printf("    {"); f(k.mass); printf(", "); f(k.invm); printf(", ");
        v3p(k.I); printf(", "); v3p(k.invI);
        printf(", %d, %d},\n", k.iso_I, k.iso_invI);
    }
    printf("};\n");
    /* geometry and drive of each joint (the h-dependent members are zero) */
    printf("constexpr JointK joint[%d] = {\n", w.nj);
    for (int j = 0; j < w.nj; ++j) {
        JointK k = joint_constants(w.j[j], 1.0f);
        printf("    {");
        v3p(k.r1); printf(", "); v3p(k.r2); printf(", "); v3p(k.a1); printf(", "); v3p(k.a2); printf(", ");
        v3p(k.b1); printf(", "); v3p(k.b2); printf(", "); q4p(k.qrel0); printf(", ");
        f(k.lo); printf(", "); f(k.hi); printf(", {");
        for (int c = 0; c < 4; ++c) { f(k.lim[c]); printf(c < 3 ? ", " : "}, "); }
        f(k.kp); printf(", "); f(k.kd); printf(", "); f(k.tmax);
        printf(", 0.0f, 0.0f, 0.0f, %d, %d},\n", k.wide, k.has_limits);
    }
    printf("};\n");
    printf("constexpr SphereK sphere[%d] = {\n", w.ns);
    for (int s = 0; s < w.ns; ++s) {
        printf("    {"); v3p(w.s[s].at); printf(", "); f(w.s[s].rad); printf(", %d},\n", w.s[s].body);
    }
    printf("};\n} /* namespace robot */\n} /* namespace xpbd */\n");
    return 0;
}
