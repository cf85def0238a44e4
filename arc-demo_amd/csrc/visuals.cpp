/*
 * visuals.cpp -- the quadruped's render geometry for the host snapshot
 * (arc_sim_world), in the reference's arc_visual layout.
 *
 * The solver never reads these; the reference's renderer draws the robot
 * from them (csrc/arc_render.c:505-511).  The shapes, sizes, materials and
 * colours are the robot's look as the reference's builder defines it
 * (csrc/arc_robot.c:38-61 colours and sizes, :111-142 per leg, :155-170 the
 * torso), computed here in binary64 with the reference's vector formulas
 * (csrc/arc_math.h: scale, normalize by 1 / |v|), bodies numbered torso 0,
 * then per leg coxa, femur, tibia.
 */
#include <math.h>

#include "../../include/arc_world.h"
#include "visuals.h"

namespace {

struct Geo {
    double r_body = 0.2, leg_angle = 0.785, femur_len = 0.50, tibia_len = 0.95, foot_r = 0.045;
};

const float C_ARMOR[3] = { 0.40f, 0.41f, 0.43f };
const float C_HULL[3] = { 0.34f, 0.345f, 0.36f };
const float C_DARK[3] = { 0.17f, 0.17f, 0.18f };
const float C_STEEL[3] = { 0.52f, 0.53f, 0.55f };
const float C_BEZEL[3] = { 0.72f, 0.73f, 0.74f };
const float C_CABLE[3] = { 0.30f, 0.30f, 0.31f };
const float C_BLACK[3] = { 0.06f, 0.06f, 0.065f };
const float C_GLOW[3] = { 1.00f, 0.10f, 0.06f };
const float C_DIM[3] = { 0.62f, 0.05f, 0.03f };

vec3 mk(double x, double y, double z)
{
    vec3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
vec3 times(vec3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
vec3 unit_of(vec3 a)
{
    const double l = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 1e-15 ? times(a, 1.0 / l) : mk(0.0, 0.0, 0.0);
}

/* appends visuals in order; aspect 1 unless given */
struct Out {
    arc_visual *v;
    int n = 0, cap;
    void add(int type, int material, int body, vec3 p0, vec3 p1, double r, double r2, const float *c,
             double aspect = 1.0)
    {
        if (n >= cap)
            return;
        arc_visual &a = v[n++];
        a.type = type;
        a.body = body;
        a.p0 = p0;
        a.p1 = p1;
        a.radius = r;
        a.radius2 = r2;
        a.aspect = aspect;
        a.material = material;
        for (int k = 0; k < 3; ++k)
            a.color[k] = c[k];
    }
};

void torso(Out &o)
{
    const double R = Geo().r_body;
    o.add(ARC_VIS_SPHERE, ARC_MAT_PANELS, 0, mk(0, 0, 0), mk(0, 0, 0), R, 0.0, C_HULL);
    o.add(ARC_VIS_RING, ARC_MAT_METAL, 0, mk(0, 0, 0.135), mk(0, 0, 1), 0.148, 0.02, C_DARK);
    o.add(ARC_VIS_RING, ARC_MAT_PAINT, 0, mk(0, 0, -0.125), mk(0, 0, 1), 0.156, 0.024, C_DARK);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_METAL, 0, mk(0, 0, -0.2), mk(0, 0, -0.17), 0.07, 0.0, C_DARK);
    for (int side = -1; side <= 1; side += 2) {
        const vec3 k = unit_of(mk(0.45, 0.62 * side, -0.64));
        o.add(ARC_VIS_CYLINDER, ARC_MAT_METAL, 0, times(k, 0.17), times(k, 0.215), 0.028, 0.0, C_STEEL);
    }
    o.add(ARC_VIS_RING, ARC_MAT_GLOW, 0, mk(0.18, 0, 0), mk(1, 0, 0), 0.08, 0.013, C_GLOW);
    o.add(ARC_VIS_RING, ARC_MAT_METAL, 0, mk(0.168, 0, 0), mk(1, 0, 0), 0.112, 0.02, C_BEZEL);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_LENS, 0, mk(0.11, 0, 0), mk(0.178, 0, 0), 0.095, 0.0, C_BLACK);
    o.add(ARC_VIS_RING, ARC_MAT_GLOW, 0, mk(0.181, 0, 0), mk(1, 0, 0), 0.032, 0.006, C_DIM);
}

/* leg l (front left, front right, rear left, rear right: x side sx, y side
 * sy) on bodies coxa bc, femur bf, tibia bt */
void leg(Out &o, int l)
{
    const Geo G;
    const double sx = l < 2 ? 1.0 : -1.0, sy = (l % 2 == 0) ? 1.0 : -1.0;
    const double pi = 3.14159265358979323846;
    const double psi = sx > 0.0 ? sy * G.leg_angle : sy * (pi - G.leg_angle);
    const vec3 d = mk(cos(psi), sin(psi), 0.0);
    const double fl = 0.5 * G.femur_len, tl = 0.5 * G.tibia_len;
    const int bc = 1 + 3 * l, bf = bc + 1, bt = bc + 2;
    o.add(ARC_VIS_CYLINDER, ARC_MAT_METAL, bc, times(d, -0.05), times(d, 0.035), 0.062, 0.0, C_DARK);
    o.add(ARC_VIS_RING, ARC_MAT_METAL, bc, times(d, 0.02), d, 0.062, 0.012, C_STEEL);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_METAL, bf, mk(0, -0.07, -fl), mk(0, 0.07, -fl), 0.045, 0.0, C_DARK);
    /* femur: armour plate over a cable sleeve */
    o.add(ARC_VIS_TAPER, ARC_MAT_PAINT, bf, mk(-0.012, 0, -fl + 0.02), mk(-0.012, 0, fl - 0.02), 0.055, 0.08,
          C_ARMOR, 0.8);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_RIBS, bf, mk(0.055, 0, -fl + 0.06), mk(0.07, 0, fl - 0.07), 0.022, 0.0, C_CABLE);
    /* knee: housing, axle and two lights */
    o.add(ARC_VIS_BOX, ARC_MAT_PAINT, bf, mk(0, 0, fl), mk(0.085, 0.07, 0.075), 0.0, 0.0, C_DARK);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_METAL, bf, mk(0, -0.085, fl), mk(0, 0.085, fl), 0.048, 0.0, C_STEEL);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_GLOW, bf, mk(0, 0.085, fl), mk(0, 0.089, fl), 0.02, 0.0, C_GLOW);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_GLOW, bf, mk(0, -0.089, fl), mk(0, -0.085, fl), 0.02, 0.0, C_GLOW);
    /* tibia: two plates meeting in a ridge, a cable sleeve */
    o.add(ARC_VIS_TAPER, ARC_MAT_PAINT, bt, mk(-0.01, 0, -tl + 0.03), mk(-0.01, 0, -0.05), 0.068, 0.082, C_ARMOR,
          0.8);
    o.add(ARC_VIS_TAPER, ARC_MAT_PAINT, bt, mk(-0.01, 0, -0.05), mk(-0.004, 0, tl - 0.11), 0.082, 0.03, C_ARMOR,
          0.85);
    o.add(ARC_VIS_CYLINDER, ARC_MAT_RIBS, bt, mk(0.065, 0, -tl + 0.08), mk(0.045, 0, 0.25 * tl), 0.02, 0.0, C_CABLE);
    /* ankle and foot */
    o.add(ARC_VIS_CYLINDER, ARC_MAT_METAL, bt, mk(0, 0, tl - 0.12), mk(0, 0, tl - 0.035), 0.026, 0.0, C_STEEL);
    o.add(ARC_VIS_RING, ARC_MAT_METAL, bt, mk(0, 0, tl - 0.09), mk(0, 0, 1), 0.03, 0.01, C_DARK);
    o.add(ARC_VIS_SPHERE, ARC_MAT_RUBBER, bt, mk(0, 0, tl), mk(0, 0, 0), G.foot_r, 0.0, C_BLACK);
    o.add(ARC_VIS_RING, ARC_MAT_METAL, bt, mk(0, 0, tl - 0.02), mk(0, 0, 1), 0.036, 0.011, C_STEEL);
}

} /* namespace */

int robot_visuals(arc_visual *out, int cap)
{
    Out o;
    o.v = out;
    o.cap = cap;
    torso(o);
    for (int l = 0; l < ARC_ROBOT_LEGS; ++l)
        leg(o, l);
    return o.n;
}
