/*
 * model.h -- host-side description of a world in binary32: body / joint /
 * sphere constants shared by every world of a simulation, plus the template
 * state the robot builder leaves behind.  Built once per simulation on the
 * host (arc_sim_create), then copied to device constant memory.
 */
#pragma once
#include <math.h>
#include <vector>
#include "xpbd_core.h"

namespace xpbd {

enum { MAX_BODIES = 16, MAX_JOINTS = 16, MAX_SPHERES = 48, MAX_CONTACTS = 128, GRID_MAX = 64 };
enum { ROBOT_LEGS = 4, ROBOT_JOINTS = 12, ROBOT_BODIES = 13, ROBOT_SPHERES = 21 };

struct HostJoint {
    int parent, child;
    V3 r1, r2, a1, a2, b1, b2;
    Q4 qrel0;
    float lo, hi, kp, kd, tmax, target;
    float angle, rate, torque;
};

struct HostSphere {
    int body;
    V3 at;
    float rad;
    int touch, anchored;
    V3 anc, ancn;
};

/* One world in binary32 on the host (the builder's workspace and the
 * generic single-world path's exchange format). */
struct HostWorld {
    V3 g;
    float mus, mud;
    int nb, nj, ns;
    BodyK bk[MAX_BODIES];
    Body b[MAX_BODIES];
    HostJoint j[MAX_JOINTS];
    HostSphere s[MAX_SPHERES];
    int unstable;
    float time;
    TerrainK t;
    float outside;
    std::vector<float> h;
};

struct RobotInfo {
    int foot[ROBOT_LEGS];
    int n_torso_spheres;
    float def[ROBOT_JOINTS], lo[ROBOT_JOINTS], hi[ROBOT_JOINTS];
    float stand, mass, footprint;
};

void world_init(HostWorld &w);
void terrain_flat(HostWorld &w, int nx, int ny, float cell, float height);
int add_body(HostWorld &w, float mass, V3 inertia, V3 pos, Q4 q);
int add_sphere(HostWorld &w, int body, V3 at, float rad);
int add_hinge(HostWorld &w, int parent, int child, V3 anchor, V3 axis);
void set_pose(HostWorld &w, V3 pos, Q4 q, const float *angles);
V3 inertia_sphere(float m, float r);
V3 inertia_capsule_z(float m, float r, float half);
V3 inertia_box(float m, V3 he);
void robot_build(HostWorld &w, RobotInfo &info);

/* worst-case columns a sphere of radius r overlaps: per axis at most
 * floor(2r / cell) + 2 indices (one more than the span, one for rounding) */
inline int window_cells(float r, float cell)
{
    int k = (int)floor((double)(2.0f * r) / (double)cell) + 2;
    return k * k;
}

/* Joint constants for a substep length h (derived products rounded in the
 * reference's evaluation order). */
JointK joint_constants(const HostJoint &j, float h);

} /* namespace xpbd */
