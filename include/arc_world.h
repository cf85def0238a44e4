/*
 * arc_world.h -- host data types of the arc_sim C API.
 *
 * These describe the memory layout that arc_sim_world() returns (a host
 * snapshot of one device world), which the reference's renderer and viewer
 * read field by field (csrc/arc_render.c:1037-1046, csrc/tools/arcview.c:43-65).
 * The layout therefore has to equal the reference's: csrc/arc_math.h:20-22
 * (vec3, quat, mat3), csrc/arc_phys.h:19-23 (capacities) and 25-112 (body,
 * sphere, joint, contact, visual, world), csrc/arc_terrain.h:10-19 (terrain),
 * csrc/arc_robot.h:18-19 (robot sizes).  sizeof(arc_world) == 62960.
 */
#ifndef ARC_WORLD_H
#define ARC_WORLD_H

#ifdef __cplusplus
extern "C" {
#endif

#define ARC_MAX_BODIES 16
#define ARC_MAX_JOINTS 16
#define ARC_MAX_SPHERES 48
#define ARC_MAX_CONTACTS 128
#define ARC_MAX_VISUALS 80
#define ARC_GRID_MAX 64
#define ARC_TERRAIN_VOID -100.0

#define ARC_ROBOT_LEGS 4
#define ARC_ROBOT_JOINTS 12

typedef struct { double x, y, z; } vec3;
typedef struct { double w, x, y, z; } quat;
typedef struct { double m[3][3]; } mat3; /* row-major */

typedef struct {
    vec3 x, x_prev;
    quat q, q_prev;
    vec3 v, w;
    double mass, inv_mass;
    vec3 inertia, inv_inertia;
    mat3 rot;
    mat3 inv_iw;
} arc_body;

typedef struct {
    int body;
    vec3 pos;
    double radius;
    int touch;
    int anchored;
    vec3 anchor;
    vec3 anchor_n;
} arc_sphere;

typedef struct {
    int parent;
    int child;
    vec3 r1, r2;
    vec3 a1, a2;
    vec3 b1, b2;
    quat q_rel0;
    double lo, hi;
    double kp, kd;
    double max_torque;
    double target;
    double angle, rate, torque;
    double torque_acc;
    double lim[4];
} arc_joint;

typedef struct {
    int body;
    vec3 local;
    vec3 n;
} arc_contact;

/* visual shapes and surfaces (csrc/arc_phys.h:69-87) */
enum { ARC_VIS_SPHERE, ARC_VIS_BOX, ARC_VIS_TAPER, ARC_VIS_CYLINDER, ARC_VIS_RING };
enum { ARC_MAT_PAINT, ARC_MAT_PANELS, ARC_MAT_METAL, ARC_MAT_RUBBER, ARC_MAT_LENS, ARC_MAT_RIBS, ARC_MAT_GLOW };

typedef struct {
    int type;
    int body;
    vec3 p0, p1;
    double radius, radius2;
    double aspect;
    int material;
    float color[3];
} arc_visual;

typedef struct {
    int nx, ny;
    double cell;
    double ox, oy;
    double outside;
    double h[ARC_GRID_MAX * ARC_GRID_MAX];
} arc_terrain;

typedef struct {
    vec3 gravity;
    double mu_static, mu_dynamic;
    int n_bodies, n_joints, n_spheres, n_contacts, n_visuals;
    int unstable;
    double time;
    arc_body bodies[ARC_MAX_BODIES];
    arc_joint joints[ARC_MAX_JOINTS];
    arc_sphere spheres[ARC_MAX_SPHERES];
    arc_contact contacts[ARC_MAX_CONTACTS];
    arc_visual visuals[ARC_MAX_VISUALS];
    arc_terrain terrain;
} arc_world;

#ifdef __cplusplus
}
#endif

#endif
