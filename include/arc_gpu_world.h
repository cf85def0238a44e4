/*
 * arc_gpu_world.h -- one hand-built world stepped on the GPU: the
 * reference's single-world physics API (csrc/arc_phys.h:114-144,
 * csrc/arc_terrain.h:22-50) behind an opaque handle.
 *
 * The reference's arc_world is a struct its callers build with arc_add_body /
 * arc_add_sphere / arc_add_hinge and then read and write field by field
 * (bodies[i].v, joints[j].target, terrain.h[c]); here the same fields are
 * reached through getters and setters, and arcx_world_step runs
 * arc_world_step on the device: the plain form's substep code (one
 * wavefront, the world in LDS) on the world's own joint targets, in binary32,
 * giving the bits of the reference compiled with double as float.
 *
 * Capacities are the reference's (16 bodies, 16 hinges, 48 spheres, a grid of
 * at most 64 x 64 columns).  Device safety adds two bounds the reference does
 * not have: a terrain cell of 1e-3 to 1e3 m, and spheres whose cell window
 * spans at most 130 columns per axis (radius below about 64 cells); a world
 * outside them is refused by arcx_world_step.
 *
 * Functions returning int return 0 (or the new index) on success and -1 on
 * an invalid argument or a full world (as the reference's builders return
 * -1), -2 on a device error; arcx_world_message describes the last failure.
 * A handle belongs to the GPU current at arcx_world_create; calls on one
 * handle must not overlap.
 */
#ifndef ARC_GPU_WORLD_H
#define ARC_GPU_WORLD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arcx_world arcx_world;

/* arc_world_init (csrc/arc_phys.c:32-39): gravity (0, 0, -9.81), mu 1.0 /
 * 0.8, a 1 x 1 flat terrain of 1 m cells at height 0.  NULL on failure. */
arcx_world *arcx_world_create(void);
void arcx_world_destroy(arcx_world *w);
const char *arcx_world_message(const arcx_world *w);

/* w->gravity and w->mu_static / w->mu_dynamic (csrc/arc_phys.h:101-102) */
int arcx_world_set_gravity(arcx_world *w, double x, double y, double z);
int arcx_world_set_mu(arcx_world *w, double mu_static, double mu_dynamic);

/* arc_terrain_flat (csrc/arc_terrain.c:17-30), grids above 64 cut to 64 */
int arcx_world_terrain_flat(arcx_world *w, int nx, int ny, double cell, double height);
/* arc_terrain_generate (csrc/arc_terrain.c:54-71) drawing from an arc_rng
 * seeded `seed` (csrc/arc_rng.h:35) */
int arcx_world_terrain_generate(arcx_world *w, int n, double cell, double difficulty, unsigned int seed);
/* w->terrain.h[i] = height */
int arcx_world_terrain_set(arcx_world *w, int i, double height);
/* arc_terrain_height (csrc/arc_terrain.h:42-50) */
double arcx_world_terrain_height(const arcx_world *w, double x, double y);
/* nx, ny, cell, ox, oy, outside */
int arcx_world_terrain_info(const arcx_world *w, double out6[6]);

/* arc_add_body, arc_add_sphere, arc_add_hinge (csrc/arc_phys.c:84-143):
 * the new index, or -1 */
int arcx_world_add_body(arcx_world *w, double mass, const double inertia3[3], const double pos3[3],
                        const double q4[4]);
int arcx_world_add_sphere(arcx_world *w, int body, const double pos3[3], double radius);
int arcx_world_add_hinge(arcx_world *w, int parent, int child, const double anchor3[3], const double axis3[3]);

/* joints[j].lo, hi, kp, kd, max_torque, target (csrc/arc_phys.h:53-56) */
int arcx_world_joint_set(arcx_world *w, int j, const double p6[6]);
/* joints[j].angle, rate, torque after the last step */
int arcx_world_joint_out(const arcx_world *w, int j, double out3[3]);
/* arc_joint_angle / arc_joint_rate of the current state (csrc/arc_phys.c:386-400) */
double arcx_world_joint_angle(const arcx_world *w, int j);
double arcx_world_joint_rate(const arcx_world *w, int j);

/* bodies[i].x (3), q (4), v (3), w (3) */
int arcx_world_body_get(const arcx_world *w, int i, double out13[13]);
int arcx_world_body_set(arcx_world *w, int i, const double in13[13]);
/* spheres[i].touch, anchored, anchor (3), anchor_n (3) */
int arcx_world_sphere_get(const arcx_world *w, int i, double out8[8]);

/* arc_world_set_pose (csrc/arc_phys.c:749-782): angles holds one per hinge */
int arcx_world_set_pose(arcx_world *w, const double pos3[3], const double q4[4], const double *angles);

/* `steps` calls of arc_world_step(w, dt, substeps) (csrc/arc_phys.c:707-747)
 * on the GPU; synchronous */
int arcx_world_step(arcx_world *w, double dt, int substeps, int steps);

/* arc_world_energy (csrc/arc_phys.c:784-796), in binary32 in the
 * reference's order */
double arcx_world_energy(const arcx_world *w);
int arcx_world_unstable(const arcx_world *w);
double arcx_world_time(const arcx_world *w);
/* n_bodies, n_joints, n_spheres */
int arcx_world_counts(const arcx_world *w, int out3[3]);

#ifdef __cplusplus
}
#endif

#endif /* ARC_GPU_WORLD_H */
