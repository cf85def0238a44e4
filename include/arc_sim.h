/*
 * arc_sim.h -- the batched simulation C API of libarcsim, a drop-in for the
 * reference's csrc/arc_sim.h behind its ctypes layer (arc/native.py:38-55).
 *
 * N independent worlds, each the quadruped on its own block terrain, stepped
 * on one GPU.  Arrays are the caller's, flat, world-major, float64 / int32 /
 * uint32, exactly as in the reference; the library converts at the boundary
 * (binary64 in, binary32 on the device, binary64 out).  Every call is
 * synchronous: when it returns, the next call sees its effect.  Calls on one
 * handle are not re-entrant; separate handles are independent.
 *
 * Differences from the reference, all at the edges of the contract:
 *   - arc_sim_create also returns NULL when no GPU is usable (with a message
 *     on stderr); the reference returns NULL only when out of memory;
 *   - out-of-range world ids in arc_sim_reset / arc_sim_terrain_at are
 *     skipped and reported (the reference does not check them);
 *   - arc_sim_reset skips and reports entries whose difficulty exceeds 6:
 *     the device stores terrain columns as int8 counts of 0.03 m, which
 *     holds every column up to that difficulty (the task uses [0, 1]);
 *   - arc_sim_world returns a snapshot copied from the device on each call
 *     (valid until the next call on the handle), not a live pointer, and it
 *     carries no render-only visuals;
 *   - arc_sim_set_threads is accepted and ignored: the device does the work.
 */
#ifndef ARC_SIM_H
#define ARC_SIM_H

#include "arc_world.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arc_sim arc_sim;

/* per-world block sizes of arc_sim_observe (csrc/arc_sim.h:18-21) */
#define ARC_SIM_ROOT_SIZE 13                        /* pos(3) quat wxyz(4) linvel(3) angvel(3) */
#define ARC_SIM_ROT_SIZE 9                          /* torso rotation, body to world, row-major */
#define ARC_SIM_JOINT_SIZE (3 * ARC_ROBOT_JOINTS)   /* angles, rates, torques */
#define ARC_SIM_CONTACT_SIZE (ARC_ROBOT_LEGS + 2)   /* feet, torso, other leg parts */

/* replaces csrc/arc_sim.h:25 (arc_sim.c:26-54): dt control period, substeps
 * XPBD substeps per period, grid x grid columns of `cell` metres (grid is
 * clamped to 64).  All worlds start as the robot at stand height on flat
 * ground.  NULL on failure. */
arc_sim *arc_sim_create(int n_worlds, double dt, int substeps, int grid, double cell);

/* replaces csrc/arc_sim.h:26; NULL-safe */
void arc_sim_destroy(arc_sim *s);

/* replaces csrc/arc_sim.h:27; kept for source compatibility, no effect */
void arc_sim_set_threads(arc_sim *s, int threads);

/* replaces csrc/arc_sim.h:28-30 */
int arc_sim_num_worlds(const arc_sim *s);
int arc_sim_num_joints(void);
int arc_sim_num_legs(void);

/* replaces csrc/arc_sim.h:33: default_angles, lo, hi: ARC_ROBOT_JOINTS each;
 * misc: stand height, total mass.  Any pointer may be NULL. */
void arc_sim_robot_info(const arc_sim *s, double *default_angles, double *lo, double *hi, double *misc);

/* replaces csrc/arc_sim.h:38-39 (arc_sim.c:88-121): regenerates the terrain
 * of the listed worlds (seed, difficulty in [0, 1]) and drops the robot at the
 * origin in its default stance with random yaw and joint noise.  A world
 * listed twice is reset once, from its last entry. */
void arc_sim_reset(arc_sim *s, const int *ids, int n_ids, const unsigned int *seeds, const double *difficulty);

/* replaces csrc/arc_sim.h:42 (arc_sim.c:123-138): targets N x 12 drive targets
 * (rad), clamped to the joint limits; unstable worlds are not stepped.  The
 * targets are read before the call returns; the step may still be running
 * on the GPU, and every function that reports state waits for it. */
void arc_sim_step(arc_sim *s, const double *targets);

/* replaces csrc/arc_sim.h:46-47 (arc_sim.c:140-183).  Any output may be NULL.
 * contacts: fraction of substeps in contact during the last step per channel;
 * flags: 1 for worlds that became unstable. */
void arc_sim_observe(const arc_sim *s, double *root, double *rot, double *joints, double *contacts, int *flags);

/* replaces csrc/arc_sim.h:51 (arc_sim.c:185-203): terrain heights at k
 * (forward, left) offsets from each torso in its heading frame; out N x k. */
void arc_sim_heights(const arc_sim *s, const double *offsets, int k, double *out);

/* replaces csrc/arc_sim.h:54 (arc_sim.c:205-210): height at m (x, y) points,
 * point i in world ids[i]. */
void arc_sim_terrain_at(const arc_sim *s, const int *ids, const double *xy, int m, double *out);

/* replaces csrc/arc_sim.h:56 (arc_sim.c:212-215): host snapshot of world i. */
const arc_world *arc_sim_world(const arc_sim *s, int i);

#ifdef __cplusplus
}
#endif

#endif
