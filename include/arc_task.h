/*
 * arc_task.h -- the checkpoint-reaching task of the reference (arc/task.py,
 * arc/env.py) run on the device over an arc_sim's worlds.
 *
 * The reference computes observations, rewards, terminations, checkpoint
 * placement and the terrain curriculum in numpy over all worlds after every
 * arc_sim_step (arc/task.py:155-280), then the vectorised env resets the
 * ended worlds and re-observes them (arc/env.py:99-117).  Here one call does
 * all of it on the GPU, with the actions read from and the observations
 * written to device memory, so a policy on the same GPU exchanges them
 * without host copies.  Checkpoint placement and reset seeds come either
 * from a per-world stream of the engine's xorshift128 generator (the
 * default: independent of which worlds end together), or, after
 * arct_seed_numpy, from one stream of numpy's default generator consumed in
 * the reference's order, so that they equal the Python task's with the same
 * seed (arc/task.py:93,117,136-137).
 */
#ifndef ARC_TASK_H
#define ARC_TASK_H

#include "arc_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arc_task arc_task;

/* TaskConfig of arc/task.py:19-62 (the simulation fields come from the sim) */
typedef struct arct_config {
    int max_steps;
    double action_scale;
    double checkpoint_radius, checkpoint_min_dist, checkpoint_max_dist, border;
    double w_progress, max_speed, fov_deg, w_turn, w_checkpoint, w_time, w_fall, w_energy, w_action_rate,
        w_body_contact, w_torso_contact, w_body_rate, w_bounce;
    double flip_up_z, void_z;
    int beached_steps;
    int curriculum;
    double difficulty, max_difficulty, level_step;
    int level_up_checkpoints, level_down_steps;
} arct_config;

/* the reference's defaults */
void arct_default_config(arct_config *c);

/* A task over every world of s (which it steps; s must outlive it).
 * offsets2k: the k (forward, left) height samples around the torso (the
 * reference's 13 x 11 grid, arc/task.py:73-76).  seed: the per-world streams.
 * NULL on failure. */
arc_task *arct_create(arc_sim *s, const arct_config *c, const double *offsets2k, int k, unsigned int seed);
void arct_destroy(arc_task *t);

/* floats per observation: 3 + 3 + 3 + 3 x 12 + 3 + k + 4 (195 for k = 143) */
int arct_obs_size(const arc_task *t);

/* Device buffers owned by the task (valid until arct_destroy):
 *   0 observations        float   [N][obs]
 *   1 terminal obs        float   [N][obs]  rows of worlds that ended in the last step
 *   2 rewards             float   [N]
 *   3 terminated          uint8   [N]
 *   4 truncated           uint8   [N]
 *   5 info                int32   [3][N]    cause (0 none, 1 unstable, 2 void, 3 flip, 4 beached),
 *                                           checkpoints, steps of the episode that just ran
 *   6 episode difficulty  float64 [N]
 *   7 actions             float   [N][12]   the default input of arct_step */
void *arct_buffer(arc_task *t, int which);

/* Resets every world (seeds from the streams, difficulty = level) and
 * writes the observations.  Asynchronous on the sim's stream. */
int arct_reset_all(arc_task *t);
/* One env step: actions (device float [N][12]; NULL = buffer 7) -> drive
 * targets, the step, rewards / terminations / checkpoints / curriculum,
 * observations, terminal observations of the ended worlds, their reset and
 * their new first observations.  graph = 1 replays a captured hipGraph of
 * the sequence (captured again when actions or the sim's launch
 * configuration change).  Asynchronous; arct_sync waits. */
int arct_step(arc_task *t, const float *actions_dev, int graph);
int arct_sync(arc_task *t);

/* Ordering with a caller's HIP stream of the same device (for example
 * torch.cuda.current_stream().cuda_stream) on the device, without a host
 * wait: after arct_wait_stream, the task's next work starts once the work
 * queued on `stream` so far has finished (the policy that wrote the
 * actions); after arct_signal_stream, work queued on `stream` later starts
 * once the task's queued work has finished (readers of the observations).
 * No counterpart in the reference, whose task runs on the host. */
int arct_wait_stream(arc_task *t, void *stream);
int arct_signal_stream(arc_task *t, void *stream);

/* Draw checkpoints and reset seeds from np.random.default_rng(seed), as the
 * reference's ArcTask (arc/task.py:93) and ArcVecEnv.reset (arc/env.py:93-95)
 * seed it; `words` is the seed as little-endian 32-bit words (numpy's
 * SeedSequence entropy; one word 0 for n_words = 0).  The draws follow the
 * reference's order: per step, new checkpoints of the worlds that reached
 * theirs (in world order: rounds of angles then distances, arc/task.py:127-150),
 * then the seeds of the ended worlds (rng.integers(0, 2**32, n, uint32),
 * arc/task.py:117), then their first checkpoints.  Takes effect from the next
 * arct_reset_all / arct_step. */
int arct_seed_numpy(arc_task *t, const unsigned int *words, int n_words);
/* the generator's state as numpy's bit_generator.state holds it: state (hi,
 * lo 64-bit words), inc (hi, lo), has_uint32, uinteger; -1 unless seeded by
 * arct_seed_numpy */
int arct_numpy_state(const arc_task *t, unsigned long long *out6);

/* curriculum levels of every world (N doubles, host): the difficulty of each
 * world's next episode (with curriculum off, its fixed difficulty); each at
 * most 6, the device terrain's range (-1 and nothing set otherwise) */
int arct_set_levels(arc_task *t, const double *levels);

/* host read-backs: checkpoints (N x 3: x, y, terrain z), curriculum levels
 * (N), and the worlds and seeds of the last resets (returns their count) */
int arct_targets(const arc_task *t, double *out3);
int arct_levels(const arc_task *t, double *out);
int arct_last_resets(const arc_task *t, int *ids, unsigned int *seeds);

#ifdef __cplusplus
}
#endif

#endif
