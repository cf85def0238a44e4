/*
 * arc_gpu.h -- additive API of libarcsim beside the reference's arc_sim.h:
 * error reporting, device-resident stepping, the task layer's fall rules with
 * auto-reset on the device, launch tuning and state read-back for tests.
 * None of these has a counterpart in the reference's C API; they exist for
 * callers that keep their inputs on the GPU (benchmarks, the step server,
 * PyTorch policies) and for the parity tests.
 */
#ifndef ARC_GPU_H
#define ARC_GPU_H

#include <stddef.h>
#include "arc_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sticky error state of a handle: 0 = none.  arcx_error_message returns a
 * static description of the first error since the last clear. */
int arcx_last_error(const arc_sim *s);
const char *arcx_error_message(const arc_sim *s);
void arcx_clear_error(arc_sim *s);

/* Number of HIP devices visible (0 when the library cannot run). */
int arcx_device_count(void);

/* Self-check of the engine's correctly rounded reciprocal on the current
 * GPU: every binary32 bit pattern x, the step's fast 1 / x against the IEEE
 * quotient; writes the number of differing results (0 expected).  Returns 0,
 * or -1 on a HIP error. */
int arcx_check_recip(unsigned long long *mismatches);

/* Self-check of the three-instruction quotient a / d the step kernels use for
 * a divisor fixed per sim (the column size) on the current GPU: every binary32
 * numerator a that is zero or has 2^-96 <= |a| < 2^100, against the IEEE
 * quotient; writes the number of differing results.  A sim enables the fast
 * quotient for a divisor only when this count is 0.  Returns 0, or -1 on a
 * HIP error. */
int arcx_check_quotient(float d, unsigned long long *mismatches);

/* Which of the sim's divisors take the checked fast quotient: bit 0 the
 * column size (the spheres' column windows). */
int arcx_fast_divisors(const arc_sim *s);

/* n successive outputs of the engine's xorshift128 generator seeded `seed`
 * (the stream the reset draws from; also used to make synthetic inputs). */
void arcx_rng_stream(unsigned int seed, long long n, unsigned int *out);

/* Plain device memory for inputs kept on the GPU (synthetic targets,
 * client buffers): hipMalloc / hipFree / synchronous copies.  to_device = 1
 * copies host -> device, 0 device -> host, 2 device -> device.  Returns NULL
 * / a hipError_t. */
void *arcx_device_alloc(size_t bytes);
int arcx_device_free(void *p);
int arcx_memcpy(void *dst, const void *src, size_t bytes, int to_device);

/* Size of the per-world flat state of arcx_state (same layout as the
 * oracle's dump): 13 bodies x (x3 q4 v3 w3), 12 joints x (target angle rate
 * torque), 21 spheres x (touch anchored anchor3 anchor_n3), unstable, time. */
int arcx_state_size(void);
/* State of world i (i = -1: the template) as float64. */
int arcx_state(const arc_sim *s, int i, double *out);
/* Terrain of world i: grid*grid heights then the outside height. */
int arcx_terrain(const arc_sim *s, int i, double *out);

/* Device-resident stepping: targets_dev is an N x 12 float32 device array
 * (world-major).  Asynchronous on the handle's stream; arcx_sync waits.
 * The entry points that take device addresses (these, arcx_time_steps,
 * arcx_step_task, arcx_step_batch, arcx_step_batch_rows) check that every
 * range the kernels will read lies inside one allocation of the sim's GPU
 * and return -1 with a sticky error otherwise, before any kernel could read
 * it. */
int arcx_step_device(arc_sim *s, const float *targets_dev);
int arcx_sync(arc_sim *s);

/* Allocation-free variant of arcx_step_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It steps every world exactly as arcx_step_device does (same
 * form, launch shape and distribution, same bits), but the step kernels'
 * scratch -- the per-substep contact lists and the persistent distribution's
 * work counter -- lives in one static device array per GPU, held from the
 * library's load until the process ends, instead of the sim's own buffers;
 * the call allocates and frees nothing (no hipMalloc, hipFree or allocating
 * library routine).  Limit: ARCX_NOALLOC_MAX_WORLDS worlds (the array holds
 * that many worlds' lists in every form); a larger sim is refused with -1
 * and a sticky error, nothing launched.  Every call on one GPU uses that
 * GPU's one array: calls that overlap in time on one GPU (two sims' streams,
 * or two host threads) race on it and must be serialised by the caller.
 * The array keeps the last call's contents; what must start at zero (the
 * work counter) is cleared by a kernel on the sim's stream before the step.
 * Asynchronous on the handle's stream, as arcx_step_device. */
#define ARCX_NOALLOC_MAX_WORLDS 8192
int arcx_step_device_noalloc(arc_sim *s, const float *targets_dev);

/* Benchmark helper: `steps` control steps back to back, step k reading
 * targets_dev + k * step_floats (the buffer must hold
 * (steps - 1) * step_floats + N * 12 floats; step_floats = 0 repeats one set);
 * kernel_ms = summed HIP-event time of the step kernels, wall_ms = event time
 * from before the first to after the last. */
int arcx_time_steps(arc_sim *s, const float *targets_dev, int steps, long long step_floats, double *kernel_ms,
                    double *wall_ms);

/* Task layer on the device (arc/task.py:180-203, arc/env.py:99-117).
 * arcx_step_task = one control step, then the fall rules in the task's
 * precedence (unstable or non-finite > void > flip > beached), then an
 * auto-reset of every fallen world with seed i + N * episode_i (episode_i
 * counts world i's auto-resets) and the difficulty of its last reset.
 * Asynchronous; the task's torso-contact counters persist across calls. */
int arcx_task_config(arc_sim *s, int beached_steps, double flip_up_z, double void_z);
int arcx_step_task(arc_sim *s, const float *targets_dev);
/* causes of the last arcx_step_task per world: 0 none, 1 unstable or
 * non-finite, 2 void, 3 flip, 4 beached (synchronous copy) */
int arcx_task_causes(const arc_sim *s, int *out);
/* terminations so far per cause (index 0 unused) */
int arcx_task_stats(const arc_sim *s, long long *out5);
int arcx_task_episodes(const arc_sim *s, unsigned int *out);

/* Form of the step kernel: 0 = one world per lane (all substeps of a
 * control step in one launch, world state in registers), 1 = the plain
 * form, one world per wavefront with the world staged in LDS, 2 = four lanes
 * per world (one leg each, torso replicated; shortest per-world latency and
 * the highest throughput untuned: the form of a new sim),
 * 3 = two lanes per world (two legs each), 4 = four lanes per world with the
 * four root joints of each block's worlds solved by the lanes of one wavefront
 * through LDS (blocks of 128 or 256 threads, no persistent distribution: form
 * 4 sets them, and arcx_set_launch / arcx_set_persistent refuse others while
 * it is on). */
int arcx_set_form(arc_sim *s, int form);

/* Launch shape of the step kernel: threads per block (a multiple of 64) and
 * blocks (0 = one pass over the worlds). */
int arcx_set_launch(arc_sim *s, int block, int grid_blocks);

/* HBM bytes the library holds per world (every form's buffers). */
double arcx_bytes_per_world(const arc_sim *s);

/* Contact-list overflow flags of world i (0 expected). */
int arcx_world_error(const arc_sim *s, int i);

/* ---- batched, persistent and graph form (register-resident forms) ---
 * No counterpart in the reference, whose arc_sim_step always steps every
 * world (csrc/arc_sim.c:123-138); these serve callers that step several
 * independent requests' worlds together. */

/* LDS layout of the lane form: tile = side of the per-world terrain tile held
 * in LDS (16), or 0 for substep-start poses only (terrain read from HBM). */
int arcx_set_layout(arc_sim *s, int tile);
/* 1: persistent distribution, a grid of resident workgroups whose
 * wavefronts take 64 worlds at a time from a device counter. */
int arcx_set_persistent(arc_sim *s, int on);
/* out5 = threads per block, tile, blocks (persistent: resident blocks),
 * persistent, form */
int arcx_get_launch(const arc_sim *s, int *out5);

/* One control step of the worlds [starts[r], starts[r] + counts[r]) of n_req
 * (<= 256) requests in one launch.  targets_dev: device float N x 12,
 * world-major (rows of worlds not in the batch are not read).  Ranges must be
 * inside [0, N) and disjoint (checked; -1 otherwise).  task = 1 adds the fall
 * rules and auto-reset of the stepped worlds (as arcx_step_task).  graph = 1
 * replays a captured hipGraph of the launch sequence: one graph per (task,
 * launch configuration), the last four kept, while the request table (ranges
 * and the address of each range's targets) is read from device memory at
 * every replay.  Asynchronous; arcx_sync waits. */
int arcx_step_batch(arc_sim *s, const int *starts, const int *counts, int n_req, const float *targets_dev, int task,
                    int graph);
/* As arcx_step_batch, with each request's targets in memory of its own:
 * request r's worlds read rows[r] + 12 k, k < counts[r] (float32 device
 * memory, e.g. the output tensor of a client's policy, read where it lies:
 * no copy).  The memory is read when the launch runs and must hold the
 * targets until the handle's stream has passed the step.  Same graphs as
 * arcx_step_batch; -1 if a non-empty request has no rows or its rows are not
 * counts[r] x 12 floats of device memory of the sim's GPU. */
int arcx_step_batch_rows(arc_sim *s, const int *starts, const int *counts, int n_req, const float *const *rows,
                         int task, int graph);
/* graphs captured so far by arcx_step_batch */
int arcx_graph_captures(const arc_sim *s);

/* Autotuner: times every (form: lane, leg, pair or shared; threads per
 * block; LDS layout; distribution) over `steps` control steps of this
 * handle's own worlds from their current state (restored afterwards), with
 * fresh uniform actions at every step, twice (all candidates in order, then
 * in reverse order, the two times averaged), keeps the fastest (a persistent
 * distribution only when at least 10 % faster than the best grid-stride one)
 * and reports it in out5 (as arcx_get_launch) and its ms per step in ms_out. */
int arcx_autotune(arc_sim *s, int steps, int *out5, double *ms_out);

#ifdef __cplusplus
}
#endif

#endif
