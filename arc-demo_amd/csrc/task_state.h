/*
 * task_state.h -- device state and launchers of the checkpoint-reaching task
 * (task_kernels.hip; arc/task.py and arc/env.py of the reference).
 */
#pragma once
#include "device_state.h"
#include "pcg64.h"

namespace xpbd {

/* the task's constants (arc/task.py:19-62), binary64, with the angles the
 * host derives once (radians, cosine) */
struct TaskCfg {
    double dt, action_scale;
    int max_steps;
    double checkpoint_radius, checkpoint_min_dist, checkpoint_max_dist, border;
    double w_progress, max_speed, fov_rad, cos_fov, w_turn, w_checkpoint, w_time, w_fall, w_energy, w_action_rate,
        w_body_contact, w_torso_contact, w_body_rate, w_bounce;
    double flip_up_z, void_z;
    int beached_steps;
    int curriculum;
    double max_difficulty, level_step;
    int level_up_checkpoints, level_down_steps;
    double half_extent, stand;
    double def[NJ];
    int np_stream; /* 1: draws from one numpy generator in the reference's order */
};

/* per-world task state, [field][stride] */
struct TaskDev {
    double *target;      /* [2] checkpoint x, y */
    double *target_z;
    double *prev_dist, *prev_gaze;
    int *in_view;
    double *action;      /* [12] this step's clipped actions */
    double *prev_action; /* [12] */
    int *torso_steps, *steps, *checkpoints;
    double *level, *difficulty;
    uint32_t *rng;       /* [4] checkpoint placement and reset seeds */
    int *list;           /* worlds to reset */
    uint32_t *seeds;
    float *diff;
    int *count;
    /* numpy-generator mode: per-world hit / ended flags, the ordered hit list,
     * the pending lists of checkpoint placement, the generator */
    int *flag, *hits, *wa, *wb;
    NpStream *np;
};

/* per-step outputs */
struct TaskOut {
    float *reward;
    uint8_t *terminated, *truncated;
    int *cause, *checkpoints, *steps;
    double *difficulty;
};

hipError_t task_launch_init(const DevWorlds &d, const TaskDev &td, uint32_t seed, double level, hipStream_t s);
hipError_t task_launch_list_all(const DevWorlds &d, const TaskDev &td, hipStream_t s);
hipError_t task_launch_actions(const DevWorlds &d, const TaskDev &td, const TaskCfg &c, const float *actions,
                               float *targets, hipStream_t s);
hipError_t task_launch_post(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c, const TaskOut &o,
                            hipStream_t s);
hipError_t task_launch_copy_listed(const DevWorlds &d, const TaskDev &td, const float *obs, float *dst, int size,
                                   hipStream_t s);
hipError_t task_launch_reset_args(const DevWorlds &d, const TaskDev &td, hipStream_t s);
hipError_t task_launch_after_reset(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                                   hipStream_t s);
hipError_t task_launch_np_hits(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                               hipStream_t s);
hipError_t task_launch_np_seeds(const TaskDev &td, hipStream_t s);
hipError_t task_launch_np_place(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                                hipStream_t s);
hipError_t task_launch_observe(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                               const double *offsets, int k_off, float *obs, bool listed, hipStream_t s);

} /* namespace xpbd */
