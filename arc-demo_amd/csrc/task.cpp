/*
 * task.cpp -- host side of the device task (include/arc_task.h): state
 * allocation, the launch sequence of one env step and its graph.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/arc_task.h"
#include "sim_internal.h"
#include "task_state.h"

struct arc_task {
    arc_sim *sim = nullptr;
    int n = 0, k = 0, size = 0;
    TaskCfg cfg{};
    TaskDev td{};
    TaskOut out{};
    void *mem = nullptr, *np_mem = nullptr;
    float *obs = nullptr, *term = nullptr, *actions = nullptr, *targets = nullptr;
    double *offsets = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const float *graph_actions = nullptr;
    int graph_key[6] = { 0, 0, 0, 0, 0, 0 };
    hipEvent_t ev_in = nullptr, ev_out = nullptr; /* ordering against a caller's stream */
};

extern "C" {

void arct_default_config(arct_config *c)
{
    c->max_steps = 1000;
    c->action_scale = 0.6;
    c->checkpoint_radius = 0.5;
    c->checkpoint_min_dist = 3.0;
    c->checkpoint_max_dist = 6.0;
    c->border = 1.5;
    c->w_progress = 10.0;
    c->max_speed = 1.2;
    c->fov_deg = 45.0;
    c->w_turn = 1.0;
    c->w_checkpoint = 10.0;
    c->w_time = 0.05;
    c->w_fall = 20.0;
    c->w_energy = 0.002;
    c->w_action_rate = 0.005;
    c->w_body_contact = 0.02;
    c->w_torso_contact = 0.5;
    c->w_body_rate = 0.01;
    c->w_bounce = 0.05;
    c->flip_up_z = 0.3;
    c->beached_steps = 30;
    c->void_z = -2.0;
    c->curriculum = 1;
    c->difficulty = 0.0;
    c->max_difficulty = 1.0;
    c->level_step = 0.1;
    c->level_up_checkpoints = 3;
    c->level_down_steps = 300;
}

void arct_destroy(arc_task *t)
{
    if (!t)
        return;
    if (t->sim && t->sim->stream)
        (void)hipStreamSynchronize(t->sim->stream);
    if (t->exec)
        (void)hipGraphExecDestroy(t->exec);
    if (t->graph)
        (void)hipGraphDestroy(t->graph);
    if (t->ev_in)
        (void)hipEventDestroy(t->ev_in);
    if (t->ev_out)
        (void)hipEventDestroy(t->ev_out);
    if (t->mem)
        (void)hipFree(t->mem);
    if (t->np_mem)
        (void)hipFree(t->np_mem);
    delete t;
}

arc_task *arct_create(arc_sim *s, const arct_config *c, const double *offsets2k, int k, unsigned int seed)
{
    if (!s || !c || k < 0 || (k && !offsets2k) || s->n <= 0)
        return nullptr;
    arc_task *t = new (std::nothrow) arc_task();
    if (!t)
        return nullptr;
    t->sim = s;
    t->n = s->n;
    t->k = k;
    t->size = 3 + 3 + 3 + 3 * NJ + 3 + k + LEGS;
    TaskCfg &g = t->cfg;
    g.dt = s->dt_d;
    g.action_scale = c->action_scale;
    g.max_steps = c->max_steps;
    g.checkpoint_radius = c->checkpoint_radius;
    g.checkpoint_min_dist = c->checkpoint_min_dist;
    g.checkpoint_max_dist = c->checkpoint_max_dist;
    g.border = c->border;
    g.w_progress = c->w_progress;
    g.max_speed = c->max_speed;
    /* numpy's radians: x * (pi / 180); the cosine on the host, once */
    g.fov_rad = c->fov_deg * (M_PI / 180.0);
    g.cos_fov = cos(g.fov_rad);
    g.w_turn = c->w_turn;
    g.w_checkpoint = c->w_checkpoint;
    g.w_time = c->w_time;
    g.w_fall = c->w_fall;
    g.w_energy = c->w_energy;
    g.w_action_rate = c->w_action_rate;
    g.w_body_contact = c->w_body_contact;
    g.w_torso_contact = c->w_torso_contact;
    g.w_body_rate = c->w_body_rate;
    g.w_bounce = c->w_bounce;
    g.flip_up_z = c->flip_up_z;
    g.void_z = c->void_z;
    g.beached_steps = c->beached_steps;
    g.curriculum = c->curriculum;
    g.max_difficulty = c->max_difficulty;
    g.level_step = c->level_step;
    g.level_up_checkpoints = c->level_up_checkpoints;
    g.level_down_steps = c->level_down_steps;
    g.half_extent = 0.5 * s->su.grid * s->cell_d;
    g.stand = (double)s->su.info.stand;
    for (int j = 0; j < NJ; ++j)
        g.def[j] = (double)s->su.info.def[j];

    const size_t S = s->dev.stride, N = (size_t)t->n;
    const size_t doubles = S * (2 + 1 + 1 + 1 + NJ + NJ + 1 + 1) + N + 2 * (size_t)(k > 0 ? k : 1);
    const size_t ints = S * (1 + 1 + 1 + 1 + 1 + 4 + 1) + 3 * N + 8;
    const size_t floats = S * 1 + 2 * N * t->size + N + 2 * N * NJ;
    const size_t bytes = doubles * 8 + ints * 4 + floats * 4 + 2 * N + 256;
    char *p = nullptr;
    if (hipMalloc((void **)&p, bytes) != hipSuccess || hipMemsetAsync(p, 0, bytes, s->stream) != hipSuccess) {
        fprintf(stderr, "libarcsim: arct_create: device allocation failed\n");
        if (p)
            (void)hipFree(p);
        delete t;
        return nullptr;
    }
    t->mem = p;
    double *dp = (double *)p;
    TaskDev &td = t->td;
    td.target = dp; dp += 2 * S;
    td.target_z = dp; dp += S;
    td.prev_dist = dp; dp += S;
    td.prev_gaze = dp; dp += S;
    td.action = dp; dp += NJ * S;
    td.prev_action = dp; dp += NJ * S;
    td.level = dp; dp += S;
    td.difficulty = dp; dp += S;
    t->out.difficulty = dp; dp += N;
    t->offsets = dp; dp += 2 * (size_t)(k > 0 ? k : 1);
    int *ip = (int *)dp;
    td.in_view = ip; ip += S;
    td.torso_steps = ip; ip += S;
    td.steps = ip; ip += S;
    td.checkpoints = ip; ip += S;
    td.list = ip; ip += S;
    td.rng = (uint32_t *)ip; ip += 4 * S;
    td.seeds = (uint32_t *)ip; ip += S;
    t->out.cause = ip; ip += N;
    t->out.checkpoints = ip; ip += N;
    t->out.steps = ip; ip += N;
    td.count = ip; ip += 8;
    float *fp = (float *)ip;
    td.diff = fp; fp += S;
    t->obs = fp; fp += N * t->size;
    t->term = fp; fp += N * t->size;
    t->out.reward = fp; fp += N;
    t->actions = fp; fp += N * NJ;
    t->targets = fp; fp += N * NJ;
    uint8_t *bp = (uint8_t *)fp;
    t->out.terminated = bp; bp += N;
    t->out.truncated = bp; bp += N;
    /* the numpy-generator mode's generator and lists */
    char *q = nullptr;
    if (hipMalloc((void **)&q, sizeof(NpStream) + 4 * N * sizeof(int)) != hipSuccess) {
        fprintf(stderr, "libarcsim: arct_create: device allocation failed\n");
        arct_destroy(t);
        return nullptr;
    }
    t->np_mem = q;
    td.np = (NpStream *)q;
    int *lp = (int *)(q + sizeof(NpStream));
    td.flag = lp; lp += N;
    td.hits = lp; lp += N;
    td.wa = lp; lp += N;
    td.wb = lp; lp += N;
    hipError_t e = hipSuccess;
    if (k)
        e = hipMemcpyAsync(t->offsets, offsets2k, 2 * (size_t)k * sizeof(double), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess)
        e = task_launch_init(s->dev, td, seed, c->curriculum ? 0.0 : c->difficulty, s->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        fprintf(stderr, "libarcsim: arct_create: %s\n", hipGetErrorString(e));
        arct_destroy(t);
        return nullptr;
    }
    return t;
}

int arct_obs_size(const arc_task *t) { return t->size; }

void *arct_buffer(arc_task *t, int which)
{
    switch (which) {
    case 0: return t->obs;
    case 1: return t->term;
    case 2: return t->out.reward;
    case 3: return t->out.terminated;
    case 4: return t->out.truncated;
    case 5: return t->out.cause; /* cause, checkpoints, steps are consecutive [N] arrays */
    case 6: return t->out.difficulty;
    case 7: return t->actions;
    default: return nullptr;
    }
}

/* reset the listed worlds and start their episodes */
static hipError_t reset_listed(arc_task *t)
{
    arc_sim *s = t->sim;
    const bool np = t->cfg.np_stream != 0;
    hipError_t e;
    if ((e = np ? task_launch_np_seeds(t->td, s->stream) : task_launch_reset_args(s->dev, t->td, s->stream))
            != hipSuccess
        || (e = launch_reset(s->dev, s->d_model, s->d_reset, t->td.list, t->td.seeds, nullptr, t->td.diff, 0,
                             t->td.count, s->reset_blocks, s->stream)) != hipSuccess
        || (e = task_launch_after_reset(s->dev, s->d_model, t->td, t->cfg, s->stream)) != hipSuccess)
        return e;
    if (np)
        return task_launch_np_place(s->dev, s->d_model, t->td, t->cfg, s->stream);
    return hipSuccess;
}

int arct_reset_all(arc_task *t)
{
    arc_sim *s = t->sim;
    HIPCHK_RET(s, task_launch_list_all(s->dev, t->td, s->stream), "arct_reset_all", -1);
    HIPCHK_RET(s, reset_listed(t), "arct_reset_all reset", -1);
    HIPCHK_RET(s, task_launch_observe(s->dev, s->d_model, t->td, t->cfg, t->offsets, t->k, t->obs, false, s->stream),
               "arct_reset_all observe", -1);
    return 0;
}

/* the kernels of one env step (arc/env.py:99-117 order) */
static hipError_t enqueue_step(arc_task *t, const float *actions)
{
    arc_sim *s = t->sim;
    hipError_t e;
    if ((e = launch_fill_int(t->td.count, 0, 1, s->stream)) != hipSuccess
        || (e = task_launch_actions(s->dev, t->td, t->cfg, actions, t->targets, s->stream)) != hipSuccess
        || (e = launch_form(s, t->targets)) != hipSuccess
        || (e = task_launch_post(s->dev, s->d_model, t->td, t->cfg, t->out, s->stream)) != hipSuccess
        || (t->cfg.np_stream && (e = task_launch_np_hits(s->dev, s->d_model, t->td, t->cfg, s->stream)) != hipSuccess)
        || (e = task_launch_observe(s->dev, s->d_model, t->td, t->cfg, t->offsets, t->k, t->obs, false, s->stream))
               != hipSuccess
        || (e = task_launch_copy_listed(s->dev, t->td, t->obs, t->term, t->size, s->stream)) != hipSuccess
        || (e = reset_listed(t)) != hipSuccess)
        return e;
    return task_launch_observe(s->dev, s->d_model, t->td, t->cfg, t->offsets, t->k, t->obs, true, s->stream);
}

int arct_step(arc_task *t, const float *actions_dev, int graph)
{
    arc_sim *s = t->sim;
    note_contacts(s);
    const float *a = actions_dev ? actions_dev : t->actions;
    if (actions_dev && (!graph || !t->exec || t->graph_actions != a) && !device_memory(s, actions_dev, (size_t)s->n * NJ * sizeof(float))) {
        set_error(s, -1, "arct_step: actions are not device memory of the sim's GPU", hipSuccess);
        return -1;
    }
    if (!graph) {
        HIPCHK_RET(s, enqueue_step(t, a), "arct_step", -1);
        return 0;
    }
    const int key[6] = { s->form, s->block, s->tile, s->persistent, s->grid_blocks, t->cfg.np_stream };
    if (!t->exec || t->graph_actions != a || memcmp(key, t->graph_key, sizeof(key))) {
        /* buffers the launch may allocate on first use are made before capture */
        HIPCHK_RET(s, prepare_launch(s), "arct_step prepare", -1);
        if (t->exec)
            (void)hipGraphExecDestroy(t->exec);
        if (t->graph)
            (void)hipGraphDestroy(t->graph);
        t->exec = nullptr;
        t->graph = nullptr;
        HIPCHK_RET(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "arct_step capture", -1);
        hipError_t e1 = enqueue_step(t, a);
        hipError_t e2 = hipStreamEndCapture(s->stream, &t->graph);
        HIPCHK_RET(s, e1, "arct_step capture", -1);
        HIPCHK_RET(s, e2, "arct_step capture", -1);
        HIPCHK_RET(s, hipGraphInstantiate(&t->exec, t->graph, nullptr, nullptr, 0), "arct_step instantiate", -1);
        t->graph_actions = a;
        memcpy(t->graph_key, key, sizeof(key));
    }
    HIPCHK_RET(s, hipGraphLaunch(t->exec, s->stream), "arct_step graph", -1);
    return 0;
}

int arct_sync(arc_task *t)
{
    arc_sim *s = t->sim;
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_sync", -1);
    return s->err;
}

/* device-side ordering with a caller's stream (events, no host wait) */
static hipError_t task_events(arc_task *t)
{
    hipError_t e = hipSuccess;
    if (!t->ev_in)
        e = hipEventCreateWithFlags(&t->ev_in, hipEventDisableTiming);
    if (e == hipSuccess && !t->ev_out)
        e = hipEventCreateWithFlags(&t->ev_out, hipEventDisableTiming);
    return e;
}

int arct_wait_stream(arc_task *t, void *stream)
{
    arc_sim *s = t->sim;
    HIPCHK_RET(s, task_events(t), "arct_wait_stream", -1);
    HIPCHK_RET(s, hipEventRecord(t->ev_in, (hipStream_t)stream), "arct_wait_stream", -1);
    HIPCHK_RET(s, hipStreamWaitEvent(s->stream, t->ev_in, 0), "arct_wait_stream", -1);
    return 0;
}

int arct_signal_stream(arc_task *t, void *stream)
{
    arc_sim *s = t->sim;
    HIPCHK_RET(s, task_events(t), "arct_signal_stream", -1);
    HIPCHK_RET(s, hipEventRecord(t->ev_out, s->stream), "arct_signal_stream", -1);
    HIPCHK_RET(s, hipStreamWaitEvent((hipStream_t)stream, t->ev_out, 0), "arct_signal_stream", -1);
    return 0;
}

int arct_targets(const arc_task *ct, double *out3)
{
    arc_task *t = const_cast<arc_task *>(ct);
    arc_sim *s = t->sim;
    const size_t S = s->dev.stride, N = (size_t)t->n;
    std::vector<double> buf(3 * S);
    HIPCHK_RET(s, hipMemcpyAsync(buf.data(), t->td.target, 2 * S * sizeof(double), hipMemcpyDeviceToHost, s->stream),
               "arct_targets", -1);
    HIPCHK_RET(s, hipMemcpyAsync(buf.data() + 2 * S, t->td.target_z, S * sizeof(double), hipMemcpyDeviceToHost,
                                 s->stream), "arct_targets", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_targets", -1);
    for (size_t i = 0; i < N; ++i) {
        out3[3 * i + 0] = buf[i];
        out3[3 * i + 1] = buf[S + i];
        out3[3 * i + 2] = buf[2 * S + i];
    }
    return 0;
}

int arct_seed_numpy(arc_task *t, const unsigned int *words, int n_words)
{
    arc_sim *s = t->sim;
    if (n_words < 0 || (n_words > 0 && !words)) {
        set_error(s, -1, "arct_seed_numpy: bad seed words", hipSuccess);
        return -1;
    }
    NpStream g;
    np_seed(g, words, n_words);
    HIPCHK_RET(s, hipMemcpyAsync(t->td.np, &g, sizeof(g), hipMemcpyHostToDevice, s->stream), "arct_seed_numpy", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_seed_numpy", -1);
    t->cfg.np_stream = 1;
    return 0;
}

int arct_numpy_state(const arc_task *ct, unsigned long long *out6)
{
    arc_task *t = const_cast<arc_task *>(ct);
    arc_sim *s = t->sim;
    if (!t->cfg.np_stream) {
        set_error(s, -1, "arct_numpy_state: the task does not draw from a numpy generator", hipSuccess);
        return -1;
    }
    NpStream g;
    HIPCHK_RET(s, hipMemcpyAsync(&g, t->td.np, offsetof(NpStream, pow2), hipMemcpyDeviceToHost, s->stream),
               "arct_numpy_state", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_numpy_state", -1);
    out6[0] = (unsigned long long)(g.state >> 64);
    out6[1] = (unsigned long long)g.state;
    out6[2] = (unsigned long long)(g.inc >> 64);
    out6[3] = (unsigned long long)g.inc;
    out6[4] = g.has_uint32;
    out6[5] = g.uinteger;
    return 0;
}

int arct_set_levels(arc_task *t, const double *levels)
{
    arc_sim *s = t->sim;
    for (int i = 0; i < t->n; ++i)
        if (levels[i] > MAX_DIFFICULTY) {
            set_error(s, -1, "arct_set_levels: level above 6, beyond the device terrain's range", hipSuccess);
            return -1;
        }
    HIPCHK_RET(s, hipMemcpyAsync(t->td.level, levels, (size_t)t->n * sizeof(double), hipMemcpyHostToDevice, s->stream),
               "arct_set_levels", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_set_levels", -1);
    return 0;
}

int arct_levels(const arc_task *ct, double *out)
{
    arc_task *t = const_cast<arc_task *>(ct);
    arc_sim *s = t->sim;
    HIPCHK_RET(s, hipMemcpyAsync(out, t->td.level, (size_t)t->n * sizeof(double), hipMemcpyDeviceToHost, s->stream),
               "arct_levels", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_levels", -1);
    return 0;
}

int arct_last_resets(const arc_task *ct, int *ids, unsigned int *seeds)
{
    arc_task *t = const_cast<arc_task *>(ct);
    arc_sim *s = t->sim;
    int count = 0;
    HIPCHK_RET(s, hipMemcpyAsync(&count, t->td.count, sizeof(int), hipMemcpyDeviceToHost, s->stream), "arct_last_resets",
               -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_last_resets", -1);
    if (count > 0) {
        if (ids)
            HIPCHK_RET(s, hipMemcpyAsync(ids, t->td.list, (size_t)count * sizeof(int), hipMemcpyDeviceToHost, s->stream),
                       "arct_last_resets", -1);
        if (seeds)
            HIPCHK_RET(s, hipMemcpyAsync(seeds, t->td.seeds, (size_t)count * sizeof(unsigned), hipMemcpyDeviceToHost,
                                         s->stream), "arct_last_resets", -1);
        HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arct_last_resets", -1);
    }
    return count;
}

} /* extern "C" */
