/*
 * batch.cpp -- the batched form of libarcsim: one launch steps the worlds of
 * several requests (uneven world ranges), optionally as a persistent grid
 * that takes work from a device counter, optionally replayed from a captured
 * hipGraph; and the autotuner that picks the launch configuration (workgroup
 * size, LDS layout, work distribution) by timing this handle's own worlds.
 *
 * The request table lives in device memory (StepBatch) and is written before
 * each launch, so one captured graph serves every batch: the graph holds the
 * kernels, the table holds what they step.
 */
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <utility>
#include <vector>

#include "noalloc.h"
#include "sim_internal.h"

namespace {

constexpr int RING = 4;   /* pinned request tables in flight */
constexpr int GRAPHS = 4; /* captured graphs kept per handle */

/* a captured launch sequence: the step kernels read every range's targets
 * through the table's row pointers, so no buffer address is part of it */
struct LaunchKey {
    int task, block, tile, grid, persistent, leg;
    bool operator==(const LaunchKey &o) const
    {
        return task == o.task && block == o.block && tile == o.tile && grid == o.grid && persistent == o.persistent
               && leg == o.leg;
    }
};

} /* namespace */

struct BatchState {
    StepBatch *d_batch = nullptr;
    StepBatch *h_batch[RING] = {};
    hipEvent_t ev[RING] = {};
    int slot = 0;
    int *d_work = nullptr;
    /* captured graphs, one per launch key, least recently used replaced */
    struct Entry {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        LaunchKey key{};
        long long used = 0;
    } cache[GRAPHS];
    long long clock = 0;
    int captures = 0;
    StepBatch last;          /* the table last uploaded */
    size_t last_bytes = 0;
};

void batch_free(arc_sim *s)
{
    BatchState *b = s->batch;
    if (!b)
        return;
    for (auto &c : b->cache) {
        if (c.exec)
            (void)hipGraphExecDestroy(c.exec);
        if (c.graph)
            (void)hipGraphDestroy(c.graph);
    }
    for (int k = 0; k < RING; ++k) {
        if (b->ev[k])
            (void)hipEventDestroy(b->ev[k]);
        if (b->h_batch[k])
            (void)hipHostFree(b->h_batch[k]);
    }
    if (b->d_batch)
        (void)hipFree(b->d_batch);
    if (b->d_work)
        (void)hipFree(b->d_work);
    delete b;
    s->batch = nullptr;
}

static BatchState *batch_of(arc_sim *s)
{
    if (s->batch)
        return s->batch;
    BatchState *b = new (std::nothrow) BatchState();
    if (!b)
        return nullptr;
    s->batch = b;
    hipError_t e = hipMalloc((void **)&b->d_batch, sizeof(StepBatch));
    if (e == hipSuccess)
        e = hipMalloc((void **)&b->d_work, 64);
    for (int k = 0; e == hipSuccess && k < RING; ++k) {
        e = hipHostMalloc((void **)&b->h_batch[k], sizeof(StepBatch), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        set_error(s, (int)e, "batch state", e);
        batch_free(s);
        return nullptr;
    }
    return b;
}

/* the group forms: 1 = four lanes per world (one leg each), 2 = two lanes
 * per world (two legs each), 4 = four lanes per world with the root joints
 * of a block's worlds solved by one wavefront; 0 = one world per lane */
static int is_leg(const arc_sim *s) { return s->form == 2 ? 1 : s->form == 3 ? 2 : s->form == 4 ? 4 : 0; }
static int lanes_per_world(const arc_sim *s) { return s->form == 2 || s->form == 4 ? 4 : s->form == 3 ? 2 : 1; }
/* LDS layout actually launched: the group forms always hold the terrain tile */
static int tile_of(const arc_sim *s) { return is_leg(s) ? 16 : s->tile; }

/* the leg form's per-lane contact lists, allocated on first use */
hipError_t ensure_leg(arc_sim *s)
{
    if (s->dev.lcbuf)
        return hipSuccess;
    const Model &M = s->su.model;
    /* the forms record contacts per lane without the per-world running
     * total: they need every sphere window to fit the 128-contact cap */
    if (M.cslots > MAXC)
        return hipErrorNotSupported;
    const size_t slots = (size_t)std::max(leg_world_slots(M, 1), leg_world_slots(M, 2));
    /* and the shared-roots form's dummy lists and anchors: one lane's slots
     * and the 21 spheres' anchors for each of a block's 256 lanes */
    const size_t dummy = ((size_t)(M.ccap[0] + M.ccap[1] + M.ccap[5]) * 6 + 2 * NS * 3) * 256;
    hipError_t e = hipMalloc((void **)&s->dev.lcbuf, (slots * 6 * s->dev.stride + dummy) * sizeof(float));
    if (e == hipSuccess)
        s->dev.lcdummy = s->dev.lcbuf + slots * 6 * s->dev.stride;
    return e;
}

/* grid of a launch over `total` worlds (graph: the worst case N) */
static int grid_for(const arc_sim *s, int total)
{
    const int items = lanes_per_world(s) * total;
    if (s->persistent) {
        int r = step_resident_blocks(s->block, tile_of(s), is_leg(s));
        return r > 0 ? r : (items + s->block - 1) / s->block;
    }
    if (s->grid_blocks > 0)
        return s->grid_blocks;
    return (items + s->block - 1) / s->block;
}

/* the kernels of one batched control step, on the handle's stream */
static hipError_t enqueue(arc_sim *s, BatchState *b, const float *targets, int task, int grid)
{
    hipError_t e;
    if (s->persistent && (e = launch_fill_int(b->d_work, 0, 1, s->stream)) != hipSuccess)
        return e;
    if ((e = launch_step(s->dev, s->d_model, targets, b->d_batch, 0, s->block, tile_of(s), grid,
                         s->persistent ? b->d_work : nullptr, is_leg(s), s->stream)) != hipSuccess)
        return e;
    if (!task)
        return hipSuccess;
    if ((e = launch_fill_int(s->ts.count, 0, 1, s->stream)) != hipSuccess
        || (e = launch_falls(s->dev, s->d_model, s->task, s->ts, b->d_batch, s->stream)) != hipSuccess)
        return e;
    return launch_reset(s->dev, s->d_model, s->d_reset, s->ts.list, s->ts.seeds, nullptr, s->ts.diff, 0, s->ts.count,
                        s->reset_blocks, s->stream);
}

/* the graph of key k, captured into the least recently used slot if absent */
static hipError_t graph_for(arc_sim *s, BatchState *b, const LaunchKey &k, hipGraphExec_t *out)
{
    BatchState::Entry *e = nullptr;
    for (auto &c : b->cache)
        if (c.exec && c.key == k)
            e = &c;
    if (!e) {
        e = &b->cache[0];
        for (auto &c : b->cache)
            if (c.used < e->used)
                e = &c;
        if (e->exec)
            (void)hipGraphExecDestroy(e->exec);
        if (e->graph)
            (void)hipGraphDestroy(e->graph);
        e->exec = nullptr;
        e->graph = nullptr;
        hipError_t r = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (r != hipSuccess)
            return r;
        hipError_t r1 = enqueue(s, b, nullptr, k.task, k.grid);
        hipError_t r2 = hipStreamEndCapture(s->stream, &e->graph);
        if (r1 != hipSuccess)
            return r1;
        if (r2 != hipSuccess)
            return r2;
        if ((r = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0)) != hipSuccess)
            return r;
        e->key = k;
        b->captures++;
    }
    e->used = ++b->clock;
    *out = e->exec;
    return hipSuccess;
}

/* buffers the configured launch allocates on first use, made now (before a
 * graph capture, which must not allocate) */
hipError_t prepare_launch(arc_sim *s)
{
    hipError_t e;
    if (is_leg(s) && (e = ensure_leg(s)) != hipSuccess)
        return e;
    if (s->persistent && !batch_of(s))
        return hipErrorOutOfMemory;
    return hipSuccess;
}

/* every world of the sim in the configured register-resident form (lane, leg,
 * pair or shared) and distribution */
hipError_t launch_lane_all(arc_sim *s, const float *targets)
{
    hipError_t e;
    note_contacts(s);
    if (is_leg(s) && (e = ensure_leg(s)) != hipSuccess)
        return e;
    if (!s->persistent)
        return launch_step(s->dev, s->d_model, targets, nullptr, s->n, s->block, tile_of(s), s->grid_blocks, nullptr,
                           is_leg(s), s->stream);
    BatchState *b = batch_of(s);
    if (!b)
        return hipErrorOutOfMemory;
    if ((e = launch_fill_int(b->d_work, 0, 1, s->stream)) != hipSuccess)
        return e;
    return launch_step(s->dev, s->d_model, targets, nullptr, s->n, s->block, tile_of(s), grid_for(s, s->n), b->d_work,
                       is_leg(s), s->stream);
}

/* bytes of contact lists (and idle-lane lists and anchors of the shared-root
 * form) a step of every world in the configured register-resident form
 * writes; the layouts are those of dev_alloc (lane form) and ensure_leg */
static size_t contact_list_bytes(const arc_sim *s)
{
    const Model &M = s->su.model;
    const size_t S = s->dev.stride;
    if (!is_leg(s))
        return (size_t)M.cslots * 6 * S * sizeof(float);
    const int k = is_leg(s) == 2 ? 2 : 1;
    size_t floats = (size_t)leg_world_slots(M, k) * 6 * S;
    if (is_leg(s) == 4)
        floats += ((size_t)(M.ccap[0] + M.ccap[1] + M.ccap[5]) * 6 + 2 * NS * 3) * 256;
    return floats * sizeof(float);
}

extern "C" {

int arcx_step_device_noalloc(arc_sim *s, const float *targets_dev)
{
    if (!device_memory(s, targets_dev, (size_t)s->n * NJ * sizeof(float))) {
        set_error(s, -1, "arcx_step_device_noalloc: targets are not device memory of the sim's GPU", hipSuccess);
        return -1;
    }
    if (s->n > ARCX_NOALLOC_MAX_WORLDS) {
        set_error(s, -1, "arcx_step_device_noalloc: more worlds than the static scratch holds (8192)", hipSuccess);
        return -1;
    }
    if (s->form == 1) {
        /* the plain form keeps its contact lists in LDS: no scratch at all */
        HIPCHK_RET(s, launch_form(s, targets_dev), "arcx_step_device_noalloc", -1);
        return 0;
    }
    if (is_leg(s) && s->su.model.cslots > MAXC) {
        set_error(s, -1, "arcx_step_device_noalloc: the robot's contact windows exceed the group forms' cap",
                  hipSuccess);
        return -1;
    }
    char *base = (char *)noalloc_scratch(s->device);
    if (!base) {
        set_error(s, -1, "arcx_step_device_noalloc: the static scratch's device address is unknown", hipSuccess);
        return -1;
    }
    const size_t lists = (contact_list_bytes(s) + 255) / 256 * 256;
    if (lists + 256 > NOALLOC_SCRATCH_BYTES) {
        set_error(s, -1, "arcx_step_device_noalloc: the sim's contact lists exceed the static scratch", hipSuccess);
        return -1;
    }
    /* the sim's layout with its scratch pointers moved into the static array */
    DevWorlds d = s->dev;
    if (is_leg(s)) {
        d.lcbuf = (float *)base;
        d.lcdummy = d.lcbuf + (size_t)leg_world_slots(s->su.model, 1) * 6 * d.stride;
    } else {
        d.cbuf = (float *)base;
    }
    int *work = nullptr;
    int grid = s->grid_blocks;
    if (s->persistent) {
        work = (int *)(base + lists);
        grid = grid_for(s, s->n);
        HIPCHK_RET(s, launch_fill_int(work, 0, 1, s->stream), "arcx_step_device_noalloc counter", -1);
    }
    HIPCHK_RET(s, launch_step(d, s->d_model, targets_dev, nullptr, s->n, s->block, tile_of(s), grid, work, is_leg(s),
                              s->stream),
               "arcx_step_device_noalloc", -1);
    s->contact_layout = 0; /* the lists are in the shared static array, which later calls overwrite */
    return 0;
}

int arcx_set_layout(arc_sim *s, int tile)
{
    if (tile != 0 && tile != 16)
        return -1;
    s->tile = tile;
    return 0;
}

int arcx_set_persistent(arc_sim *s, int on)
{
    if (on && s->form == 4)
        return -1; /* the shared-roots form moves whole blocks, not wavefronts */
    s->persistent = on ? 1 : 0;
    return 0;
}

int arcx_get_launch(const arc_sim *s, int *out5)
{
    out5[0] = s->block;
    out5[1] = tile_of(s);
    out5[2] = s->persistent ? step_resident_blocks(s->block, tile_of(s), is_leg(s)) : s->grid_blocks;
    out5[3] = s->persistent;
    out5[4] = s->form;
    return 0;
}

/* arcx_step_batch and arcx_step_batch_rows: request r's targets start at
 * rows[r] (rows given), else at targets_dev + starts[r] x 12 */
static int step_batch(arc_sim *s, const int *starts, const int *counts, int n_req, const float *targets_dev,
                      const float *const *rows, int task, int graph, bool rows_checked)
{
    if (n_req < 0 || n_req > MAX_BATCH || (n_req && (!starts || !counts)) || (!targets_dev && !rows)) {
        set_error(s, -1, "arcx_step_batch: request list", hipSuccess);
        return -1;
    }
    for (int r = 0; rows && r < n_req; ++r)
        if (counts[r] > 0 && !rows[r]) {
            set_error(s, -1, "arcx_step_batch_rows: a request without targets", hipSuccess);
            return -1;
        }
    /* ranges inside [0, N) and disjoint: overlapping ranges would race */
    std::vector<std::pair<int, int>> iv;
    iv.reserve(n_req);
    size_t end = 0; /* targets_dev must hold the rows up to the last range's end */
    for (int r = 0; r < n_req; ++r) {
        if (counts[r] < 0 || starts[r] < 0 || starts[r] > s->n - counts[r]) {
            set_error(s, -1, "arcx_step_batch: range outside the sim", hipSuccess);
            return -1;
        }
        if (counts[r]) {
            iv.emplace_back(starts[r], counts[r]);
            end = std::max(end, (size_t)starts[r] + (size_t)counts[r]);
        }
    }
    if (targets_dev && !device_memory(s, targets_dev, end * NJ * sizeof(float))) {
        set_error(s, -1, "arcx_step_batch: targets are not device memory of the sim's GPU", hipSuccess);
        return -1;
    }
    /* every request's row block: counts[r] x 12 floats inside one allocation
     * of the sim's GPU (a host address would fault the kernels) */
    for (int r = 0; rows && !rows_checked && r < n_req; ++r)
        if (counts[r] > 0 && !device_memory(s, rows[r], (size_t)counts[r] * NJ * sizeof(float))) {
            set_error(s, -1, "arcx_step_batch_rows: a request's targets are not device memory of the sim's GPU",
                      hipSuccess);
            return -1;
        }
    std::sort(iv.begin(), iv.end());
    for (size_t k = 1; k < iv.size(); ++k)
        if (iv[k - 1].first + iv[k - 1].second > iv[k].first) {
            set_error(s, -1, "arcx_step_batch: overlapping ranges", hipSuccess);
            return -1;
        }
    if (s->form == 1) {
        set_error(s, -1, "arcx_step_batch: the batched launch runs the register-resident forms, not the plain form",
                  hipSuccess);
        return -1;
    }
    if (is_leg(s))
        HIPCHK_RET(s, ensure_leg(s), "arcx_step_batch: leg-form buffers", -1);
    BatchState *b = batch_of(s);
    if (!b)
        return -1;
    const int slot = b->slot;
    b->slot = (slot + 1) % RING;
    HIPCHK_RET(s, hipEventSynchronize(b->ev[slot]), "arcx_step_batch", -1);
    StepBatch *h = b->h_batch[slot];
    int total = 0, m = 0;
    for (int r = 0; r < n_req; ++r) {
        if (!counts[r])
            continue;
        h->r[m].start = starts[r];
        h->r[m].count = counts[r];
        h->r[m].prefix = total;
        h->r[m].pad = 0;
        h->r[m].targets = rows ? rows[r] : targets_dev + (size_t)starts[r] * NJ;
        total += counts[r];
        ++m;
    }
    h->n_ranges = m;
    h->total = total;
    if (m == 0) {
        /* an empty table still steps nothing: the kernels read total = 0 */
        h->r[0].start = h->r[0].count = h->r[0].prefix = h->r[0].pad = 0;
        h->r[0].targets = nullptr;
    }
    const size_t bytes = offsetof(StepBatch, r) + sizeof(Range) * (size_t)std::max(m, 1);
    /* the device table is rewritten only when the batch differs from the last
     * one uploaded (stream order keeps the kernels reading the right one) */
    if (bytes != b->last_bytes || memcmp(h, &b->last, bytes) != 0) {
        HIPCHK_RET(s, hipMemcpyAsync(b->d_batch, h, bytes, hipMemcpyHostToDevice, s->stream), "arcx_step_batch table",
                   -1);
        memcpy(&b->last, h, bytes);
        b->last_bytes = bytes;
    }
    HIPCHK_RET(s, hipEventRecord(b->ev[slot], s->stream), "arcx_step_batch", -1);
    note_contacts(s);
    if (!graph) {
        if (total == 0 && !task)
            return 0;
        HIPCHK_RET(s, enqueue(s, b, nullptr, task, grid_for(s, std::max(total, 1))), "arcx_step_batch launch", -1);
        return 0;
    }
    LaunchKey k{ task ? 1 : 0, s->block, tile_of(s), grid_for(s, std::max(s->n, 1)), s->persistent, is_leg(s) };
    hipGraphExec_t exec = nullptr;
    HIPCHK_RET(s, graph_for(s, b, k, &exec), "arcx_step_batch capture", -1);
    HIPCHK_RET(s, hipGraphLaunch(exec, s->stream), "arcx_step_batch graph", -1);
    return 0;
}

int arcx_step_batch(arc_sim *s, const int *starts, const int *counts, int n_req, const float *targets_dev, int task,
                    int graph)
{
    return step_batch(s, starts, counts, n_req, targets_dev, nullptr, task, graph, false);
}

int arcx_step_batch_rows(arc_sim *s, const int *starts, const int *counts, int n_req, const float *const *rows,
                         int task, int graph)
{
    return step_batch(s, starts, counts, n_req, nullptr, rows, task, graph, false);
}

} /* extern "C" */

int step_batch_rows_checked(arc_sim *s, const int *starts, const int *counts, int n_req, const float *const *rows,
                            int task, int graph)
{
    return step_batch(s, starts, counts, n_req, nullptr, rows, task, graph, true);
}

extern "C" {

int arcx_graph_captures(const arc_sim *s) { return s->batch ? s->batch->captures : 0; }

/* Times every launch configuration of the register-resident forms on this
 * handle's worlds and keeps the fastest.  The device state is copied aside
 * first and put back afterwards, so the worlds end where they started. */
int arcx_autotune(arc_sim *s, int steps, int *out5, double *ms_out)
{
    if (s->n <= 0 || steps <= 0) {
        set_error(s, -1, "arcx_autotune: nothing to time", hipSuccess);
        return -1;
    }
    if (s->form == 1) {
        set_error(s, -1, "arcx_autotune: the tuner configures the register-resident forms, not the plain form",
                  hipSuccess);
        return -1;
    }
    const size_t nt = (size_t)s->n * NJ;
    /* default pose plus 0.6 x uniform actions, a fresh draw for each step of a
     * timing run (set j from the action stream seeded j): the step's cost
     * depends on the action law, and a law that repeats a few sets makes the
     * legs oscillate and the contact load grow */
    const int sets = std::min(steps, 16) + 1;
    std::vector<float> ht(nt * (size_t)sets);
    {
        std::vector<unsigned> u(nt);
        for (int j = 0; j < sets; ++j) {
            arcx_rng_stream((unsigned)j, (long long)nt, u.data());
            for (size_t k = 0; k < nt; ++k) {
                double a = (double)(u[k] >> 8) * (1.0 / 8388608.0) - 1.0;
                ht[(size_t)j * nt + k] = (float)((double)s->su.info.def[k % NJ] + 0.6 * a);
            }
        }
    }
    float *dt = nullptr;
    void *snap = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -1;
    const int keep_block = s->block, keep_tile = s->tile, keep_grid = s->grid_blocks, keep_pers = s->persistent,
              keep_form = s->form;
    int best[5] = { keep_block, keep_tile, keep_grid, keep_pers, keep_form };
    double best_ms = 1e30;
    do {
        if (hipMalloc((void **)&dt, ht.size() * sizeof(float)) != hipSuccess || hipMalloc(&snap, s->dev.bytes) != hipSuccess
            || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            set_error(s, -1, "arcx_autotune: allocation", hipSuccess);
            break;
        }
        bool ok = hipMemcpyAsync(dt, ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice, s->stream) == hipSuccess
                  && hipMemcpyAsync(snap, s->dev.base, s->dev.bytes, hipMemcpyDeviceToDevice, s->stream) == hipSuccess
                  && hipStreamSynchronize(s->stream) == hipSuccess;
        const int blocks[3] = { 64, 128, 256 }, tiles[2] = { 16, 0 }, forms[4] = { 0, 2, 3, 4 };
        const bool leg_ok = ensure_leg(s) == hipSuccess; /* else lane form only */
        struct Cand {
            int form, block, tile, pers;
            double ms;
        };
        std::vector<Cand> cand;
        for (int fi = 0; fi < 4; ++fi)
            for (int bi = 0; bi < 3 && (forms[fi] == 0 || leg_ok); ++bi)
                for (int ti = 0; ti < (forms[fi] == 0 ? 2 : 1); ++ti)
                    for (int pi = 0; pi < 2; ++pi)
                        /* the shared-roots form: blocks of 128 or 256, no work counter */
                        if (!(forms[fi] == 4 && (blocks[bi] == 64 || pi)))
                            cand.push_back(Cand{ forms[fi], blocks[bi], tiles[ti], pi, 0.0 });
        /* every candidate timed twice, the candidates in order and then in
         * reverse order, so that a drift of the device's speed over the run
         * favours none of them */
        const int nc = (int)cand.size();
        for (int pass = 0; pass < 2 && ok; ++pass)
            for (int ci = 0; ci < nc && ok; ++ci) {
                Cand &c = cand[pass == 0 ? ci : nc - 1 - ci];
                s->form = c.form;
                s->block = c.block;
                s->tile = c.tile;
                s->grid_blocks = 0;
                s->persistent = c.pers;
                const int start = 0, count = s->n;
                /* from the saved state: one untimed step, then `steps` timed ones */
                float ms = 0.0f;
                ok = hipMemcpyAsync(s->dev.base, snap, s->dev.bytes, hipMemcpyDeviceToDevice, s->stream) == hipSuccess;
                ok = ok && arcx_step_batch(s, &start, &count, 1, dt, 0, 0) == 0;
                ok = ok && hipEventRecord(e0, s->stream) == hipSuccess;
                for (int k = 0; k < steps && ok; ++k)
                    ok = arcx_step_batch(s, &start, &count, 1, dt + (size_t)(1 + k % (sets - 1)) * nt, 0, 0) == 0;
                ok = ok && hipEventRecord(e1, s->stream) == hipSuccess;
                ok = ok && hipEventSynchronize(e1) == hipSuccess;
                ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
                c.ms += ms / (float)steps / 2.0f;
            }
        for (const Cand &c : cand) {
            /* the persistent distribution only when far faster: it pays where the
             * ranges are many and small, while this one full range has timed it a few
             * per cent faster than it then ran in steady serving (graph replays,
             * a counter reset every step) */
            if (ok && c.ms < best_ms * (c.pers ? 0.9 : 1.0)) {
                best_ms = c.ms;
                best[0] = c.block;
                best[1] = c.tile;
                best[2] = 0;
                best[3] = c.pers;
                best[4] = c.form;
            }
        }
        if (!ok) {
            set_error(s, -1, "arcx_autotune: timing run", hipSuccess);
            break;
        }
        rc = 0;
    } while (0);
    if (snap)
        (void)hipMemcpyAsync(s->dev.base, snap, s->dev.bytes, hipMemcpyDeviceToDevice, s->stream);
    (void)hipStreamSynchronize(s->stream);
    if (rc == 0) {
        s->block = best[0];
        s->tile = best[1];
        s->grid_blocks = best[2];
        s->persistent = best[3];
        s->form = best[4];
        arcx_get_launch(s, out5);
        if (ms_out)
            *ms_out = best_ms;
    } else {
        s->block = keep_block;
        s->tile = keep_tile;
        s->grid_blocks = keep_grid;
        s->persistent = keep_pers;
        s->form = keep_form;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (dt) (void)hipFree(dt);
    if (snap) (void)hipFree(snap);
    return rc;
}

} /* extern "C" */
