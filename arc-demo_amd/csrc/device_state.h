/*
 * device_state.h -- HBM layout of a simulation's worlds and the launchers of
 * the kernels that work on it.
 *
 * Dynamic state is structure-of-arrays with the world index fastest
 * (field f of world i at f * stride + i), so a wavefront of 64 lanes handling
 * 64 consecutive worlds reads and writes 256-byte runs.  Terrain is stored
 * world-major (one world's grid contiguous), since each world only touches
 * the few columns under its robot.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "lane_engine.h"
#include "reset_engine.h"

namespace xpbd {

/* LDS sizing of the reset kernel: grid up to 64 x 64 columns; value-noise
 * lattices of (64/8+2)^2 and (64/3+2)^2 points */
constexpr int GRID_LIMIT_CELLS = 64 * 64;
constexpr int LAT_LIMIT = (64 / 8 + 2) * (64 / 8 + 2) + (64 / 3 + 2) * (64 / 3 + 2);

struct DevWorlds {
    int n;            /* worlds */
    size_t stride;    /* SoA stride (n rounded up to a multiple of 64) */
    int cells;        /* nx * ny terrain columns per world */
    int tpitch;       /* nx + 2 * TPAD: row pitch of the stored terrain */
    int tcells;       /* tpitch^2 bytes of stored terrain per world */
    void *base;       /* the one allocation holding every array below */
    size_t bytes;     /* its size */
    float *x;         /* [NB*3][stride] body position */
    float *q;         /* [NB*4][stride] body orientation w,x,y,z */
    float *v;         /* [NB*3][stride] linear velocity */
    float *w;         /* [NB*3][stride] angular velocity */
    float *target;    /* [NJ][stride] drive targets (clamped) */
    float *jout;      /* [3*NJ][stride] joint angle, rate, torque */
    uint8_t *touch;   /* [NS][stride] substeps in contact during the last step */
    uint8_t *ncontact; /* [NCB][stride] contacts of each contact body recorded in the last substep of the
                          last step (0 after a reset), the records themselves in cbuf or lcbuf */
    uint32_t *anchored; /* [stride] friction-anchor bit per sphere */
    float *anc;       /* [NS*3][stride] friction anchor */
    float *ancn;      /* [NS*3][stride] anchor normal */
    int *unstable;    /* [stride] */
    float *time;      /* [stride] */
    int8_t *tq;       /* [n][tcells] column steps (TerrainRef), world-major */
    float *tbase;     /* [stride] terrain base (TerrainRef) */
    float *outside;   /* [stride] height outside the grid */
    float *difficulty; /* [stride] terrain difficulty of the last reset */
    float *cbuf;      /* [cslots*6][stride] per-substep contact list */
    float *lcbuf;     /* leg form's contact lists, [slot*6+field][4*stride] (allocated on first use) */
    float *lcdummy;   /* shared-roots form: contact lists [slot*6+field][256] and anchors [2][NS*3][256]
                         of lanes that step no world (block tails, frozen worlds) */
    int *err;         /* [stride] contact-list overflow bits (sticky) */
};

/* Robot data the reset needs beyond the Model. */
struct ResetK {
    float def[NJ], lo[NJ], hi[NJ];
    float stand, footprint;
    int grid;         /* generated grid size n (n x n columns) */
    float cell;
    int chunk;        /* draws per lane of a generated reset (RESET_LANES lanes) */
    const uint32_t *jtab; /* per-lane jump tables (reset_engine.h); device memory in the device copy */
};

/* Template state broadcast at creation (the robot as built, flat ground). */
struct TemplateK {
    V3 x[NB];
    Q4 q[NB];
    float target[NJ], angle[NJ];
};

/* Task layer (arc/task.py:180-203, arc/env.py:99-117): fall rules and the
 * auto-reset of fallen worlds, kept on the device. */
struct TaskK {
    double half_extent;   /* 0.5 * grid * cell, from the caller's double cell */
    double void_z;        /* torso below this height: fell into the void */
    double flip_up_z;     /* torso up-axis z below this: flipped */
    int beached_steps;    /* more consecutive steps of torso contact: beached */
    int n;                /* worlds (seed schedule: seed = i + n * episode) */
};
struct TaskState {
    int *torso_steps;     /* [n] consecutive steps with torso contact */
    uint32_t *episode;    /* [n] resets so far */
    int *cause;           /* [n] cause of the last step's termination (0 none) */
    int *list;            /* [n] worlds to reset this step */
    uint32_t *seeds;      /* [n] their seeds */
    float *diff;          /* [n] their difficulties */
    int *count;           /* [1] entries in list */
    unsigned long long *stats; /* [5] terminations per cause (index 0: steps) */
};

/* Slices of worlds a batched launch steps: worlds [start, start+count)
 * for each range; prefix = sum of the counts before it; the range's targets
 * are count x 12 floats from `targets` (its request's own device buffer, or
 * the rows of a shared buffer). */
struct Range {
    int start, count, prefix, pad;
    const float *targets;
};
/* requests per batched launch */
constexpr int MAX_BATCH = 256;
/* the batch a launch reads from device memory when it starts */
struct StepBatch {
    int n_ranges, total, pad[2];
    Range r[MAX_BATCH];
};
/* contact-list slots per world of the group form with K legs per lane (its
 * 4 / K lanes each hold the torso's, and their legs' femur and tibia lists) */
inline __host__ __device__ int leg_world_slots(const Model &M, int K)
{
    return (LEGS / K) * (M.ccap[0] + K * (M.ccap[1] + M.ccap[5]));
}

/* world index of work item t of a batch: binary search over the prefixes */
__device__ __forceinline__ int batch_world(const StepBatch *__restrict__ b, int n_ranges, int t)
{
    int lo = 0, hi = n_ranges - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (b->r[mid].prefix <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return b->r[lo].start + (t - b->r[lo].prefix);
}
/* the same, with the first of the world's 12 targets */
__device__ __forceinline__ int batch_world(const StepBatch *__restrict__ b, int n_ranges, int t,
                                           const float *__restrict__ &row)
{
    int lo = 0, hi = n_ranges - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (b->r[mid].prefix <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int off = t - b->r[lo].prefix;
    row = b->r[lo].targets + (size_t)off * NJ;
    return b->r[lo].start + off;
}

/* Kernel launchers (kernels.hip).  All are asynchronous on `stream`. */
size_t dev_bytes_per_world(int grid, int cslots);
hipError_t dev_alloc(DevWorlds &d, int n, int grid, int cslots);
/* world i's terrain view */
__host__ __device__ inline TerrainRef terrain_ref(const DevWorlds &d, size_t i)
{
    TerrainRef r;
    r.q = d.tq + i * (size_t)d.tcells + TPAD + TPAD * d.tpitch;
    r.pitch = d.tpitch;
    r.base = 0.0f;
    r.outside = 0.0f;
    return r;
}
void dev_free(DevWorlds &d);
hipError_t launch_init_template(const DevWorlds &d, const TemplateK *tk_dev, float outside, hipStream_t s);
/* targets: n*NJ world-major, double (host ABI) or float (device inputs) */
/* step kernel.  batch (device) or worlds 0..total-1; tile 0 or 16 (LDS
 * terrain tile side); grid_blocks 0 = one work item per thread; work (device
 * counter, zeroed beforehand) selects the persistent distribution; leg = 1:
 * four lanes per world, 2: two lanes per world (leg_engine.h; need d.lcbuf
 * and tile 16), 0: one. */
hipError_t launch_step(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch, int total,
                       int block, int tile, int grid_blocks, int *work, int leg, hipStream_t s);
hipError_t launch_step_b64(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch,
                           int total, int tile, int grid_blocks, int *work, int leg, hipStream_t s);
hipError_t launch_step_b128(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch,
                            int total, int tile, int grid_blocks, int *work, int leg, hipStream_t s);
hipError_t launch_step_b256(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch,
                            int total, int tile, int grid_blocks, int *work, int leg, hipStream_t s);
/* workgroups of the step kernel (block, tile) resident on the device at once */
int step_resident_blocks(int block, int tile, int leg);
int step_resident_blocks_b64(int tile, int leg);
int step_resident_blocks_b128(int tile, int leg);
int step_resident_blocks_b256(int tile, int leg);
/* Resets ids[k] with seeds[k] and difficulty (double or float array), for k <
 * *count_dev when count_dev is given (entries decided on the device), else
 * k < n_ids; at most `max_blocks` blocks loop over the entries. */
hipError_t launch_reset(const DevWorlds &d, const Model *m_dev, const ResetK *r_dev, const int *ids,
                        const uint32_t *seeds, const double *difficulty_d, const float *difficulty_f, int n_ids,
                        const int *count_dev, int max_blocks, hipStream_t s);
/* p[0..n) = v with a kernel (safe inside captured graphs, unlike memset nodes) */
hipError_t launch_fill_int(int *p, int v, int n, hipStream_t s);
/* Fall rules after a step; fallen worlds are appended to the task list
 * (ts.count must be zero before). */
/* fall rules over the worlds of batch (device), or all worlds when null */
hipError_t launch_falls(const DevWorlds &d, const Model *m_dev, const TaskK &tk, const TaskState &ts,
                        const StepBatch *batch, hipStream_t s);
hipError_t launch_observe(const DevWorlds &d, const Model *m_dev, double *root, double *rot, double *joints,
                          double *contacts, int *flags, hipStream_t s);
hipError_t launch_heights(const DevWorlds &d, const Model *m_dev, const double *offsets, int k, double *out,
                          hipStream_t s);
hipError_t launch_terrain_at(const DevWorlds &d, const Model *m_dev, const int *ids, const double *xy, int m,
                             double *out, hipStream_t s);
/* full state of world i as floats (layout of the oracle's state dump) */
hipError_t launch_dump(const DevWorlds &d, int i, float *out, hipStream_t s);
/* the last substep's contacts of world i (k_dump_contacts): out[0] count,
 * then 7 floats per contact (body, local point, normal), at most 128 */
hipError_t launch_dump_contacts(const DevWorlds &d, const Model *m_dev, int i, int layout, float *out, hipStream_t s);

} /* namespace xpbd */
